/* ref_shim.c — appended to the reference's physics functions by build_ref.sh (same translation
 * unit) to export two static inline helpers of MCnptColloid.c for the tests. Test-only. */
double ref_compute_distance_sq(double x1, double y1, double z1, double x2, double y2, double z2,
                               double *d)
{
    return compute_distance_sq(x1, y1, z1, x2, y2, z2, &d[0], &d[1], &d[2]);
}

double ref_g_omega(double eiej, double amasb2, double amenb2, double omega)
{
    return g_omega(eiej, amasb2, amenb2, omega);
}
