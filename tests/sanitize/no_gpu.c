/*
 * no_gpu.c — the GPU entry points of include/mcgb.h that the C driver calls, each failing with
 * MCGB_ERR_HIP as libmcgb does on a machine without a GPU, so that the driver and the host
 * part (csrc/mcgb_host.c) can be built with AddressSanitizer and UndefinedBehaviorSanitizer on
 * their own (tests/test_sanitizers.py) and run up to the point where they need the device.
 */
#include "mcgb.h"
#include "mcgb_internal.h"

#define NO_GPU                                                      \
    do {                                                            \
        mcgb_set_error("no HIP device available (host-only build)"); \
        return MCGB_ERR_HIP;                                        \
    } while (0)

int mcgb_create(const mcgb_params *p, int n, mcgb_ctx **out)
{
    (void)p; (void)n;
    if (out) *out = 0;
    NO_GPU;
}
void mcgb_destroy(mcgb_ctx *c) { (void)c; }
int mcgb_set_state(mcgb_ctx *c, int r, int n, const double *x, const double *y, const double *z,
                   const double *ux, const double *uy, const double *uz, const double box[3])
{
    (void)c; (void)r; (void)n; (void)x; (void)y; (void)z; (void)ux; (void)uy; (void)uz; (void)box;
    NO_GPU;
}
int mcgb_get_state(mcgb_ctx *c, int r, int n, double *x, double *y, double *z, double *ux,
                   double *uy, double *uz, double box[3])
{
    (void)c; (void)r; (void)n; (void)x; (void)y; (void)z; (void)ux; (void)uy; (void)uz; (void)box;
    NO_GPU;
}
int mcgb_get_state_fixed(mcgb_ctx *c, int r, int n, uint32_t *qx, uint32_t *qy, uint32_t *qz,
                         double *ux, double *uy, double *uz, double box[3])
{
    (void)c; (void)r; (void)n; (void)qx; (void)qy; (void)qz; (void)ux; (void)uy; (void)uz; (void)box;
    NO_GPU;
}
int mcgb_get_stats(mcgb_ctx *c, int r, mcgb_stats *s) { (void)c; (void)r; (void)s; NO_GPU; }
int mcgb_get_resume(mcgb_ctx *c, int r, mcgb_resume *o) { (void)c; (void)r; (void)o; NO_GPU; }
int mcgb_resume_state(mcgb_ctx *c, int r, int n, const uint32_t *qx, const uint32_t *qy,
                      const uint32_t *qz, const double *ux, const double *uy, const double *uz,
                      const double box[3], const mcgb_resume *rec)
{
    (void)c; (void)r; (void)n; (void)qx; (void)qy; (void)qz; (void)ux; (void)uy; (void)uz;
    (void)box; (void)rec;
    NO_GPU;
}
int mcgb_order_parameter(mcgb_ctx *c, int r, double *p2, double d[3])
{
    (void)c; (void)r; (void)p2; (void)d;
    NO_GPU;
}
int mcgb_sweep_async(mcgb_ctx *c, int n) { (void)c; (void)n; NO_GPU; }
int mcgb_sync(mcgb_ctx *c) { (void)c; NO_GPU; }
int mcgb_snapshot_begin(mcgb_ctx *c, int r) { (void)c; (void)r; NO_GPU; }
int mcgb_snapshot_wait(mcgb_ctx *c, int r, int n, double *x, double *y, double *z, double *ux,
                       double *uy, double *uz, double box[3])
{
    (void)c; (void)r; (void)n; (void)x; (void)y; (void)z; (void)ux; (void)uy; (void)uz; (void)box;
    NO_GPU;
}
