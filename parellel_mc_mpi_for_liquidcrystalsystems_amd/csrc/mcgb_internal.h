/*
 * mcgb_internal.h — declarations shared by the C host part (mcgb_host.c) and the HIP part
 * (mcgb_api.hip) of libmcgb.  Not installed; callers use include/mcgb.h.
 */
#ifndef MCGB_INTERNAL_H
#define MCGB_INTERNAL_H

#include "mcgb.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Accepted moves of one sweep whose sum of |finite dE| exceeds this trigger an FP64
 * re-summation of the running energy at the end of the sweep (near-cap pairs carry finite
 * energies up to ~1e36, far beyond what an FP64 running sum can track). */
#define MCGB_RESUM_THRESHOLD 1.0e10

void     mcgb_set_error(const char *fmt, ...);
void     mcgb_philox_host(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
/* Largest even n with L/n >= rc*(1+1e-6); 0 when that n is below 4. */
int      mcgb_cell_dim(double L, double rc);
/* x -> 32-bit fixed-point fraction of the box L (x wrapped into [0, L)). */
uint32_t mcgb_to_fixed(double x, double L);

#ifdef __cplusplus
}
#endif
#endif
