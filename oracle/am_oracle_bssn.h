/*
 * Shared declarations of the oracle's BSSN part (test infrastructure only).
 *
 * Variable order of the 24 evolved fields:
 *   0 chi | 1-6 gt_xx,xy,xz,yy,yz,zz | 7 K | 8-13 At_xx..zz | 14-16 Gt^x,y,z
 *   17 alpha | 18-20 beta^x,y,z | 21-23 B^x,y,z
 * The 11 fields with a Hessian are chi, gt_ij, alpha, beta^i (state indices
 * 0-6, 17-20), in that order.
 */
#ifndef AM_ORACLE_BSSN_H
#define AM_ORACLE_BSSN_H

#include <stdint.h>

#define BSSN_NV 24
#define BSSN_NH 11

typedef struct {
    double eta;          /* Gamma-driver damping */
    double shift_coeff;  /* d_t beta^i = shift_coeff B^i (+ advection) */
    double chi_floor;    /* chi used in divisions is max(chi, chi_floor) */
    int32_t advect_shift;/* 1: beta^k d_k on beta and B, and d_0 Gt in the B equation */
    int32_t slicing;     /* 0: 1+log (d_t alpha = -2 alpha K), 1: harmonic (-alpha^2 K) */
} orc_bssn_par;

#endif
