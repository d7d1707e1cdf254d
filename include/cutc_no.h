/*
 * Normal ordering of the three-body transcorrelated term: MI355X (gfx950)
 * library, drop-in for CuTC's cutc_no.  Host code is C; the kernels are HIP.
 *
 * Shapes (Fortran order, as in CuTC):
 *   wr1(n_grid1), mos_l_in_r(n_grid1, n_mo), mos_r_in_r(n_grid1, n_mo),
 *   int2_grad1_u12(n_grid1, 3, n_mo, n_mo)      (the MO-basis integrals)
 *   no_2e(n_mo, n_mo, n_mo, n_mo), no_1e(n_mo, n_mo), no_0e(1)
 * Orbitals 1..ne_b are doubly occupied, ne_b+1..ne_a singly (ne_a == ne_b:
 * closed shell).
 */
#ifndef CUTC_NO_H
#define CUTC_NO_H

#include <stddef.h>
#include <stdint.h>

#include "cutc_int.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Replaces src/cutc_no.c:14-16 (exported there without a declaration).
 * Host arrays in and out.  Returns 0, or CUTC_ERR_ARG (-2) for negative
 * sizes or occupations beyond n_mo, CUTC_ERR_NOMEM (-3) when even a 16-point
 * grid chunk does not fit the device, CUTC_ERR_HIP (-4). */
int cutc_no(int n_grid1, int n_mo, int ne_a, int ne_b, double *h_wr1, double *h_mos_l_in_r, double *h_mos_r_in_r,
            double *h_int2_grad1_u12, double *h_no_2e, double *h_no_1e, double *h_no_0e);

/* ---------------------------------------------------------------------
 * Additions (no counterpart in the reference)
 * ------------------------------------------------------------------- */

typedef struct cutc_no_stats {
    double ms_total;  /* stream-ordered hipEvents around the whole build */
    double ms_point;  /* per-point and per-orbital intermediates             */
    double ms_pair;   /* the pair operands of the 2e contractions          */
    double ms_gemm2e; /* the 2e contractions (FP64 MFMA)                     */
    double ms_1e;     /* the 1e products (GEMV, GEMM, split-K sums)         */
    double ms_symm;   /* N + N^T with the (q, s) index swap, no_0e sum      */
    double flops_gemm2e; /* 2 n_mo^4 (3 + 1) n_grid1: the 2e contraction work */
    int64_t chunk;    /* grid points per chunk                             */
    int64_t n_chunks;
    size_t workspace_bytes;
} cutc_no_stats;

/* cutc_no on device pointers and a caller stream (NULL: the null stream):
 * no allocation of the caller's arrays, no host copies.  mem_limit: device
 * bytes the workspace may use (0: free memory); grid_chunk: grid points per
 * chunk (0: as many as fit).  The device workspace is kept between calls
 * (cutc_release_workspace frees it). */
int cutc_no_device(int n_grid1, int n_mo, int ne_a, int ne_b, const double *wr1, const double *mos_l_in_r,
                   const double *mos_r_in_r, const double *int2_grad1_u12, double *no_2e, double *no_1e,
                   double *no_0e, size_t mem_limit, int64_t grid_chunk, void *stream, cutc_no_stats *stats);

#ifdef __cplusplus
}
#endif
#endif
