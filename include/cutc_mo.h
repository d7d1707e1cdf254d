/*
 * AO -> MO transform of int2_grad1_u12: MI355X (gfx950) library, the step
 * between cutc_int's output and cutc_no's input.  Host code is C; the
 * contractions run on the library's FP64 MFMA GEMM.
 *
 * CuTC has no such function: its callers transform cutc_int's AO-basis
 * int2_grad1_u12_ao(n_ao, n_ao, n_grid1, 3) themselves and hand cutc_no the
 * MO-basis int2_grad1_u12(n_grid1, 3, n_mo, n_mo) (layout read by
 * src/no_tmpJ_cs_kernel.cu:18).  Definition (Fortran order, 1-based):
 *
 *   int2_grad1_u12(g, x, i, j) = sum_{k,l} mo_l_coef(k, i)
 *                                 * int2_grad1_u12_ao(k, l, g, x)
 *                                 * mo_r_coef(l, j)
 *
 * i.e. C_L^T A C_R for every grid point g and component x, C_L and C_R the
 * left and right MO coefficients (n_ao, n_mo) of a bi-orthogonal basis (pass
 * the same matrix twice for ordinary orbitals).  int2_grad1_u12_ao from
 * cutc_int is symmetric in (k, l), so exchanging the two coefficient
 * arguments transposes (i, j).
 */
#ifndef CUTC_MO_H
#define CUTC_MO_H

#include <stddef.h>
#include <stdint.h>

#include "cutc_int.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Host arrays in and out (the library allocates, copies and frees device
 * memory).  Like cutc_int and cutc_no it returns holding no device memory:
 * the cached workspace is released too (its pass buffer alone is
 * 8 n_ao * 3 n_grid1 * n_mo bytes when it fits), including what an earlier
 * cutc_int_ex call kept.  Returns 0, CUTC_ERR_ARG (-2) for negative sizes or
 * missing arrays, CUTC_ERR_NOMEM (-3), CUTC_ERR_HIP (-4). */
int cutc_ao_to_mo(int n_grid1, int n_ao, int n_mo, double *h_mo_l_coef, double *h_mo_r_coef,
                  double *h_int2_grad1_u12_ao, double *h_int2_grad1_u12);

typedef struct cutc_mo_stats {
    double ms_total;  /* stream-ordered hipEvents around the whole transform */
    double ms_left;   /* the contractions with mo_l_coef (over k)         */
    double ms_right;  /* the contractions with mo_r_coef (over l)         */
    double flops;     /* 6 n_grid1 n_ao n_mo (n_ao + n_mo)                 */
    int64_t mo_chunk; /* MO indices i per pass                              */
    int64_t n_chunks;
    size_t workspace_bytes;
} cutc_mo_stats;

/* The transform on device pointers and a caller stream (NULL: the null
 * stream).  mem_limit: device bytes the workspace may use (0: free memory);
 * mo_chunk: MO indices i per pass (0: as many as fit; each pass reads all of
 * int2_grad1_u12_ao).  The workspace (8 n_ao * 3 n_grid1 bytes per MO index
 * of a pass) is kept between calls (cutc_release_workspace frees it). */
int cutc_ao_to_mo_device(int n_grid1, int n_ao, int n_mo, const double *mo_l_coef, const double *mo_r_coef,
                         const double *int2_grad1_u12_ao, double *int2_grad1_u12, size_t mem_limit, int64_t mo_chunk,
                         void *stream, cutc_mo_stats *stats);

#ifdef __cplusplus
}
#endif
#endif
