/* Internal launch interface between the C host layer (cutc_int.c) and the
 * gfx950 kernels (*.hip).  Plain C: pointers, 64-bit sizes, a stream. */
#ifndef CUTC_LAUNCH_H
#define CUTC_LAUNCH_H

#include <hip/hip_runtime_api.h>
#include <stdint.h>

#include "cutc_int.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Jastrow monomial program (jastrow_prog.c, host) and kernels (jastrow.hip).
 * A monomial coef f1^i (x grad f1 unless is_r) is added to table slot `slot`
 * of its r1 row (slot .. slot+2 for grad-f1 terms, slot+3 for is_r terms). */
typedef struct cutc_jmono {
    int A, slot, i, is_r;
    double coef;
} cutc_jmono;

typedef struct cutc_jprog {
    cutc_jmono *mono; /* sorted by nucleus */
    int *cls;         /* (j, k) per class */
    int n_mono, n_cls, kmax;
    int64_t ts;       /* doubles per r1 row of the table */
} cutc_jprog;

int cutc_jast_program(int n_nuc, int size_bh, const double *c_bh, const int *m_bh, const int *n_bh, const int *o_bh,
                      cutc_jprog *prog);
void cutc_jast_program_free(cutc_jprog *prog);

/* Per-r1 table rows for r1 points [ii0, ii0 + n1) into tab (n1 x ts). */
int cutc_launch_jast_prep(const double *r1, int64_t ii0, int64_t n1, const double *rn, int n_nuc,
                          const cutc_jmono *mono, int n_mono, double *tab, int64_t ts, hipStream_t st);
/* G for r1 rows [ii0, ii0 + n1) x r2 rows [jj0, jj0 + n2), written as
 * G[m * plane + i1 * ldg + i2]; columns [n2, n2_fill) get 0.  tab from
 * cutc_launch_jast_prep for the same r1 rows; cls: device copy of the
 * program's classes.  f2tab (n_nuc x n2_fill) is scratch used when
 * n_nuc > 32, NULL otherwise. */
int cutc_launch_jast_pairs(const double *r1, const double *r2, const double *rn, int n_nuc, const double *tab,
                           int64_t ts, const int *cls, int n_cls, int kmax, int64_t ii0, int64_t n1, int64_t jj0,
                           int64_t n2, int64_t n2_fill, double *G, int64_t ldg, int64_t plane, double *f2tab,
                           hipStream_t st);

/* Integrands (integrands.hip).  pairs: (k, l) per packed index, k <= l. */
int cutc_launch_pair_table(int n_ao, int *pairs, hipStream_t st);
int cutc_launch_pair_weights(const double *wr2, const double *aos2, int64_t n2_ld, int64_t jj0, int64_t n2,
                             int64_t n2_fill, int n_ao, double *Apk, int64_t lda, hipStream_t st);
int cutc_launch_r1_integrands(const double *wr1, const double *aos1, int64_t n1_ld, int64_t ii0, int64_t n1,
                              int n_ao, const int *pairs, double *Hp, double *NHp, int64_t ldq, int64_t nh_plane,
                              hipStream_t st);
int cutc_launch_long_range_ref(int64_t jj0, int64_t n2_eff, int64_t n2_tot, int64_t n2, int n_ao, const double *wr2,
                               const double *aos2, double *out, hipStream_t st);
int cutc_launch_herm_ref(int64_t n1, int n_ao, const double *wr1, const double *aos1, double *out, hipStream_t st);
int cutc_launch_nonherm_ref(int64_t n1, int n_ao, const double *wr1, const double *aos1, double *out,
                            hipStream_t st);

/* dst[0..n) = value. */
int cutc_launch_fill(double *dst, int64_t n, double value, hipStream_t st);

/* M <- M + M^T for an n x n column-major matrix (symm.hip). */
int cutc_launch_trans_inplace(double *M, int64_t n, hipStream_t st);

/* FP64 MFMA GEMMs (gemm_f64.hip). */
/* Step 1: C[p][j] = sum_k A[p][k] G[j][k], both K-contiguous (ld lda, ldg;
 * K % 16 == 0).  j = m * n1c + i1; scattered into X (4, n1, n_ao, n_ao) at
 * r1 offset i0 (both (k,l) and (l,k)).  accumulate != 0 adds to X. */
int cutc_gemm1_f64(const double *Apk, int64_t lda, const double *G, int64_t ldg, int64_t P, int64_t n1c, int64_t K,
                   const int *pairs, int n_ao, double *X, int64_t n1, int64_t i0, int accumulate, hipStream_t st);
/* Step 2: C[p][q] = alpha sum_k X[k][idx(p)] B[k][q] (rows k of X at stride
 * xld, of B at stride ldb), scattered into the n_ao^2 x n_ao^2 matrix M.
 * X has K valid rows; B must hold roundup(K, 16) rows, zero past row K.
 * mode 0: store C at (kl,ab) (lk,ab) (kl,ba) (lk,ba); 1: add there;
 * 2: add at (*,ab) and subtract at (*,ba) for a < b. */
int cutc_gemm2_f64(const double *X, int64_t xld, const double *B, int64_t ldb, int64_t P, int64_t K, double alpha,
                   const int *pairs, int n_ao, double *M, int mode, hipStream_t st);

/* Split-precision (Ozaki) path (split.hip). */
int cutc_split_supported(int precision, int slices);

#ifdef __cplusplus
}
#endif
#endif
