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
/* The same into process-wide static host arrays (no allocation) for tables
 * of at most CUTC_JPROG_STATIC_TERMS terms (n_nuc x size_bh), CUTC_ERR_ARG
 * beyond; the program stays valid until the next call and is not freed.
 * Callers hold the library lock. */
#define CUTC_JPROG_STATIC_TERMS 4096
int cutc_jast_program_static(int n_nuc, int size_bh, const double *c_bh, const int *m_bh, const int *n_bh,
                             const int *o_bh, cutc_jprog *prog);

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
/* partial sums per G row written by cutc_launch_jast_pairs_norms: one per
 * block of 512 of the n2_fill points */
int64_t cutc_jast_norm_blocks(int64_t n2_fill);
/* cutc_launch_jast_pairs (r2 points from 0) that also writes, for every G
 * row m n1 + i1, its sums of squares over each block of 512 points into
 * part[b 4 n1 + m n1 + i1] (b < cutc_jast_norm_blocks(n2_fill)), for
 * cutc_rowexp_finish; CUTC_ERR_UNSUPPORTED (nothing launched) for the
 * shapes of the unstaged kernel (more than 32 nuclei, large programs) */
int cutc_launch_jast_pairs_norms(const double *r1, const double *r2, const double *rn, int n_nuc, const double *tab,
                                 int64_t ts, const int *cls, int n_cls, int kmax, int64_t ii0, int64_t n1, int64_t n2,
                                 int64_t n2_fill, double *G, int64_t ldg, int64_t plane, double *part,
                                 hipStream_t st);

/* Integrands (integrands.hip).  pairs: (k, l) per packed index, k <= l. */
int cutc_launch_pair_table(int n_ao, int *pairs, hipStream_t st);
int cutc_launch_pair_weights(const double *wr2, const double *aos2, int64_t n2_ld, int64_t jj0, int64_t n2,
                             int64_t n2_fill, int n_ao, double *Apk, int64_t lda, hipStream_t st);
/* perm (optional): row r of Hp / NHp is point ii0 + perm[r] instead of ii0 + r;
 * NHp NULL: Hp only */
int cutc_launch_r1_integrands(const double *wr1, const double *aos1, int64_t n1_ld, int64_t ii0, int64_t n1,
                              int n_ao, const int *pairs, double *Hp, double *NHp, int64_t ldq, int64_t nh_plane,
                              const uint32_t *perm, hipStream_t st);
int cutc_launch_long_range_ref(int64_t jj0, int64_t n2_eff, int64_t n2_tot, int64_t n2, int n_ao, const double *wr2,
                               const double *aos2, double *out, hipStream_t st);
int cutc_launch_herm_ref(int64_t n1, int n_ao, const double *wr1, const double *aos1, double *out, hipStream_t st);
int cutc_launch_nonherm_ref(int64_t n1, int n_ao, const double *wr1, const double *aos1, double *out,
                            hipStream_t st);

/* Contracted Cartesian Gaussian AOs and gradients (ao.hip): out[g + n (i +
 * n_ao c)], c = 0 value, 1..3 d/dx, d/dy, d/dz; nucl 1-based, power (n_ao,
 * 3), expo / coef (pmax, n_ao); an AO whose entries are out of range gets
 * NaN; n_ao <= 65535. */
int cutc_launch_aos(int64_t n, const double *r, const double *rn, int n_nuc, int n_ao, const int *nucl,
                    const int *power, const int *nprim, int pmax, const double *expo, const double *coef, double *out,
                    hipStream_t st);

/* Device scratch of the allocation-free entry point (scratch.hip): the
 * device address and size of this GPU's file-scope array, looked up once
 * per device (with that device current) and kept.  Callers hold the
 * library lock. */
#define CUTC_SCRATCH_BYTES CUTC_NOALLOC_SCRATCH_BYTES
int cutc_scratch(void **ptr, size_t *bytes);

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
/* Split-K step 2: unscaled partial sums of cutc_gemm2_split_count(K, nsplit)
 * K slices into work[y][q][p] (P x P each), reduced by cutc_scatter_m. */
int cutc_gemm2_f64_split(const double *X, int64_t xld, const double *B, int64_t ldb, int64_t P, int64_t K,
                         const int *pairs, int n_ao, int nsplit, double *work, hipStream_t st);
int cutc_gemm2_split_count(int64_t K, int nsplit);
int cutc_gemm2_f64(const double *X, int64_t xld, const double *B, int64_t ldb, int64_t P, int64_t K, double alpha,
                   const int *pairs, int n_ao, double *M, int mode, hipStream_t st);

/* Modular (CRT) split-precision contractions: crt_host.c (host constants) and
 * crt.hip (kernels).  Residue planes are [n][packed rows x kpad] in the
 * fragment-packed layouts below, zero past the valid K.  Each operand row x
 * is scaled by 2^(bits - e) with ||x||_2 < 2^e (its row exponent), so that
 * ||q||_2 < 2^bits + sqrt(K)/2 for the rounded row q and, by Cauchy-Schwarz,
 * every integer dot product stays below M/2 in magnitude.
 *
 * Kinds: CUTC_CRT_I8, moduli <= 256, int8 residues in [-128, 127] and
 * v_mfma_i32_32x32x32_i8 (exact int32 sums); CUTC_CRT_BF16 / _FP16, moduli
 * <= 255, residues |r| <= 127 stored as bf16 / fp16 values (exact in both)
 * and v_mfma_f32_32x32x16_{bf16,f16}, whose FP32 sums stay exact because
 * every accumulator is reduced modulo m at least once per 1024 K (at most
 * 1024 * 127^2 + 1.5 * 255 < 2^24 in between).  GEMM results are int8
 * residues for every kind. */
#define CUTC_CRT_MAX 16
/* Moduli (modular kinds) or digit slices (slice kinds) a precision takes
 * when cutc_options.slices is 0.  int8: 11 moduli; with 2-norm row scaling
 * they keep the contractions of the 20-nucleus, n_ao = 200 workload near
 * 1e-12 relative error, two orders of magnitude inside 1e-10.  bf16 / fp16
 * modular: 10 moduli (moduli <= 255, about 7.9 bits each).  Slices: 4. */
#define CUTC_INT8_DEFAULT_MODULI 11
#define CUTC_CRT16_DEFAULT_MODULI 10
#define CUTC_SLICE_DEFAULT 4
static inline int cutc_default_slices(int precision)
{
    if (precision == CUTC_PREC_INT8)
        return CUTC_INT8_DEFAULT_MODULI;
    if (precision == CUTC_PREC_BF16_CRT || precision == CUTC_PREC_FP16_CRT)
        return CUTC_CRT16_DEFAULT_MODULI;
    if (precision == CUTC_PREC_BF16 || precision == CUTC_PREC_FP16)
        return CUTC_SLICE_DEFAULT;
    return 0;
}
#define CUTC_CRT_I8 0
#define CUTC_CRT_BF16 1
#define CUTC_CRT_FP16 2
#define CUTC_CRT_RBYTES(kind) ((kind) == CUTC_CRT_I8 ? 1 : 2)
typedef struct cutc_crt {
    int n, bits;
    int m[CUTC_CRT_MAX];
    int y[CUTC_CRT_MAX];
    double inv_m[CUTC_CRT_MAX];
    double Mi[CUTC_CRT_MAX][3];
    double M[3];
    double Minv;
    int kind; /* CUTC_CRT_* */
} cutc_crt;

int cutc_crt_setup(int64_t K, int n, int kind, cutc_crt *c);

/* column-major sources for the stage-2 operands: element (row r, k) of
 * segment s is p[s][kp(k) * ld + col(r)], col(r) = k_r + n_ao l_r through the
 * pair table when pairs != NULL, else r; kp(k) = kperm[k] when kperm != NULL
 * (the r1 points of a chunk in another order), else k. */
typedef struct cutc_colsrc {
    const double *p[3];
    int nsrc;
    int64_t ld, K;
    const int *pairs;
    int n_ao;
    const uint32_t *kperm;
} cutc_colsrc;

/* Row exponents: CUTC_ROWEXP_NORM gives e with ||x_row||_2 < 2^e (the
 * modular path), CUTC_ROWEXP_MAX e with max |x_row| < 2^e (the slice path). */
#define CUTC_ROWEXP_NORM 0
#define CUTC_ROWEXP_MAX 1
int cutc_rowexp_k(const double *X, int64_t ld, int64_t rows, int64_t K, int mode, int *rexp, hipStream_t st);
/* row exponents from ks partial sums per row, partial[y rows + r] (sums of
 * squares, CUTC_ROWEXP_NORM, or maxima), joined in order */
int cutc_rowexp_finish(const double *partial, int64_t ks, int64_t rows, int mode, int *rexp, hipStream_t st);
int cutc_rowexp_pairw(const double *wr2, const double *aos2, int64_t n2, int n_ao, int mode, int *rexp,
                      hipStream_t st);
/* partial: CUTC_CRT_KSPLIT x rows doubles of scratch */
#define CUTC_CRT_KSPLIT 64
int cutc_rowexp_colk(const cutc_colsrc *src, int64_t rows, int mode, double *partial, int *rexp, hipStream_t st);
/* Fragment-packed residue planes.  int8: rows are
 * grouped in blocks of 32 and K in steps of 64 bytes; the 2048-byte tile of
 * (row block rb, step kt) sits at (rb * KT + kt) * 2048 (KT = kpad / 64) and
 * holds byte k = 64 kt + 32 h + 16 c + e of row 32 rb + r at
 * c * 1024 + (32 h + r) * 16 + e: the 16 bytes lane 32 h + r of a wave feeds
 * to the c-th v_mfma_i32_32x32x32_i8 of the step, and every 1 KB half-tile is
 * one fully coalesced wave load.  Both operands use the same K order, so the
 * dot products are unchanged.  kpad is a multiple of CUTC_PK_KALIGN (the
 * K steps of every pipeline depth of the GEMM tile it); the
 * G-side planes (A) hold CUTC_PK_ROWS_A(rows) rows, the pair-side planes (B)
 * CUTC_PK_ROWS_B(rows); rows past the valid ones are never stored.
 * 16-bit kinds: the same 2048-byte tiles, of 32 rows x 32 elements (KT =
 * kpad / 32), element k = 32 kt + 16 c + 8 h + e of row 32 rb + r at byte
 * c * 1024 + (32 h + r) * 16 + 2 e (the operands of the c-th
 * v_mfma_f32_32x32x16 of the step).  Plane bytes: rows x kpad x
 * CUTC_CRT_RBYTES(kind); every entry point below takes the kind from c. */
#define CUTC_PK_KALIGN 768
#define CUTC_PK_ROWS_A(rows) ((((int64_t)(rows)) + 511) / 512 * 512)
#define CUTC_PK_ROWS_B(rows) ((((int64_t)(rows)) + 255) / 256 * 256)
int cutc_crt_quant_k_pk(const cutc_crt *c, const double *X, int64_t ld, int64_t rows, int64_t K, int64_t kpad,
                        const int *rexp, int8_t *planes, hipStream_t st);
int cutc_crt_quant_pairw_pk(const cutc_crt *c, const double *wr2, const double *aos2, int64_t n2, int n_ao,
                            int64_t kpad, const int *rexp, int8_t *planes, hipStream_t st);
/* the same for r2 points [0, n2) of arrays whose AO columns are ld2 apart
 * (one K pass: wr2 and aos2 offset to its first point) */
int cutc_crt_quant_pairw_pk_range(const cutc_crt *c, const double *wr2, const double *aos2, int64_t ld2, int64_t n2,
                                  int n_ao, int64_t kpad, const int *rexp, int8_t *planes, hipStream_t st);
/* column-major sources into packed planes of plane_rows rows (CUTC_PK_ROWS_A or _B of rows, by the operand's side); the
 * K of segment s starts at s kp, kp a multiple of 128 and nsrc kp one of
 * CUTC_PK_KALIGN. */
int cutc_crt_quant_colk_pk(const cutc_crt *c, const cutc_colsrc *src, int64_t rows, int64_t kp, int64_t plane_rows,
                           const int *rexp, int8_t *planes, hipStream_t st);
/* the same with the non-zero flags occ[tile][kt] of every (CUTC_SP_ROWS-row
 * tile, K step) of the planes (cleared by the caller) */
int cutc_crt_quant_colk_pk_occ(const cutc_crt *c, const cutc_colsrc *src, int64_t rows, int64_t kp,
                               int64_t plane_rows, const int *rexp, int8_t *planes, uint8_t *occ, hipStream_t st);
/* C[i][r][p] = (sum_k A[i][r][k] B[i][p][k]) mod m_i as int8 in
 * [-(m_i / 2), (m_i - 1) / 2] for r < M, p < N, packed planes as above; C
 * holds n M N bytes (row r of the A side is contiguous over p).
 * accumulate: add the residues already in C (one K pass of several). */
int cutc_crt_gemm_pk(const cutc_crt *c, const int8_t *A, int64_t M, const int8_t *B, int64_t N, int64_t kpad,
                     int8_t *C, int accumulate, hipStream_t st);
/* Block-sparse stage 1 (order.hip).  The r2 points are reordered along a
 * Morton curve and the pair rows clustered by where their quantised values
 * are non-zero, so that whole (CUTC_SP_ROWS-pair B tile, K step) tiles of
 * the pair planes are zero; the quantiser flags the others
 * (occ[tb][kt], one byte each, cleared first) and the GEMM runs, per B tile,
 * only the listed steps (blocks of 256 A x CUTC_SP_ROWS B rows).  A zero
 * tile adds 0 to every modular sum: the results are those of the dense
 * GEMM, bit for bit.
 * Steps: CUTC_CRT_KSTEP(kind) K each (kpad / KSTEP per plane); lists are
 * padded to CUTC_CRT_LIST_ALIGN(kind) steps (the GEMM pipelines' unroll). */
#define CUTC_CRT_KSTEP(kind) ((kind) == CUTC_CRT_I8 ? 64 : 32)
#define CUTC_CRT_LIST_ALIGN(kind) ((kind) == CUTC_CRT_I8 ? 6 : 24)
#define CUTC_SP_ROWS 256  /* pair rows per B tile of the step lists */
/* the same GEMM over the steps klist[t KT + s], s < kcount[t], of tile t
 * (KT = kpad / KSTEP) of the B side (list_a = 0) or of the A side
 * (list_a = 1); klist = kcount = NULL: every step.  nsplit > 1 (lists
 * only): each tile's list runs in nsplit consecutive parts (whole padding
 * units), part s into its own residues at C + s n M N, whose sum modulo
 * m_i is the whole result (cutc_crt_rebuild_m adds them).  accumulate:
 * bit 0 adds the products into C (modulo m_i); CUTC_GEMM_KPASS marks one of
 * several K passes.  The moduli and parts share one launch (a launch per
 * modulus for the B-list contraction over 1024 tiles or more in int8, or
 * in a K pass) */
#define CUTC_GEMM_KPASS 2
int cutc_crt_gemm_pk_list(const cutc_crt *c, const int8_t *A, int64_t M, const int8_t *B, int64_t N, int64_t kpad,
                          const uint32_t *klist, const int *kcount, int list_a, int nsplit, int8_t *C,
                          int accumulate, hipStream_t st);
/* pair planes with (k, l) of row p from ptab (NULL: the packed pair index)
 * and, with occ, the non-zero flags of every (B tile, step) */
int cutc_crt_quant_pairw_pk_occ(const cutc_crt *c, const double *wr2, const double *aos2, int64_t ld2, int64_t n2,
                                int n_ao, int64_t kpad, const int *rexp, const int *ptab, int8_t *planes,
                                uint8_t *occ, hipStream_t st);
/* sort length for n keys: the power of two >= max(n, 2048) */
int64_t cutc_sort_len(int64_t n);
/* ascending sort of n (key, value) pairs, n = cutc_sort_len(n) */
int cutc_sort_u64(uint64_t *keys, uint32_t *vals, int64_t n, hipStream_t st);
/* point order of the n points r (n x 3): Morton (Z-order) code in their
 * bounding box (box: 6 doubles of scratch); the permutation in vals[0 .. n);
 * keys / vals hold nsort = cutc_sort_len(n) */
int cutc_order_points(const double *r, int64_t n, double *box, uint64_t *keys, uint32_t *vals, int64_t nsort,
                      hipStream_t st);
/* r2 order (cutc_order_points): perm in vals[0 .. n2) and the permuted
 * copies r2p (n2 x 3), wr2p (n2) and aos2p (n_ao x n2) of r2, wr2 and the AO
 * values aos2 (n_ao x n2) */
int cutc_order_r2(const double *r2, const double *wr2, const double *aos2, int64_t n2, int n_ao, double *box,
                  uint64_t *keys, uint32_t *vals, int64_t nsort, double *r2p, double *wr2p, double *aos2p,
                  hipStream_t st);
/* scratch bytes of the pair clustering of P rows over nsteps K steps */
size_t cutc_cluster_bytes(int64_t P, int64_t nsteps);
/* pair order: the support of every pair (the K steps of kstep points, 32 or
 * 64, where rint(w chi_k chi_l 2^(bits - rexp_nat[p])) is non-zero, nsteps
 * steps) and a greedy clustering of the pairs into tiles of CUTC_SP_ROWS
 * with similar supports; the ordered pair table pairs (2 ints per row) and
 * exponents rexp; ws: cutc_cluster_bytes(P, nsteps) bytes */
int cutc_order_pairs(const double *wr2, const double *aos2, int64_t n2, int n_ao, int bits, const int *rexp_nat,
                     int kstep, int64_t nsteps, void *ws, int *pairs, int *rexp, hipStream_t st);
/* the same over a one-segment column source (rows = packed pairs, K = the
 * chunk's points): supports of rint(x 2^(bits - rexp[row])); the ordered
 * pair table in pairs */
int cutc_order_pairs_colk(const cutc_colsrc *src, int n_ao, int bits, const int *rexp, int kstep, int64_t nsteps,
                          void *ws, int *pairs, hipStream_t st);
int cutc_zero_u8(uint8_t *p, int64_t n, hipStream_t st);
/* per-tile step lists from the flags occ (tiles x KT), KT a multiple of align */
int cutc_klists(const uint8_t *occ, int64_t tiles, int64_t KT, int align, uint32_t *list, int *count,
                hipStream_t st);
/* Slice (Ozaki scheme I) contractions, ozaki1.hip: S = 2..4 digit planes of
 * 16-bit elements (bf16 or fp16 bit patterns, f16 != 0), rows padded to
 * CUTC_OZ_ROWS, kpad a multiple of CUTC_PK_KALIGN.  Row exponents from
 * cutc_rowexp_* with CUTC_ROWEXP_MAX. */
#define CUTC_OZ_ROWS(rows) ((((int64_t)(rows)) + 127) / 128 * 128)
int cutc_oz_slice_k(int S, int f16, const double *X, int64_t ld, int64_t rows, int64_t K, int64_t kpad,
                    int64_t plane_rows, const int *rexp, uint16_t *planes, hipStream_t st);
int cutc_oz_slice_pairw(int S, int f16, const double *wr2, const double *aos2, int64_t n2, int n_ao, int64_t kpad,
                        int64_t plane_rows, const int *rexp, uint16_t *planes, hipStream_t st);
int cutc_oz_slice_colk(int S, int f16, const cutc_colsrc *src, int64_t rows, int64_t kp, int64_t plane_rows,
                       const int *rexp, uint16_t *planes, hipStream_t st);
/* C[r][c] (FP64, row-major, ld N) = sum_k A[r][k] B[c][k] from the digit planes */
int cutc_oz_gemm(int S, int f16, const uint16_t *A, int64_t M, const uint16_t *B, int64_t N, int64_t kpad,
                 const int *eA, const int *eB, double *C, hipStream_t st);
/* scatter of FP64 contraction results, as cutc_crt_rebuild_x / _m */
int cutc_oz_scatter_x(const double *C, int64_t P, int64_t N, const int *pairs, int n_ao, double *X, int64_t n1,
                      int64_t n1c, int64_t i0, hipStream_t st);
/* alpha sum_{y < nsplit} C[y][q][p] into M (C slices of P x P) */
int cutc_scatter_m(const double *C, int nsplit, int64_t P, double alpha, const int *pairs, int n_ao, double *M,
                   int mode, hipStream_t st);
/* Rebuild C = V 2^(eA + eB - 2 bits) and scatter into X (4, n1, n_ao, n_ao):
 * column j = m n1c + i1 -> X[m][i0 + i1][(k,l) and (l,k)]. */
int cutc_crt_rebuild_x(const cutc_crt *c, const int8_t *C, int64_t P, int64_t N, const int *rexpA,
                       const int *rexpB, const int *pairs, int n_ao, double *X, int64_t n1, int64_t n1c,
                       int64_t i0, hipStream_t st);
/* Rebuild alpha C into halves of the n_ao^2 x n_ao^2 matrix M: part 0 (the
 * symmetric X^3 H) into the columns a + n_ao b, part 1 (the antisymmetric
 * X^{0..2} NH) into the columns b + n_ao a, a < b, rows k + n_ao l, k <= l
 * only; mode 0 stores, 1 adds.  pairs: the packed-order table of the B side
 * (rexpA); pairsB (optional): the table of the rows of the GEMM's A side
 * (rexpB), when it differs.  cutc_crt_finalize_m then forms the full M.
 * nsplit: the parts of a K-split GEMM (C + s n P^2, s < nsplit <= 256),
 * whose residues are added before the rebuild. */
int cutc_crt_rebuild_m(const cutc_crt *c, const int8_t *C, int64_t P, const int *rexpA,
                       const int *rexpB, double alpha, const int *pairs, const int *pairsB, int n_ao, double *M,
                       int part, int mode, int nsplit, hipStream_t st);
/* M[:, ab] = S + A, M[:, ba] = S - A (a < b) and the rows k > l from k < l,
 * from the halves cutc_crt_rebuild_m writes */
int cutc_crt_finalize_m(double *M, int n_ao, hipStream_t st);

/* Dense TN GEMM (gemm_f64.hip): C = alpha A^T B + beta C, K-contiguous rows
 * (lda, ldb >= K).  K a multiple of 16 with even leading dimensions and
 * 16-byte aligned operands reads whole 16-wide K steps (operands zero past K);
 * anything else reads exactly K.  nsplit > 1 writes the unscaled partial sums
 * of cutc_gemm_tn_split_count(K, nsplit) K slices to work[y][j][i]. */
int cutc_gemm_tn_f64(const double *A, int64_t lda, const double *B, int64_t ldb, int64_t M, int64_t N, int64_t K,
                     double alpha, double beta, double *C, int64_t ldc, int nsplit, double *work, hipStream_t st);
int cutc_gemm_tn_split_count(int64_t K, int nsplit);

/* Normal ordering (no.hip): one grid chunk [g0, g0 + ngc) of the inputs
 * (w, mol, mor, G in the reference's layouts over ng points), intermediates
 * in chunk-local layouts (see no.hip), GEMM operands padded to ldk points.
 * ne = max(na, nb) weighted orbitals; nblk: blocks of cutc_no_point over all
 * chunks (the partial-sum stride). */
typedef struct cutc_no_args {
    int64_t ng, g0, ngc, ldk, nblk;
    int nm, na, nb, ne;
    const double *w, *mol, *mor, *G;
    double *J, *O, *S, *TE, *TF, *Dv, *A, *Rt, *E, *F, *Gt, *Ht, *C1x, *Gx, *C2, *D2;
} cutc_no_args;
int cutc_no_point(const cutc_no_args *a, double *part, hipStream_t st);
int cutc_no_orb(const cutc_no_args *a, hipStream_t st);
int cutc_no_pair(const cutc_no_args *a, hipStream_t st);
int cutc_no_gemv(const double *Gx, const double *C2, const double *Dv, int nm, int64_t ldk, int64_t ngc,
                 double *no_1e, hipStream_t st);
int cutc_no_add_slices(const double *work, int ns, int64_t n, double *out, hipStream_t st);
int cutc_no_sym_swap(const double *N, int nm, double *out, hipStream_t st);
int cutc_no_finish_0e(const double *part, int64_t nblk, double *no_0e, hipStream_t st);

#ifdef __cplusplus
}
#endif
#endif
