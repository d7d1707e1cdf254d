/*
 * FP64 CPU ORACLE for the transcorrelated integral path -- TEST INFRASTRUCTURE.
 *
 * This file is the checker, never the product: only tests, the smoke check
 * and the CPU-baseline leg of the benchmark may load it.  The HIP library
 * (libcutcint.so) never links or calls anything here.
 *
 * It restates, in plain C + OpenMP + host DGEMM, the reference algorithm in
 * the reference's own loop order (citations are to the CuTC sources):
 *   oracle_powd_int        src/jast_bh.cu:3-28
 *   oracle_jast_bh         src/jast_bh.cu:88-262
 *   oracle_long_range      src/long_range_integ.cu:17-46
 *   oracle_herm            src/short_range_integ_herm.cu:16-45
 *   oracle_nonherm         src/short_range_integ_nonherm.cu:19-63
 *   oracle_trans_inplace   src/add_trans_inplace.cu:6-19
 *   oracle_int2_grad1_u12_ao  src/build_int2_grad1_u12_ao.cu:97-331 (chunk loop, TN DGEMMs)
 *   oracle_int_2e_ao       src/build_int_2e_ao.cu:44-124 (two NN DGEMMs + M + M^T)
 *   oracle_int_2e_ao_cols  the same for one block of columns (large shapes)
 *   oracle_dd_2e           sum D D int_2e_ao through the factors of M (large shapes)
 *   oracle_cutc_int        src/cutc_int.c:94-105,174-176,210-211 (what is returned)
 * It implements the intended semantics where the reference is defective:
 * 64-bit sizes everywhere and beta reset to 0 at the first r2 chunk of every
 * r1 chunk (the reference keeps beta = 1 after its first GEMM,
 * src/build_int2_grad1_u12_ao.cu:110-111,166).
 *
 * How each function is pinned (tests/test_oracle.py):
 *   jast_bh: central finite differences of u(r1,r2) evaluated from the
 *     Boys-Handy formula (an independent expression), the closed form of a
 *     pure e-e term, the closed form of a pure e-n term, G3 == -|G|^2/2,
 *     coincident points, |c| < 1e-10 skipping, translation invariance.
 *   integrands: elementwise closed forms, (anti)symmetry in the AO pair.
 *   DGEMM (OpenBLAS or the fallback below): numpy matmul on random operands
 *     for every transpose combination, alpha and beta.
 *   stage drivers: brute-force quadruple loops at tiny shapes, chunk
 *     invariance, linearity in the weights, X_kl == X_lk.
 *   trans_inplace: symmetry, diagonal doubling.
 *   int_2e_ao_cols and dd_2e: against the columns of oracle_int_2e_ao and an
 *     einsum over it (tests/test_oracle.py), so that a wrong offset, a
 *     missing transpose or a dropped non-herm term fails.
 * The reference ships no golden vectors; its only known answer (all-zero
 * inputs with c = 0 give all-zero outputs, src/cutc_int_f.f90:41-55 with
 * src/jast_bh.cu:203) is checked too.
 */
#include <dlfcn.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------ DGEMM */

typedef void (*fdgemm64_t)(const char *, const char *, const int64_t *, const int64_t *, const int64_t *,
                           const double *, const double *, const int64_t *, const double *, const int64_t *,
                           const double *, double *, const int64_t *, size_t, size_t);

static fdgemm64_t g_blas_dgemm = NULL;

/* Load an ILP64 Fortran-ABI dgemm (e.g. numpy's bundled OpenBLAS,
 * symbol scipy_dgemm_64_).  Returns 0 on success. */
int oracle_use_blas(const char *path, const char *symbol)
{
    void *h = dlopen(path, RTLD_NOW | RTLD_LOCAL);
    if (!h)
        return -1;
    void *f = dlsym(h, symbol);
    if (!f)
        return -2;
    g_blas_dgemm = (fdgemm64_t)f;
    return 0;
}

void oracle_use_builtin_dgemm(void) { g_blas_dgemm = NULL; }
int oracle_blas_loaded(void) { return g_blas_dgemm != NULL; }

int oracle_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* Column-major C = alpha op(A) op(B) + beta C (BLAS semantics), the
 * project's own blocked OpenMP fallback. */
static void dgemm_builtin(char ta, char tb, int64_t m, int64_t n, int64_t k, double alpha, const double *A,
                          int64_t lda, const double *B, int64_t ldb, double beta, double *C, int64_t ldc)
{
    const int tA = (ta == 'T' || ta == 't'), tB = (tb == 'T' || tb == 't');
    const int64_t MB = 64, NB = 64, KB = 256;
#pragma omp parallel
    {
        double *ap = malloc(sizeof(double) * MB * KB);
        double *bp = malloc(sizeof(double) * KB * NB);
        double *cb = malloc(sizeof(double) * MB * NB);
#pragma omp for collapse(2) schedule(dynamic)
        for (int64_t j0 = 0; j0 < n; j0 += NB) {
            for (int64_t i0 = 0; i0 < m; i0 += MB) {
                int64_t nb = (n - j0 < NB) ? n - j0 : NB;
                int64_t mb = (m - i0 < MB) ? m - i0 : MB;
                memset(cb, 0, sizeof(double) * MB * NB);
                for (int64_t p0 = 0; p0 < k; p0 += KB) {
                    int64_t kb = (k - p0 < KB) ? k - p0 : KB;
                    /* ap[i*KB + p] = op(A)(i0+i, p0+p) ; bp[j*KB + p] = op(B)(p0+p, j0+j) */
                    for (int64_t i = 0; i < mb; i++)
                        for (int64_t p = 0; p < kb; p++)
                            ap[i * KB + p] = tA ? A[(p0 + p) + lda * (i0 + i)] : A[(i0 + i) + lda * (p0 + p)];
                    for (int64_t j = 0; j < nb; j++)
                        for (int64_t p = 0; p < kb; p++)
                            bp[j * KB + p] = tB ? B[(j0 + j) + ldb * (p0 + p)] : B[(p0 + p) + ldb * (j0 + j)];
                    for (int64_t j = 0; j < nb; j++)
                        for (int64_t i = 0; i < mb; i++) {
                            const double *a = ap + i * KB, *b = bp + j * KB;
                            double s = 0.0;
                            for (int64_t p = 0; p < kb; p++)
                                s += a[p] * b[p];
                            cb[i + MB * j] += s;
                        }
                }
                for (int64_t j = 0; j < nb; j++)
                    for (int64_t i = 0; i < mb; i++) {
                        double *c = C + (i0 + i) + ldc * (j0 + j);
                        *c = alpha * cb[i + MB * j] + (beta == 0.0 ? 0.0 : beta * *c);
                    }
            }
        }
        free(ap);
        free(bp);
        free(cb);
    }
}

void oracle_dgemm(char ta, char tb, int64_t m, int64_t n, int64_t k, double alpha, const double *A, int64_t lda,
                  const double *B, int64_t ldb, double beta, double *C, int64_t ldc)
{
    if (m <= 0 || n <= 0)
        return;
    if (k <= 0) {
        for (int64_t j = 0; j < n; j++)
            for (int64_t i = 0; i < m; i++)
                C[i + ldc * j] = (beta == 0.0) ? 0.0 : beta * C[i + ldc * j];
        return;
    }
    if (g_blas_dgemm) {
        g_blas_dgemm(&ta, &tb, &m, &n, &k, &alpha, A, &lda, B, &ldb, &beta, C, &ldc, 1, 1);
        return;
    }
    dgemm_builtin(ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
}

/* --------------------------------------------------------------- Jastrow */

/* a^b for the small non-negative integer exponents of the Boys-Handy table,
 * with the reference's multiplication order for b <= 6 (src/jast_bh.cu:3-28). */
double oracle_powd_int(double a, int b)
{
    double a2;
    switch (b) {
    case 0: return 1.0;
    case 1: return a;
    case 2: return a * a;
    case 3: return a * a * a;
    case 4: a2 = a * a; return a2 * a2;
    case 5: a2 = a * a; return a * a2 * a2;
    case 6: a2 = a * a; return a2 * a2 * a2;
    default: {
        double r = 1.0;
        for (int i = 0; i < b; i++)
            r *= a;
        return r;
    }
    }
}

/* rbar(d) = d / (1 + d) for a displacement (dx,dy,dz); its gradient with
 * respect to the first point is (dx,dy,dz) / (d (1 + d)^2).  Squared
 * distances <= 1e-15 give 0 and a zero gradient (src/jast_bh.cu:139-156). */
static inline void rbar_grad(double dx, double dy, double dz, double *f, double *gx, double *gy, double *gz)
{
    double d2 = dx * dx + dy * dy + dz * dz;
    if (d2 > 1e-15) {
        double d = sqrt(d2);
        double t1 = 1.0 / (1.0 + d);
        double t2 = t1 * t1 / d;
        *f = d * t1;
        *gx = t2 * dx;
        *gy = t2 * dy;
        *gz = t2 * dz;
    } else {
        *f = *gx = *gy = *gz = 0.0;
    }
}

/* G[4][n1_eff][n2_eff] (r2 fastest) for r1 rows ii0.. and r2 rows jj0..:
 * G0..2 = grad_1 u(r1, r2), G3 = -1/2 |grad_1 u|^2 (src/jast_bh.cu:38-43,88-262). */
void oracle_jast_bh(int64_t ii0, int64_t n1_eff, int64_t jj0, int64_t n2_eff, int n_nuc, int size_bh,
                    const double *r1, const double *r2, const double *rn, const double *c_bh, const int *m_bh,
                    const int *n_bh, const int *o_bh, double *G)
{
    const size_t plane = (size_t)n1_eff * (size_t)n2_eff;
#pragma omp parallel for schedule(static)
    for (int64_t i1 = 0; i1 < n1_eff; i1++) {
        const double *p1 = r1 + 3 * (ii0 + i1);
        for (int64_t i2 = 0; i2 < n2_eff; i2++) {
            const double *p2 = r2 + 3 * (jj0 + i2);
            double g, gx, gy, gz;
            rbar_grad(p1[0] - p2[0], p1[1] - p2[1], p1[2] - p2[2], &g, &gx, &gy, &gz);
            double ux = 0.0, uy = 0.0, uz = 0.0;
            for (int a = 0; a < n_nuc; a++) {
                const double *A = rn + 3 * a;
                double f1, f1x, f1y, f1z, f2;
                rbar_grad(p1[0] - A[0], p1[1] - A[1], p1[2] - A[2], &f1, &f1x, &f1y, &f1z);
                /* e2-n term: value only, with the reference's direct quotient (src/jast_bh.cu:185-194) */
                {
                    double dx = p2[0] - A[0], dy = p2[1] - A[1], dz = p2[2] - A[2];
                    double d2 = dx * dx + dy * dy + dz * dz;
                    if (d2 > 1e-15) {
                        double d = sqrt(d2);
                        f2 = d / (1.0 + d);
                    } else {
                        f2 = 0.0;
                    }
                }
                for (int p = 0; p < size_bh; p++) {
                    int kk = p + size_bh * a;
                    double c = c_bh[kk];
                    if (fabs(c) < 1e-10)
                        continue;
                    int m = m_bh[kk], n = n_bh[kk], o = o_bh[kk];
                    double f1_mm1 = 0.0, f1_m = 1.0, f2_m = 1.0;
                    double f1_nm1 = 0.0, f1_n = 1.0, f2_n = 1.0;
                    if (m > 0) {
                        f1_mm1 = oracle_powd_int(f1, m - 1);
                        f1_m = f1_mm1 * f1;
                        f2_m = oracle_powd_int(f2, m);
                    }
                    if (n > 0) {
                        f1_nm1 = oracle_powd_int(f1, n - 1);
                        f1_n = f1_nm1 * f1;
                        f2_n = oracle_powd_int(f2, n);
                    }
                    /* d/dr1 of (f1^m f2^n + f1^n f2^m) g^o, split into the e-n part (t1)
                     * and the e-e part (t2) -- src/jast_bh.cu:225-246 */
                    double t1 = 0.0, t2 = 0.0;
                    if (m > 0)
                        t1 += (double)m * f1_mm1 * f2_n;
                    if (n > 0)
                        t1 += (double)n * f1_nm1 * f2_m;
                    if (o > 0) {
                        double g_om1 = oracle_powd_int(g, o - 1);
                        double g_o = g_om1 * g;
                        t2 = c * (double)o * g_om1 * (f1_m * f2_n + f1_n * f2_m);
                        t1 = c * t1 * g_o;
                    } else {
                        t1 *= c;
                    }
                    ux += t1 * f1x + t2 * gx;
                    uy += t1 * f1y + t2 * gy;
                    uz += t1 * f1z + t2 * gz;
                }
            }
            size_t q = (size_t)i1 * n2_eff + i2;
            G[q] = ux;
            G[q + plane] = uy;
            G[q + 2 * plane] = uz;
            G[q + 3 * plane] = -0.5 * (ux * ux + uy * uy + uz * uz);
        }
    }
}

/* ------------------------------------------------------------ integrands */

/* A[i2 + ld*(j + n_ao*i)] = w2 chi_i chi_j for r2 rows jj0..jj0+n2_eff
 * (src/long_range_integ.cu:17-46; aos2 is (n2, n_ao, *), value component). */
void oracle_long_range(int64_t jj0, int64_t n2_eff, int64_t ld, int64_t n2, int n_ao, const double *wr2,
                       const double *aos2, double *A)
{
    /* every element is the same product w * vi * vj; the column loop is
     * outside so that the stores run along i2, contiguous in A */
#pragma omp parallel for collapse(2) schedule(static)
    for (int i = 0; i < n_ao; i++) {
        for (int j = 0; j < n_ao; j++) {
            double *col = A + ld * ((int64_t)j + (int64_t)n_ao * i);
            const double *ai = aos2 + jj0 + n2 * (int64_t)i, *aj = aos2 + jj0 + n2 * (int64_t)j;
            const double *w = wr2 + jj0;
            for (int64_t i2 = 0; i2 < n2_eff; i2++) {
                double vi = ai[i2], vj = aj[i2];
                col[i2] = w[i2] * vi * vj;
            }
        }
    }
}

/* H[i1 + n1*(j + n_ao*i)] = w1 chi_i chi_j (src/short_range_integ_herm.cu:16-45). */
void oracle_herm(int64_t n1, int n_ao, const double *wr1, const double *aos1, double *H)
{
#pragma omp parallel for schedule(static)
    for (int64_t i1 = 0; i1 < n1; i1++) {
        double w = wr1[i1];
        for (int i = 0; i < n_ao; i++) {
            double vi = aos1[i1 + n1 * i];
            for (int j = 0; j < n_ao; j++)
                H[i1 + n1 * ((int64_t)j + (int64_t)n_ao * i)] = w * vi * aos1[i1 + n1 * j];
        }
    }
}

/* NH[(i1 + m*n1) + 3*n1*(j + n_ao*i)] = w1 (chi_j d_m chi_i - chi_i d_m chi_j)
 * (src/short_range_integ_nonherm.cu:19-63; derivative m is aos1 component m+1). */
void oracle_nonherm(int64_t n1, int n_ao, const double *wr1, const double *aos1, double *NH)
{
    const int64_t comp = n1 * (int64_t)n_ao;
#pragma omp parallel for schedule(static)
    for (int64_t i1 = 0; i1 < n1; i1++) {
        double w = wr1[i1];
        for (int i = 0; i < n_ao; i++) {
            double vi = aos1[i1 + n1 * i];
            for (int j = 0; j < n_ao; j++) {
                double vj = aos1[i1 + n1 * j];
                for (int m = 0; m < 3; m++) {
                    double di = aos1[i1 + n1 * i + (m + 1) * comp];
                    double dj = aos1[i1 + n1 * j + (m + 1) * comp];
                    NH[(i1 + m * n1) + 3 * n1 * ((int64_t)j + (int64_t)n_ao * i)] = w * (vj * di - vi * dj);
                }
            }
        }
    }
}

/* data <- data + data^T for a size x size column-major matrix
 * (src/add_trans_inplace.cu:6-19). */
void oracle_trans_inplace(double *data, int64_t size)
{
    /* same pairwise update, visited in 64 x 64 tile pairs for cache reuse */
    const int64_t T = 64, nt = (size + T - 1) / T;
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t bi = 0; bi < nt; bi++) {
        for (int64_t bj = 0; bj <= bi; bj++) {
            int64_t i1 = (bi + 1) * T < size ? (bi + 1) * T : size;
            int64_t j1 = (bj + 1) * T < size ? (bj + 1) * T : size;
            for (int64_t j = bj * T; j < j1; j++)
                for (int64_t i = bi * T; i < i1; i++) {
                    if (j > i)
                        continue;
                    double s = data[i + size * j] + data[j + size * i];
                    data[i + size * j] = s;
                    data[j + size * i] = s;
                }
        }
    }
}

/* ---------------------------------------------------------- stage drivers */

/* X[4][n1][n_ao^2] (X^m at n_ao^2 (r1 + m n1)), chunked over r1 (n1c) and r2
 * (n2c) in the reference's loop order (src/build_int2_grad1_u12_ao.cu:97-331):
 * for each r1 chunk, for each r2 chunk: long-range integrand, Jastrow pair
 * block, then 4 DGEMM(T, N, n_ao^2, n1 chunk, n2 chunk) with beta = 0 at the
 * first r2 chunk and 1 after.  n1c/n2c <= 0 mean "one chunk". */
int oracle_int2_grad1_u12_ao(int64_t n1, int64_t n2, int n_ao, int n_nuc, int size_bh, const double *r1,
                             const double *r2, const double *wr2, const double *rn, const double *aos2,
                             const double *c_bh, const int *m_bh, const int *n_bh, const int *o_bh, double *X,
                             int64_t n1c, int64_t n2c)
{
    const int64_t nao2 = (int64_t)n_ao * n_ao;
    if (n1c <= 0 || n1c > n1)
        n1c = n1;
    if (n2c <= 0 || n2c > n2)
        n2c = n2;
    double *A = malloc(sizeof(double) * (size_t)nao2 * n2c);
    double *G = malloc(sizeof(double) * 4 * (size_t)n1c * n2c);
    if (!A || !G) {
        free(A);
        free(G);
        return -3;
    }
    for (int64_t ii = 0; ii < n1; ii += n1c) {
        int64_t n1e = (n1 - ii < n1c) ? n1 - ii : n1c;
        for (int64_t jj = 0; jj < n2; jj += n2c) {
            int64_t n2e = (n2 - jj < n2c) ? n2 - jj : n2c;
            oracle_long_range(jj, n2e, n2c, n2, n_ao, wr2, aos2, A);
            oracle_jast_bh(ii, n1e, jj, n2e, n_nuc, size_bh, r1, r2, rn, c_bh, m_bh, n_bh, o_bh, G);
            double beta = (jj == 0) ? 0.0 : 1.0;
            for (int m = 0; m < 4; m++)
                oracle_dgemm('T', 'N', nao2, n1e, n2e, 1.0, A, n2c, G + (size_t)m * n1e * n2e, n2e, beta,
                             X + (size_t)nao2 * (ii + m * n1), nao2);
        }
    }
    free(A);
    free(G);
    return 0;
}

/* int_2e_ao (n_ao^2 x n_ao^2, column-major) = M + M^T with
 * M = X^3 H - 1/2 X^{0..2} NH (src/build_int_2e_ao.cu:44-124). */
int oracle_int_2e_ao(int64_t n1, int n_ao, const double *wr1, const double *aos1, const double *X, double *out)
{
    const int64_t nao2 = (int64_t)n_ao * n_ao;
    double *H = malloc(sizeof(double) * (size_t)n1 * nao2);
    if (!H)
        return -3;
    oracle_herm(n1, n_ao, wr1, aos1, H);
    oracle_dgemm('N', 'N', nao2, nao2, n1, 1.0, X + (size_t)3 * nao2 * n1, nao2, H, n1, 0.0, out, nao2);
    free(H);
    double *NH = malloc(sizeof(double) * 3 * (size_t)n1 * nao2);
    if (!NH)
        return -3;
    oracle_nonherm(n1, n_ao, wr1, aos1, NH);
    oracle_dgemm('N', 'N', nao2, nao2, 3 * n1, -0.5, X, nao2, NH, 3 * n1, 1.0, out, nao2);
    free(NH);
    oracle_trans_inplace(out, nao2);
    return 0;
}

/* Columns [c0, c0 + nc) of int_2e_ao = M + M^T, as an n_ao^2 x nc
 * column-major block (ld n_ao^2), without forming the n_ao^4 array: the two
 * NN DGEMMs of src/build_int_2e_ao.cu:56-124 restricted to those columns of
 * M, the same two DGEMMs for the matching rows of M, and the pairwise sum
 * of src/add_trans_inplace.cu:6-19 for that block.  For the large shapes
 * whose n_ao^4 array the tests do not hold on the host. */
int oracle_int_2e_ao_cols(int64_t n1, int n_ao, const double *wr1, const double *aos1, const double *X, int64_t c0,
                          int64_t nc, double *out)
{
    const int64_t nao2 = (int64_t)n_ao * n_ao;
    if (c0 < 0 || nc < 0 || c0 + nc > nao2)
        return -2;
    if (nc == 0)
        return 0;
    double *H = malloc(sizeof(double) * (size_t)n1 * nao2);
    double *NH = malloc(sizeof(double) * 3 * (size_t)n1 * nao2);
    double *R = malloc(sizeof(double) * (size_t)nc * nao2); /* rows c0.. of M, nc x n_ao^2, ld nc */
    if (!H || !NH || !R) {
        free(H);
        free(NH);
        free(R);
        return -3;
    }
    oracle_herm(n1, n_ao, wr1, aos1, H);
    oracle_nonherm(n1, n_ao, wr1, aos1, NH);
    const double *X3 = X + (size_t)3 * nao2 * n1;
    /* M[:, c0..] */
    oracle_dgemm('N', 'N', nao2, nc, n1, 1.0, X3, nao2, H + (size_t)n1 * c0, n1, 0.0, out, nao2);
    oracle_dgemm('N', 'N', nao2, nc, 3 * n1, -0.5, X, nao2, NH + (size_t)3 * n1 * c0, 3 * n1, 1.0, out, nao2);
    /* M[c0.., :] */
    oracle_dgemm('N', 'N', nc, nao2, n1, 1.0, X3 + c0, nao2, H, n1, 0.0, R, nc);
    oracle_dgemm('N', 'N', nc, nao2, 3 * n1, -0.5, X + c0, nao2, NH, 3 * n1, 1.0, R, nc);
#pragma omp parallel for schedule(static)
    for (int64_t j = 0; j < nc; j++)
        for (int64_t i = 0; i < nao2; i++)
            out[i + nao2 * j] += R[j + nc * i];
    free(H);
    free(NH);
    free(R);
    return 0;
}

/* sum_{kl,ab} d_kl d_ab int_2e_ao[kl][ab] for a vector d of n_ao^2 entries,
 * without forming int_2e_ao: with int_2e_ao = M + M^T and
 * M = X^3 H - 1/2 X^{0..2} NH (src/build_int_2e_ao.cu:56-124) the sum is
 * 2 d^T M d = 2 [ (X^3T d) . (H^T d) - 1/2 (X^{0..2}T d) . (NH^T d) ],
 * four DGEMM-vector products. */
int oracle_dd_2e(int64_t n1, int n_ao, const double *wr1, const double *aos1, const double *X, const double *d,
                 double *result)
{
    const int64_t nao2 = (int64_t)n_ao * n_ao;
    double *H = malloc(sizeof(double) * (size_t)n1 * nao2);
    double *NH = malloc(sizeof(double) * 3 * (size_t)n1 * nao2);
    double *u = malloc(sizeof(double) * (3 * (size_t)n1 + 1));
    double *v = malloc(sizeof(double) * (3 * (size_t)n1 + 1));
    if (!H || !NH || !u || !v) {
        free(H);
        free(NH);
        free(u);
        free(v);
        return -3;
    }
    oracle_herm(n1, n_ao, wr1, aos1, H);
    oracle_nonherm(n1, n_ao, wr1, aos1, NH);
    double s = 0.0;
    oracle_dgemm('T', 'N', n1, 1, nao2, 1.0, X + (size_t)3 * nao2 * n1, nao2, d, nao2, 0.0, u, n1);
    oracle_dgemm('N', 'N', n1, 1, nao2, 1.0, H, n1, d, nao2, 0.0, v, n1);
    for (int64_t r = 0; r < n1; r++)
        s += u[r] * v[r];
    oracle_dgemm('T', 'N', 3 * n1, 1, nao2, 1.0, X, nao2, d, nao2, 0.0, u, 3 * n1);
    oracle_dgemm('N', 'N', 3 * n1, 1, nao2, 1.0, NH, 3 * n1, d, nao2, 0.0, v, 3 * n1);
    for (int64_t r = 0; r < 3 * n1; r++)
        s -= 0.5 * u[r] * v[r];
    *result = 2.0 * s;
    free(H);
    free(NH);
    free(u);
    free(v);
    return 0;
}

/* What cutc_int returns (src/cutc_int.c:94-105,174-176,210-211): components
 * 0..2 of X (3 n_ao^2 n1 doubles) and int_2e_ao (n_ao^4 doubles). */
int oracle_cutc_int(int64_t n1, int64_t n2, int n_ao, int n_nuc, int size_bh, const double *r1, const double *wr1,
                    const double *r2, const double *wr2, const double *rn, const double *aos1, const double *aos2,
                    const double *c_bh, const int *m_bh, const int *n_bh, const int *o_bh, double *X3out,
                    double *int_2e_ao, int64_t n1c, int64_t n2c)
{
    const int64_t nao2 = (int64_t)n_ao * n_ao;
    double *X = malloc(sizeof(double) * 4 * (size_t)nao2 * n1);
    if (!X)
        return -3;
    int rc = oracle_int2_grad1_u12_ao(n1, n2, n_ao, n_nuc, size_bh, r1, r2, wr2, rn, aos2, c_bh, m_bh, n_bh, o_bh,
                                      X, n1c, n2c);
    if (rc == 0)
        rc = oracle_int_2e_ao(n1, n_ao, wr1, aos1, X, int_2e_ao);
    if (rc == 0)
        memcpy(X3out, X, sizeof(double) * 3 * (size_t)nao2 * n1);
    free(X);
    return rc;
}
