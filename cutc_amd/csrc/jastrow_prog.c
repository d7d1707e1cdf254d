/*
 * Boys-Handy term table -> monomial program for the Jastrow pair kernel
 * (pure host code).
 *
 * The reference evaluates, per pair and per (nucleus A, term p), with
 * f1 = rbar(r1 - R_A), f2 = rbar(r2 - R_A), g = rbar(r12)
 * (src/jast_bh.cu:198-246):
 *   t1 = m f1^(m-1) f2^n [m > 0] + n f1^(n-1) f2^m [n > 0]
 *   grad_1 u += c t1 g^o grad f1 + c o g^(o-1) (f1^m f2^n + f1^n f2^m) grad g  [o > 0]
 * skipping |c| < 1e-10.  Expanding gives monomials coef * f1^i f2^j g^k that
 * multiply either grad f1 ("alpha" part) or grad g ("beta" part).  The
 * program sorts them by what they depend on:
 *   j == 0 : r1-only factor; summed over nuclei per r1 into Qbar_k (3-vector,
 *            alpha) and Rbar_k (scalar, beta), k = 0..kmax;
 *   j >= 1 : class (j, k); per (r1, A) a 3-vector Q and a scalar R.
 * The pair kernel then needs, per pair, g and its powers and per (class, A)
 * one f2^j: sum_k g^k (Qbar_k + Rbar_k grad g)
 *        + sum_class g^k sum_A f2A^j (Q_A + R_A grad g).
 * Same sums as the reference, summed in a different order.
 *
 * Per-r1 table layout (ts doubles per r1 point):
 *   [4k .. 4k+3]                       Qbar_k (x, y, z), Rbar_k     k = 0..kmax
 *   [4(kmax+1) + 4(c n_nuc + A) .. +3] Q (x, y, z), R of class c, nucleus A
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "cutc_int.h"
#include "kernels/launch.h"

static int find_or_add_class(int *cls, int *n_cls, int j, int k, int cap)
{
    for (int c = 0; c < *n_cls; c++)
        if (cls[2 * c] == j && cls[2 * c + 1] == k)
            return c;
    if (*n_cls >= cap)
        return -1;
    cls[2 * *n_cls] = j;
    cls[2 * *n_cls + 1] = k;
    return (*n_cls)++;
}

/* raw monomials: A, i, j, k, is_r, coef */
typedef struct {
    int A, i, j, k, is_r;
    double coef;
} raw_t;

/* The expansion into caller memory: raw (4 nt + 1 entries) and cls (2 (4 nt
 * + 1) ints) are work space, prog->mono (4 nt + 1 entries) and prog->cls
 * (2 (4 nt + 1) ints) must already point at the output arrays. */
static int program_core(int n_nuc, int size_bh, const double *c_bh, const int *m_bh, const int *n_bh,
                        const int *o_bh, raw_t *raw, int *cls, cutc_jprog *prog)
{
    const int64_t nt = (int64_t)n_nuc * size_bh;
    int64_t nr = 0;
    int kmax = 0, n_cls = 0;
    for (int A = 0; A < n_nuc; A++) {
        for (int p = 0; p < size_bh; p++) {
            const int64_t kk = p + (int64_t)size_bh * A;
            const double c = c_bh[kk];
            if (fabs(c) < 1e-10) /* src/jast_bh.cu:203-204 */
                continue;
            const int m = m_bh[kk], n = n_bh[kk], o = o_bh[kk];
            if (m > 0)
                raw[nr++] = (raw_t){A, m - 1, n, o, 0, c * m};
            if (n > 0)
                raw[nr++] = (raw_t){A, n - 1, m, o, 0, c * n};
            if (o > 0) {
                raw[nr++] = (raw_t){A, m, n, o - 1, 1, c * o};
                raw[nr++] = (raw_t){A, n, m, o - 1, 1, c * o};
            }
        }
    }
    for (int64_t t = 0; t < nr; t++)
        if (raw[t].j == 0 && raw[t].k > kmax)
            kmax = raw[t].k;
    for (int64_t t = 0; t < nr; t++)
        if (raw[t].j > 0 && find_or_add_class(cls, &n_cls, raw[t].j, raw[t].k, (int)(4 * nt + 1)) < 0)
            return CUTC_ERR_ARG;
    const int base = 4 * (kmax + 1);
    for (int64_t t = 0; t < nr; t++) {
        cutc_jmono *mo = &prog->mono[t];
        mo->A = raw[t].A;
        mo->i = raw[t].i;
        mo->is_r = raw[t].is_r;
        mo->coef = raw[t].coef;
        if (raw[t].j == 0)
            mo->slot = 4 * raw[t].k;
        else
            mo->slot = base + 4 * (find_or_add_class(cls, &n_cls, raw[t].j, raw[t].k, n_cls) * n_nuc + raw[t].A);
    }
    memcpy(prog->cls, cls, sizeof(int) * 2 * n_cls);
    prog->n_mono = (int)nr;
    prog->n_cls = n_cls;
    prog->kmax = kmax;
    prog->ts = base + 4 * (int64_t)n_cls * n_nuc;
    return CUTC_OK;
}

int cutc_jast_program(int n_nuc, int size_bh, const double *c_bh, const int *m_bh, const int *n_bh, const int *o_bh,
                      cutc_jprog *prog)
{
    memset(prog, 0, sizeof *prog);
    if (n_nuc < 0 || size_bh < 0)
        return CUTC_ERR_ARG;
    const int64_t nt = (int64_t)n_nuc * size_bh;
    raw_t *raw = malloc(sizeof(raw_t) * (4 * nt + 1));
    int *cls = malloc(sizeof(int) * 2 * (4 * nt + 1));
    prog->mono = malloc(sizeof(cutc_jmono) * (4 * nt + 1));
    prog->cls = malloc(sizeof(int) * 2 * (4 * nt + 1));
    int rc = (raw && cls && prog->mono && prog->cls) ? CUTC_OK : CUTC_ERR_NOMEM;
    if (rc == CUTC_OK)
        rc = program_core(n_nuc, size_bh, c_bh, m_bh, n_bh, o_bh, raw, cls, prog);
    free(raw);
    free(cls);
    if (rc != CUTC_OK)
        cutc_jast_program_free(prog);
    return rc;
}

/* Host work space of cutc_jast_program_static: terms of up to
 * CUTC_JPROG_STATIC_TERMS (n_nuc x size_bh).  One set for the process; the
 * library's callers hold the library lock while they use it. */
static raw_t g_raw[4 * CUTC_JPROG_STATIC_TERMS + 1];
static int g_cls_work[2 * (4 * CUTC_JPROG_STATIC_TERMS + 1)];
static cutc_jmono g_mono[4 * CUTC_JPROG_STATIC_TERMS + 1];
static int g_cls[2 * (4 * CUTC_JPROG_STATIC_TERMS + 1)];

int cutc_jast_program_static(int n_nuc, int size_bh, const double *c_bh, const int *m_bh, const int *n_bh,
                             const int *o_bh, cutc_jprog *prog)
{
    memset(prog, 0, sizeof *prog);
    if (n_nuc < 0 || size_bh < 0 || (int64_t)n_nuc * size_bh > CUTC_JPROG_STATIC_TERMS)
        return CUTC_ERR_ARG;
    prog->mono = g_mono;
    prog->cls = g_cls;
    return program_core(n_nuc, size_bh, c_bh, m_bh, n_bh, o_bh, g_raw, g_cls_work, prog);
}

void cutc_jast_program_free(cutc_jprog *prog)
{
    free(prog->mono);
    free(prog->cls);
    prog->mono = NULL;
    prog->cls = NULL;
}
