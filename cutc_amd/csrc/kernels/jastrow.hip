// Boys-Handy Jastrow pair kernel for gfx950: G(r1, r2) = grad_1 u and
// -1/2 |grad_1 u|^2 over an r1 chunk x all r2 points (FP64 throughout).
//
// Reference: src/jast_bh.cu:31-264 (kernel) and :268-303 (wrapper cutc_int_bh).
// Same formula, thresholds (1e-15 on squared distances, |c| < 1e-10 skip) and
// term algebra; the work is laid out for CDNA4 instead:
//   - lanes run along r2, so every store is a coalesced 512-byte row segment
//     (the reference puts r1 on threadIdx.x and stores with stride n2);
//   - a block owns 256 r2 points x R1B r1 points.  The e-n factors of its r1
//     points (f1A and grad f1A) are computed once into LDS, and each lane keeps
//     f2A of its own r2 point in registers across the R1B rows, so a pair costs
//     one sqrt/div for r12 plus the term algebra;
//   - accumulation is in registers, one store per output element (the
//     reference read-modify-writes global memory once per term).
#include "common.h"
#include "launch.h"

#define JAST_THREADS 256

struct JastArgs {
    const double *r1;   // (3, *) xyz rows, first row used = ii0
    const double *r2;   // (3, *) xyz rows, first row used = jj0
    const double *rn;   // (3, n_nuc)
    const double *c_bh; // (size_bh, n_nuc)
    const int *m_bh, *n_bh, *o_bh;
    const double *f2tab; // (n_nuc, ld_f2) e2-n factors, generic path only
    double *G;           // G[m * plane + i1 * ldg + i2]
    int64_t ii0, jj0;
    int64_t n1, n2;      // valid r1 rows, valid r2 columns
    int64_t n2_fill;     // columns written (>= n2; the extra ones get 0)
    int64_t ldg, plane;
    int64_t ld_f2;
    int n_nuc, size_bh;
};

// Sum over nuclei and terms for one pair, given the e-n factors.
__device__ static inline void jast_terms(const JastArgs &a, int A, double f1, double f1x, double f1y, double f1z,
                                         double f2, double g, double gx, double gy, double gz, double &ux,
                                         double &uy, double &uz)
{
    const int base = a.size_bh * A;
    for (int p = 0; p < a.size_bh; p++) {
        const int kk = base + p;
        const double c = a.c_bh[kk];
        if (fabs(c) < 1e-10)
            continue;
        const int m = a.m_bh[kk], n = a.n_bh[kk], o = a.o_bh[kk];
        double f1_mm1 = 0.0, f1_m = 1.0, f2_m = 1.0;
        double f1_nm1 = 0.0, f1_n = 1.0, f2_n = 1.0;
        if (m > 0) {
            f1_mm1 = cutc_powi(f1, m - 1);
            f1_m = f1_mm1 * f1;
            f2_m = cutc_powi(f2, m);
        }
        if (n > 0) {
            f1_nm1 = cutc_powi(f1, n - 1);
            f1_n = f1_nm1 * f1;
            f2_n = cutc_powi(f2, n);
        }
        double t1 = 0.0, t2 = 0.0;
        if (m > 0)
            t1 += (double)m * f1_mm1 * f2_n;
        if (n > 0)
            t1 += (double)n * f1_nm1 * f2_m;
        if (o > 0) {
            const double g_om1 = cutc_powi(g, o - 1);
            const double g_o = g_om1 * g;
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

// NMAX > 0: f2A held in registers (n_nuc <= NMAX).  NMAX == 0: f2A read from
// the precomputed table a.f2tab.
template <int NMAX, int R1B>
__global__ __launch_bounds__(JAST_THREADS) void jast_pair_kernel(JastArgs a)
{
    extern __shared__ double s_r1[]; // [R1B][n_nuc][4]  (f1, d/dx, d/dy, d/dz)
    const int64_t i2 = (int64_t)blockIdx.x * JAST_THREADS + threadIdx.x;
    const bool live = i2 < a.n2;
    const bool fill = i2 < a.n2_fill;

    double p2x = 0.0, p2y = 0.0, p2z = 0.0;
    if (live) {
        const double *p = a.r2 + 3 * (a.jj0 + i2);
        p2x = p[0];
        p2y = p[1];
        p2z = p[2];
    }
    double f2r[NMAX > 0 ? NMAX : 1];
    if constexpr (NMAX > 0) {
#pragma unroll
        for (int A = 0; A < NMAX; A++) {
            f2r[A] = 0.0;
            if (A < a.n_nuc)
                f2r[A] = cutc_rbar(p2x - a.rn[3 * A], p2y - a.rn[3 * A + 1], p2z - a.rn[3 * A + 2]);
        }
    }

    for (int64_t i1_0 = (int64_t)blockIdx.y * R1B; i1_0 < a.n1; i1_0 += (int64_t)gridDim.y * R1B) {
        const int nr = (a.n1 - i1_0 < R1B) ? (int)(a.n1 - i1_0) : R1B;
        __syncthreads();
        for (int t = threadIdx.x; t < nr * a.n_nuc; t += JAST_THREADS) {
            const int r = t / a.n_nuc, A = t - r * a.n_nuc;
            const double *p = a.r1 + 3 * (a.ii0 + i1_0 + r);
            double f, gx, gy, gz;
            cutc_rbar_grad(p[0] - a.rn[3 * A], p[1] - a.rn[3 * A + 1], p[2] - a.rn[3 * A + 2], f, gx, gy, gz);
            double *d = s_r1 + 4 * t;
            d[0] = f;
            d[1] = gx;
            d[2] = gy;
            d[3] = gz;
        }
        __syncthreads();
        if (!fill)
            continue;
        for (int r = 0; r < nr; r++) {
            const int64_t i1 = i1_0 + r;
            double ux = 0.0, uy = 0.0, uz = 0.0;
            if (live) {
                const double *p = a.r1 + 3 * (a.ii0 + i1);
                double g, gx, gy, gz;
                cutc_rbar_grad(p[0] - p2x, p[1] - p2y, p[2] - p2z, g, gx, gy, gz);
                const double *s = s_r1 + 4 * (int64_t)r * a.n_nuc;
                if constexpr (NMAX > 0) {
#pragma unroll
                    for (int A = 0; A < NMAX; A++) {
                        if (A < a.n_nuc)
                            jast_terms(a, A, s[4 * A], s[4 * A + 1], s[4 * A + 2], s[4 * A + 3], f2r[A], g, gx, gy,
                                       gz, ux, uy, uz);
                    }
                } else {
                    for (int A = 0; A < a.n_nuc; A++)
                        jast_terms(a, A, s[4 * A], s[4 * A + 1], s[4 * A + 2], s[4 * A + 3],
                                   a.f2tab[(int64_t)A * a.ld_f2 + i2], g, gx, gy, gz, ux, uy, uz);
                }
            }
            double *o = a.G + i1 * a.ldg + i2;
            o[0] = ux;
            o[a.plane] = uy;
            o[2 * a.plane] = uz;
            o[3 * a.plane] = -0.5 * (ux * ux + uy * uy + uz * uz);
        }
    }
}

// e2-n factors f2A(r2) for the generic (many-nuclei) path: f2tab[A * ld + i2].
__global__ void jast_f2_kernel(const double *r2, int64_t jj0, int64_t n2, const double *rn, int n_nuc, double *f2tab,
                               int64_t ld)
{
    const int64_t i2 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i2 >= n2)
        return;
    const double *p = r2 + 3 * (jj0 + i2);
    for (int A = 0; A < n_nuc; A++)
        f2tab[(int64_t)A * ld + i2] = cutc_rbar(p[0] - rn[3 * A], p[1] - rn[3 * A + 1], p[2] - rn[3 * A + 2]);
}

template <int NMAX, int R1B>
static int launch_pairs(const JastArgs &a, hipStream_t st)
{
    const size_t lds = sizeof(double) * 4 * (size_t)R1B * a.n_nuc;
    if (lds > 64 * 1024)
        return CUTC_ERR_ARG;
    int64_t gx = (a.n2_fill + JAST_THREADS - 1) / JAST_THREADS;
    int64_t gy = (a.n1 + R1B - 1) / R1B;
    if (gx == 0 || gy == 0)
        return 0;
    if (gx > 2147483647)
        return CUTC_ERR_ARG;
    if (gy > 65535)
        gy = 65535;
    hipLaunchKernelGGL((jast_pair_kernel<NMAX, R1B>), dim3((unsigned)gx, (unsigned)gy), dim3(JAST_THREADS), lds, st,
                       a);
    return hipGetLastError() == hipSuccess ? 0 : CUTC_ERR_HIP;
}

// Compute G for r1 rows [ii0, ii0 + n1) x r2 rows [jj0, jj0 + n2), written as
// G[m * plane + i1 * ldg + i2]; columns [n2, n2_fill) are set to 0.
// f2tab (n_nuc x n2_fill doubles) is scratch for n_nuc > 32 and may be NULL
// otherwise.
extern "C" int cutc_launch_jastrow(const double *r1, const double *r2, const double *rn, const double *c_bh,
                                   const int *m_bh, const int *n_bh, const int *o_bh, int n_nuc, int size_bh,
                                   int64_t ii0, int64_t n1, int64_t jj0, int64_t n2, int64_t n2_fill, double *G,
                                   int64_t ldg, int64_t plane, double *f2tab, hipStream_t st)
{
    JastArgs a;
    a.r1 = r1;
    a.r2 = r2;
    a.rn = rn;
    a.c_bh = c_bh;
    a.m_bh = m_bh;
    a.n_bh = n_bh;
    a.o_bh = o_bh;
    a.f2tab = f2tab;
    a.G = G;
    a.ii0 = ii0;
    a.jj0 = jj0;
    a.n1 = n1;
    a.n2 = n2;
    a.n2_fill = n2_fill;
    a.ldg = ldg;
    a.plane = plane;
    a.ld_f2 = n2_fill;
    a.n_nuc = n_nuc;
    a.size_bh = size_bh;
    if (n1 <= 0 || n2_fill <= 0)
        return 0;
    if (n_nuc <= 8)
        return launch_pairs<8, 16>(a, st);
    if (n_nuc <= 16)
        return launch_pairs<16, 16>(a, st);
    if (n_nuc <= 32)
        return launch_pairs<32, 16>(a, st);
    if (!f2tab)
        return CUTC_ERR_ARG;
    if (n2 > 0) {
        int64_t nb = (n2 + 255) / 256;
        hipLaunchKernelGGL(jast_f2_kernel, dim3((unsigned)nb), dim3(256), 0, st, r2, jj0, n2, rn, n_nuc, f2tab,
                           a.ld_f2);
    }
    return launch_pairs<0, 4>(a, st);
}
