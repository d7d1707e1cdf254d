// Boys-Handy Jastrow pair kernels for gfx950: G(r1, r2) = grad_1 u and
// -1/2 |grad_1 u|^2 over an r1 chunk x all r2 points (FP64 throughout).
//
// Reference: src/jast_bh.cu:31-264 (kernel) and :268-303 (wrapper).  Same
// formula, thresholds (1e-15 on squared distances, |c| < 1e-10 skip) and
// special cases; the work is reorganised for CDNA4:
//   - the term table is expanded once on the host into monomials
//     coef f1^i f2^j g^k (jastrow_prog.c).  Everything that depends on r1
//     alone (f1A, grad f1A and their powers, the nucleus sums of the r1-only
//     monomials) is computed once per r1 point by jast_prep_kernel into a
//     table row, so a pair costs one sqrt/div for r12, the powers of g, and
//     for each (class, nucleus) one f2^j times a 4-wide FMA;
//   - the table row is uniform across a block and is read with scalar loads;
//   - lanes run along r2: every store is a coalesced row segment, and each
//     lane keeps f2A of its r2 point in registers across R1B rows;
//   - accumulation in registers, one store per output element.
#include "common.h"
#include "launch.h"

#define JAST_THREADS 256

// sum over the 64 lanes of a wave (the device library's DPP reduction, a
// fixed order), returned to every lane
extern "C" __device__ double __ockl_wfred_add_f64(double);

// Row i1 of the per-r1 table: zero it, then add every monomial.
__global__ __launch_bounds__(256) void jast_prep_kernel(const double *__restrict__ r1, int64_t ii0, int64_t n1,
                                                        const double *__restrict__ rn, int n_nuc,
                                                        const cutc_jmono *__restrict__ mono, int n_mono,
                                                        double *__restrict__ tab, int64_t ts)
{
    const int64_t i1 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i1 >= n1)
        return;
    double *row = tab + i1 * ts;
    for (int64_t t = 0; t < ts; t++)
        row[t] = 0.0;
    const double *p = r1 + 3 * (ii0 + i1);
    const double x = p[0], y = p[1], z = p[2];
    int Aprev = -1;
    double f = 0.0, gx = 0.0, gy = 0.0, gz = 0.0;
    for (int t = 0; t < n_mono; t++) {
        const cutc_jmono mo = mono[t];
        if (mo.A != Aprev) {
            const double *R = rn + 3 * mo.A;
            cutc_rbar_grad(x - R[0], y - R[1], z - R[2], f, gx, gy, gz);
            Aprev = mo.A;
        }
        double fi = 1.0;
        for (int e = 0; e < mo.i; e++)
            fi *= f;
        const double v = mo.coef * fi;
        if (mo.is_r) {
            row[mo.slot + 3] += v;
        } else {
            row[mo.slot] += v * gx;
            row[mo.slot + 1] += v * gy;
            row[mo.slot + 2] += v * gz;
        }
    }
}

extern "C" int cutc_launch_jast_prep(const double *r1, int64_t ii0, int64_t n1, const double *rn, int n_nuc,
                                     const cutc_jmono *mono, int n_mono, double *tab, int64_t ts, hipStream_t st)
{
    if (n1 <= 0)
        return 0;
    hipLaunchKernelGGL(jast_prep_kernel, dim3((unsigned)((n1 + 255) / 256)), dim3(256), 0, st, r1, ii0, n1, rn,
                       n_nuc, mono, n_mono, tab, ts);
    return hipGetLastError() == hipSuccess ? 0 : CUTC_ERR_HIP;
}

// e2-n factors f2A(r2) for the many-nuclei path: f2tab[A * ld + i2].
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

__device__ static inline double ipow_u(double x, int e)
{
    double r = 1.0;
    for (int t = 0; t < e; t++)
        r *= x;
    return r;
}

// NMAX > 0: f2A of the lane's r2 point held in registers (n_nuc <= NMAX).
// NMAX == 0: f2A read from f2tab.
template <int NMAX, int R1B>
__global__ __launch_bounds__(JAST_THREADS) void jast_pair_kernel(
    const double *__restrict__ r1, const double *__restrict__ r2, const double *__restrict__ rn, int n_nuc,
    const double *__restrict__ tab, int64_t ts, const int *__restrict__ cls, int n_cls, int kmax,
    const double *__restrict__ f2tab, int64_t ld_f2, double *__restrict__ G, int64_t ii0, int64_t n1, int64_t jj0,
    int64_t n2, int64_t n2_fill, int64_t ldg, int64_t plane)
{
    const int64_t i2 = (int64_t)blockIdx.x * JAST_THREADS + threadIdx.x;
    if (i2 >= n2_fill)
        return;
    const bool live = i2 < n2;
    double p2x = 0.0, p2y = 0.0, p2z = 0.0;
    if (live) {
        const double *p = r2 + 3 * (jj0 + i2);
        p2x = p[0];
        p2y = p[1];
        p2z = p[2];
    }
    double f2r[NMAX > 0 ? NMAX : 1];
    if constexpr (NMAX > 0) {
#pragma unroll
        for (int A = 0; A < NMAX; A++) {
            f2r[A] = 0.0;
            if (A < n_nuc)
                f2r[A] = cutc_rbar(p2x - rn[3 * A], p2y - rn[3 * A + 1], p2z - rn[3 * A + 2]);
        }
    }
    const int base = 4 * (kmax + 1);
    for (int64_t i1_0 = (int64_t)blockIdx.y * R1B; i1_0 < n1; i1_0 += (int64_t)gridDim.y * R1B) {
        const int nr = (n1 - i1_0 < R1B) ? (int)(n1 - i1_0) : R1B;
        for (int r = 0; r < nr; r++) {
            const int64_t i1 = i1_0 + r;
            double ux = 0.0, uy = 0.0, uz = 0.0;
            if (live) {
                const double *p = r1 + 3 * (ii0 + i1);
                double g, gx, gy, gz;
                cutc_rbar_grad(p[0] - p2x, p[1] - p2y, p[2] - p2z, g, gx, gy, gz);
                const double *row = tab + i1 * ts;
                double s = 0.0, gp = 1.0;
                for (int k = 0; k <= kmax; k++) {
                    ux += gp * row[4 * k];
                    uy += gp * row[4 * k + 1];
                    uz += gp * row[4 * k + 2];
                    s += gp * row[4 * k + 3];
                    gp *= g;
                }
                for (int c = 0; c < n_cls; c++) {
                    const int j = cls[2 * c], k = cls[2 * c + 1];
                    const double gk = ipow_u(g, k);
                    const double *q = row + base + 4 * (int64_t)c * n_nuc;
                    if constexpr (NMAX > 0) {
#pragma unroll
                        for (int A = 0; A < NMAX; A++) {
                            if (A < n_nuc) {
                                const double w = ipow_u(f2r[A], j) * gk;
                                ux += w * q[4 * A];
                                uy += w * q[4 * A + 1];
                                uz += w * q[4 * A + 2];
                                s += w * q[4 * A + 3];
                            }
                        }
                    } else {
                        for (int A = 0; A < n_nuc; A++) {
                            const double w = ipow_u(f2tab[(int64_t)A * ld_f2 + i2], j) * gk;
                            ux += w * q[4 * A];
                            uy += w * q[4 * A + 1];
                            uz += w * q[4 * A + 2];
                            s += w * q[4 * A + 3];
                        }
                    }
                }
                ux += s * gx;
                uy += s * gy;
                uz += s * gz;
            }
            double *o = G + i1 * ldg + i2;
            o[0] = ux;
            o[plane] = uy;
            o[2 * plane] = uz;
            o[3 * plane] = -0.5 * (ux * ux + uy * uy + uz * uz);
        }
    }
}

// rbar(r12) = d / (1 + d) and its gradient (dx,dy,dz) / (d (1 + d)^2) with
// one division: q = 1 / (d (1 + d)), then 1 / (1 + d) = d q and
// 1 / (d (1 + d)^2) = d q^2; squared distances <= 1e-15 give zeros
// (reference: src/jast_bh.cu:139-156).
__device__ static inline void rbar_grad_1div(double dx, double dy, double dz, double &f, double &gx, double &gy,
                                             double &gz)
{
    const double d2 = dx * dx + dy * dy + dz * dz;
    if (d2 > 1e-15) {
        const double d = sqrt(d2);
        const double q = 1.0 / (d * (1.0 + d));
        const double t1 = d * q, t2 = t1 * q;
        f = d * t1;
        gx = t2 * dx;
        gy = t2 * dy;
        gz = t2 * dz;
    } else {
        f = gx = gy = gz = 0.0;
    }
}

// The pair kernel with the block's r1 table rows staged in LDS and NMAX the
// nucleus count rounded up to a multiple of 4.  The table entries are read
// as LDS broadcasts, which bound the kernel, so every thread takes two r2
// points (i2 and i2 + 256 of a 512-point block) and each entry read serves
// both.  ONE: the program has a single class (j, k), whose f2A^j is formed
// once per r2 point (with k = 0 that power is the weight itself).  A block
// stages JAST_LDS_R1B rows at a time and walks JAST_ROWS_PER_BLOCK rows, so
// that an r2 point's nucleus factors are formed once per that many pairs.
// The sums are those of jast_pair_kernel, in the same order.
#define JAST_LDS_R1B 16
#define JAST_ROWS_PER_BLOCK 128
#define JAST_PPT 2 /* r2 points per thread */
//
// part (optional): the block's sums of squares of every G row it writes
// (row m n1 + i1 of component m), over its 512 points, at
// part[blockIdx.x 4 n1 + m n1 + i1], for the row exponents of the modular
// quantisers (rowexp_finish_kernel sums them over the blocks in order).
// Fixed-order reductions: the wave sum of the device library, then the 4
// waves in order.
template <int NMAX, bool ONE>
__global__ __launch_bounds__(JAST_THREADS) void jast_pair_lds_kernel(
    const double *__restrict__ r1, const double *__restrict__ r2, const double *__restrict__ rn, int n_nuc,
    const double *__restrict__ tab, int64_t ts, const int *__restrict__ cls, int n_cls, int kmax,
    double *__restrict__ G, int64_t ii0, int64_t n1, int64_t jj0, int64_t n2, int64_t n2_fill, int64_t ldg,
    int64_t plane, double *__restrict__ part)
{
    extern __shared__ double s_tab[]; // JAST_LDS_R1B rows x ts
    __shared__ double s_part[JAST_LDS_R1B][4][JAST_THREADS / 64]; // [row][component][wave]
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    // the previous batch's row sums: the 4 waves in order, one value per (row, component)
    auto flush = [&](int64_t row0, int nrows) {
        const int t = threadIdx.x;
        if (part && t < 4 * nrows) {
            const int r = t >> 2, m = t & 3;
            const double v = ((s_part[r][m][0] + s_part[r][m][1]) + s_part[r][m][2]) + s_part[r][m][3];
            part[(int64_t)blockIdx.x * 4 * n1 + (int64_t)m * n1 + row0 + r] = v;
        }
    };
    int64_t prev_row0 = 0;
    int prev_nr = 0;
    const int j1 = ONE ? cls[0] : 0, k1 = ONE ? cls[1] : 0;
    int64_t i2v[JAST_PPT];
    bool act[JAST_PPT], live[JAST_PPT];
    double px[JAST_PPT], py[JAST_PPT], pz[JAST_PPT];
    double fj[JAST_PPT][NMAX]; // f2A, or f2A^j for a single class
#pragma unroll
    for (int h = 0; h < JAST_PPT; h++) {
        i2v[h] = (int64_t)blockIdx.x * JAST_PPT * JAST_THREADS + h * JAST_THREADS + threadIdx.x;
        act[h] = i2v[h] < n2_fill;
        live[h] = i2v[h] < n2;
        px[h] = py[h] = pz[h] = 0.0;
        if (live[h]) {
            const double *p = r2 + 3 * (jj0 + i2v[h]);
            px[h] = p[0];
            py[h] = p[1];
            pz[h] = p[2];
        }
#pragma unroll
        for (int A = 0; A < NMAX; A++) {
            fj[h][A] = 0.0;
            if (A < n_nuc) {
                const double f = cutc_rbar(px[h] - rn[3 * A], py[h] - rn[3 * A + 1], pz[h] - rn[3 * A + 2]);
                fj[h][A] = ONE ? ipow_u(f, j1) : f;
            }
        }
    }
    const int base = 4 * (kmax + 1);
    for (int64_t i1_0 = (int64_t)blockIdx.y * JAST_LDS_R1B; i1_0 < n1; i1_0 += (int64_t)gridDim.y * JAST_LDS_R1B) {
        const int nr = (n1 - i1_0 < JAST_LDS_R1B) ? (int)(n1 - i1_0) : JAST_LDS_R1B;
        __syncthreads(); // the previous rows are no longer read
        flush(prev_row0, prev_nr);
        for (int64_t t = threadIdx.x; t < nr * ts; t += JAST_THREADS)
            s_tab[t] = tab[i1_0 * ts + t];
        __syncthreads();
        for (int r = 0; r < nr; r++) {
            const double *p = r1 + 3 * (ii0 + i1_0 + r);
            const double *row = s_tab + r * ts;
            double ux[JAST_PPT], uy[JAST_PPT], uz[JAST_PPT], s[JAST_PPT], g[JAST_PPT], gx[JAST_PPT], gy[JAST_PPT],
                gz[JAST_PPT];
#pragma unroll
            for (int h = 0; h < JAST_PPT; h++) {
                rbar_grad_1div(p[0] - px[h], p[1] - py[h], p[2] - pz[h], g[h], gx[h], gy[h], gz[h]);
                ux[h] = uy[h] = uz[h] = s[h] = 0.0;
                double gp = 1.0;
                for (int k = 0; k <= kmax; k++) {
                    ux[h] += gp * row[4 * k];
                    uy[h] += gp * row[4 * k + 1];
                    uz[h] += gp * row[4 * k + 2];
                    s[h] += gp * row[4 * k + 3];
                    gp *= g[h];
                }
            }
            if constexpr (ONE) {
                double gk[JAST_PPT];
#pragma unroll
                for (int h = 0; h < JAST_PPT; h++)
                    gk[h] = ipow_u(g[h], k1);
#pragma unroll
                for (int A = 0; A < NMAX; A++) {
                    if (A < n_nuc) {
                        const double *qa = row + base + 4 * A;
                        const double q0 = qa[0], q1 = qa[1], q2 = qa[2], q3 = qa[3];
#pragma unroll
                        for (int h = 0; h < JAST_PPT; h++) {
                            const double w = k1 == 0 ? fj[h][A] : fj[h][A] * gk[h];
                            ux[h] += w * q0;
                            uy[h] += w * q1;
                            uz[h] += w * q2;
                            s[h] += w * q3;
                        }
                    }
                }
            } else {
                for (int c = 0; c < n_cls; c++) {
                    const int j = cls[2 * c], k = cls[2 * c + 1];
                    double gk[JAST_PPT];
#pragma unroll
                    for (int h = 0; h < JAST_PPT; h++)
                        gk[h] = ipow_u(g[h], k);
                    const double *qc = row + base + 4 * (int64_t)c * n_nuc;
#pragma unroll
                    for (int A = 0; A < NMAX; A++) {
                        if (A < n_nuc) {
                            const double q0 = qc[4 * A], q1 = qc[4 * A + 1], q2 = qc[4 * A + 2], q3 = qc[4 * A + 3];
#pragma unroll
                            for (int h = 0; h < JAST_PPT; h++) {
                                const double w = ipow_u(fj[h][A], j) * gk[h];
                                ux[h] += w * q0;
                                uy[h] += w * q1;
                                uz[h] += w * q2;
                                s[h] += w * q3;
                            }
                        }
                    }
                }
            }
            double sq[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int h = 0; h < JAST_PPT; h++) {
                if (!act[h])
                    continue;
                double vx = ux[h] + s[h] * gx[h], vy = uy[h] + s[h] * gy[h], vz = uz[h] + s[h] * gz[h];
                if (!live[h])
                    vx = vy = vz = 0.0;
                const double v3 = -0.5 * (vx * vx + vy * vy + vz * vz);
                double *o = G + (i1_0 + r) * ldg + i2v[h];
                o[0] = vx;
                o[plane] = vy;
                o[2 * plane] = vz;
                o[3 * plane] = v3;
                sq[0] = fma(vx, vx, sq[0]);
                sq[1] = fma(vy, vy, sq[1]);
                sq[2] = fma(vz, vz, sq[2]);
                sq[3] = fma(v3, v3, sq[3]);
            }
            if (part) {
#pragma unroll
                for (int m = 0; m < 4; m++) {
                    const double v = __ockl_wfred_add_f64(sq[m]);
                    if (lane == 0)
                        s_part[r][m][wave] = v;
                }
            }
        }
        prev_row0 = i1_0;
        prev_nr = nr;
    }
    __syncthreads();
    flush(prev_row0, prev_nr);
}

template <int NMAX, bool ONE>
static int launch_pairs_lds(const double *r1, const double *r2, const double *rn, int n_nuc, const double *tab,
                            int64_t ts, const int *cls, int n_cls, int kmax, double *G, int64_t ii0, int64_t n1,
                            int64_t jj0, int64_t n2, int64_t n2_fill, int64_t ldg, int64_t plane, double *part,
                            hipStream_t st)
{
    int64_t gx = (n2_fill + JAST_PPT * JAST_THREADS - 1) / (JAST_PPT * JAST_THREADS);
    int64_t gy = (n1 + JAST_ROWS_PER_BLOCK - 1) / JAST_ROWS_PER_BLOCK;
    if (gx > 2147483647)
        return CUTC_ERR_ARG;
    if (gy > 65535)
        gy = 65535;
    const size_t lds = sizeof(double) * JAST_LDS_R1B * (size_t)ts;
    hipLaunchKernelGGL((jast_pair_lds_kernel<NMAX, ONE>), dim3((unsigned)gx, (unsigned)gy),
                       dim3(JAST_THREADS), lds, st, r1, r2, rn, n_nuc, tab, ts, cls, n_cls, kmax, G, ii0, n1, jj0, n2,
                       n2_fill, ldg, plane, part);
    return hipGetLastError() == hipSuccess ? 0 : CUTC_ERR_HIP;
}

template <int NMAX, int R1B>
static int launch_pairs(const double *r1, const double *r2, const double *rn, int n_nuc, const double *tab,
                        int64_t ts, const int *cls, int n_cls, int kmax, const double *f2tab, int64_t ld_f2,
                        double *G, int64_t ii0, int64_t n1, int64_t jj0, int64_t n2, int64_t n2_fill, int64_t ldg,
                        int64_t plane, hipStream_t st)
{
    int64_t gx = (n2_fill + JAST_THREADS - 1) / JAST_THREADS;
    int64_t gy = (n1 + R1B - 1) / R1B;
    if (gx > 2147483647)
        return CUTC_ERR_ARG;
    if (gy > 65535)
        gy = 65535;
    hipLaunchKernelGGL((jast_pair_kernel<NMAX, R1B>), dim3((unsigned)gx, (unsigned)gy), dim3(JAST_THREADS), 0, st,
                       r1, r2, rn, n_nuc, tab, ts, cls, n_cls, kmax, f2tab, ld_f2, G, ii0, n1, jj0, n2, n2_fill, ldg,
                       plane);
    return hipGetLastError() == hipSuccess ? 0 : CUTC_ERR_HIP;
}

// x blocks (512 points each) of the staged kernel over n2_fill points: the
// number of partial row sums per G row it writes
extern "C" int64_t cutc_jast_norm_blocks(int64_t n2_fill)
{
    return (n2_fill + JAST_PPT * JAST_THREADS - 1) / (JAST_PPT * JAST_THREADS);
}

static int jast_staged(int n_nuc, int64_t ts) { return n_nuc <= 48 && sizeof(double) * JAST_LDS_R1B * ts <= 48 * 1024; }

static int jast_pairs(const double *r1, const double *r2, const double *rn, int n_nuc, const double *tab, int64_t ts,
                      const int *cls, int n_cls, int kmax, int64_t ii0, int64_t n1, int64_t jj0, int64_t n2,
                      int64_t n2_fill, double *G, int64_t ldg, int64_t plane, double *f2tab, double *part,
                      hipStream_t st)
{
    if (n1 <= 0 || n2_fill <= 0)
        return 0;
    /* up to 48 nuclei and a table block that fits 48 KB of LDS: the staged
     * kernel for the nucleus count rounded up to 4 (40 nuclei: 233 VGPRs, no
     * scratch, 2 waves per SIMD; 36 ms -> 15 ms per 7273-point chunk of
     * 300000 r2 points against the unstaged kernel) */
    if (jast_staged(n_nuc, ts)) {
        const int one = n_cls == 1;
#define LL(NM)                                                                                                       \
    return one ? launch_pairs_lds<NM, true>(r1, r2, rn, n_nuc, tab, ts, cls, n_cls, kmax, G, ii0, n1, jj0, n2,       \
                                            n2_fill, ldg, plane, part, st)                                            \
               : launch_pairs_lds<NM, false>(r1, r2, rn, n_nuc, tab, ts, cls, n_cls, kmax, G, ii0, n1, jj0, n2,      \
                                             n2_fill, ldg, plane, part, st)
        switch ((n_nuc + 3) / 4) {
        case 0:
        case 1: LL(4);
        case 2: LL(8);
        case 3: LL(12);
        case 4: LL(16);
        case 5: LL(20);
        case 6: LL(24);
        case 7: LL(28);
        case 8: LL(32);
        case 9: LL(36);
        case 10: LL(40);
        case 11: LL(44);
        default: LL(48);
        }
#undef LL
    }
#define LP(NM, RB)                                                                                                   \
    launch_pairs<NM, RB>(r1, r2, rn, n_nuc, tab, ts, cls, n_cls, kmax, f2tab, n2_fill, G, ii0, n1, jj0, n2, n2_fill,  \
                         ldg, plane, st)
    if (n_nuc <= 8)
        return LP(8, 32);
    if (n_nuc <= 16)
        return LP(16, 32);
    if (n_nuc <= 32)
        return LP(32, 32);
#undef LP
    if (!f2tab)
        return CUTC_ERR_ARG;
    if (n2 > 0) {
        int64_t nb = (n2 + 255) / 256;
        hipLaunchKernelGGL(jast_f2_kernel, dim3((unsigned)nb), dim3(256), 0, st, r2, jj0, n2, rn, n_nuc, f2tab,
                           n2_fill);
    }
    return launch_pairs<0, 32>(r1, r2, rn, n_nuc, tab, ts, cls, n_cls, kmax, f2tab, n2_fill, G, ii0, n1, jj0, n2,
                               n2_fill, ldg, plane, st);
}

extern "C" int cutc_launch_jast_pairs(const double *r1, const double *r2, const double *rn, int n_nuc,
                                      const double *tab, int64_t ts, const int *cls, int n_cls, int kmax, int64_t ii0,
                                      int64_t n1, int64_t jj0, int64_t n2, int64_t n2_fill, double *G, int64_t ldg,
                                      int64_t plane, double *f2tab, hipStream_t st)
{
    return jast_pairs(r1, r2, rn, n_nuc, tab, ts, cls, n_cls, kmax, ii0, n1, jj0, n2, n2_fill, G, ldg, plane, f2tab,
                      nullptr, st);
}

extern "C" int cutc_launch_jast_pairs_norms(const double *r1, const double *r2, const double *rn, int n_nuc,
                                            const double *tab, int64_t ts, const int *cls, int n_cls, int kmax,
                                            int64_t ii0, int64_t n1, int64_t n2, int64_t n2_fill, double *G,
                                            int64_t ldg, int64_t plane, double *part, hipStream_t st)
{
    if (!jast_staged(n_nuc, ts) || !part)
        return CUTC_ERR_UNSUPPORTED;
    return jast_pairs(r1, r2, rn, n_nuc, tab, ts, cls, n_cls, kmax, ii0, n1, 0, n2, n2_fill, G, ldg, plane, nullptr,
                      part, st);
}
