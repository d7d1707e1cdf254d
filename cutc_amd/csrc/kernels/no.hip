// Normal ordering (cutc_no) for gfx950: the per-grid-point intermediates of
// src/cutc_no.c:169-406 and the assembly steps of :419-694 that are not
// GEMMs.  The GEMMs themselves run on gemm_f64.hip (cutc_gemm_tn_f64).
//
// Inputs (device, the reference's Fortran layouts): w[g], mol[g + e ng],
// mor[g + e ng], G[g + x ng + i 3ng + j 3ng nm] (int2_grad1_u12 in the MO
// basis).  Orbitals e < nb are doubly occupied, nb <= e < na singly; with
// the pair weight W(e, f) = 1 if both are doubly occupied and 1/2 otherwise
// and the orbital weight w1(e) = 1 or 1/2, the reference's closed-shell
// (na == nb) and open-shell kernels are the same sums:
//   J_x = sum_e w1 G_x(e,e)            O = sum_e w1 L_e R_e
//   M_x = sum_ef W L_e R_f G_x(f,e)    S = 2 |J|^2 - sum_ef W sum_x G_x(f,e) G_x(e,f)
//   A_x(p) = w sum_e w1 L_e G_x(p,e)   Rt_x(p) = sum_e w1 R_e G_x(e,p)
//   C2(p,s) = sum_ex w1 G_x(p,e) G_x(e,s)
//   E(p) = [w L_p, -2 A(p).J + w sum_ex G_x(p,e) TE_x(e), A(p)],
//          TE_x(e) = sum_f W(e,f) L_f G_x(e,f)
//   F(p) = [-2 Rt(p).J + R_p S + sum_fx G_x(f,p) TF_x(f), R_p, Rt(p)],
//          TF_x(f) = sum_e W(e,f) R_e G_x(e,f)
//   C1_x(p,s) = R_s A_x(p) + L_p w Rt_x(s) - w O G_x(p,s) - 2 w L_p R_s J_x
//   D2(p,s) = w L_p R_s,  Dv = [w (2 O J - M), -w O]
// (src/no_tmp{J,O,M,S,C2}_os_kernel.cu, no_2e_tmp{A,B}_os_kernel.cu,
// no_1e_tmp{L,R,E,F}_os_kernel.cu, no_2e_tmpC1_kernel.cu:24-70,
// no_2e_tmpD2_kernel.cu:18-36, no_1e_tmpD_kernel.cu:18-35; the sums over
// e, f are regrouped as above, the same terms with the same weights).
//
// GEMM operands are written K-contiguous per row, padded with zeros to ldk
// (a multiple of 16) points of the chunk: C1x[x][p + s nm][k], Gx[x][..][k],
// C2[p + s nm][k], D2[...][k], E[p][c][k] (5 rows), F[...], Gt[p][x][k],
// Ht[...] (open shell).
#include "common.h"
#include "launch.h"

#define NO_TILE 8 // orbitals e held in registers per pass of the pair kernel

// (x, i, j) of int2_grad1_u12 at point g of the whole grid
#define GIJ(x, i, j) a.G[g + (x) * a.ng + (int64_t)(i) * n1 + (int64_t)(j) * n2]

__device__ static inline double w1(int e, int nb) { return e < nb ? 1.0 : 0.5; }
__device__ static inline double W2(int e, int f, int nb) { return (e < nb && f < nb) ? 1.0 : 0.5; }

// Per point: J, O, M, S, TE, TF, Dv and the point's share of no_0e
// (sum_{p < nb} A(p).Rt(p) + w (O S - 2 J.M), and for the open shell
// sum_f X(f).Y(f)), reduced per block in a fixed order into part[0][blk],
// part[1][blk].  One thread per point of the chunk.
__global__ __launch_bounds__(256) void no_point_kernel(cutc_no_args a, double *part)
{
    const int64_t gl = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t n1 = 3 * a.ng, n2 = n1 * a.nm;
    const int na = a.na, nb = a.nb, ne = a.ne;
    double s0 = 0.0, s1 = 0.0;
    if (gl < a.ngc) {
        const int64_t g = a.g0 + gl;
        const double wg = a.w[g];
        double J[3] = {0.0, 0.0, 0.0}, M[3] = {0.0, 0.0, 0.0}, O = 0.0, S = 0.0;
        for (int e = 0; e < ne; e++) {
            const double c = w1(e, nb);
            for (int x = 0; x < 3; x++)
                J[x] += c * GIJ(x, e, e);
            O += c * a.mol[g + e * a.ng] * a.mor[g + e * a.ng];
        }
        for (int e = 0; e < ne; e++) {
            const double Le = a.mol[g + e * a.ng];
            double te[3] = {0.0, 0.0, 0.0}, tf[3] = {0.0, 0.0, 0.0};
            for (int f = 0; f < ne; f++) {
                const double c = W2(e, f, nb), Lf = a.mol[g + f * a.ng], Rf = a.mor[g + f * a.ng];
                for (int x = 0; x < 3; x++) {
                    const double gfe = GIJ(x, f, e), gef = GIJ(x, e, f);
                    M[x] += c * Le * Rf * gfe;
                    S -= c * gfe * gef;
                    te[x] += c * Lf * gef;            // TE_x(e): sum_f W L_f G_x(e,f)
                    tf[x] += c * Rf * gfe;            // TF_x(e): sum_f W R_f G_x(f,e)
                }
            }
            for (int x = 0; x < 3; x++) {
                a.TE[(e * 3 + x) * a.ngc + gl] = te[x];
                a.TF[(e * 3 + x) * a.ngc + gl] = tf[x];
            }
        }
        S += 2.0 * (J[0] * J[0] + J[1] * J[1] + J[2] * J[2]);
        for (int x = 0; x < 3; x++) {
            a.J[x * a.ngc + gl] = J[x];
            a.Dv[x * a.ldk + gl] = wg * (2.0 * O * J[x] - M[x]);
        }
        a.Dv[3 * a.ldk + gl] = -wg * O;
        a.O[gl] = O;
        a.S[gl] = S;
        // 0e: the closed-shell DDOT runs over the nb first orbitals
        for (int p = 0; p < nb; p++)
            for (int x = 0; x < 3; x++) {
                double l = 0.0, r = 0.0;
                for (int e = 0; e < ne; e++) {
                    const double c = w1(e, nb);
                    l += c * a.mol[g + e * a.ng] * GIJ(x, p, e);
                    r += c * a.mor[g + e * a.ng] * GIJ(x, e, p);
                }
                s0 += wg * l * r;
            }
        s0 += wg * (O * S - 2.0 * (J[0] * M[0] + J[1] * M[1] + J[2] * M[2]));
        // open shell: X(f).Y(f) (src/no_0e_tmp{X,Y}_os_kernel.cu)
        if (na > nb) {
            for (int f = 0; f < na; f++)
                for (int x = 0; x < 3; x++) {
                    double xs = 0.0, ys = 0.0;
                    const int e0 = (f < nb) ? nb : 0;
                    const double cx = (f < nb) ? 0.25 : 0.5;
                    for (int e = e0; e < na; e++) {
                        xs += a.mol[g + e * a.ng] * GIJ(x, f, e);
                        ys += a.mor[g + e * a.ng] * GIJ(x, e, f);
                    }
                    s1 += cx * wg * xs * ys;
                }
        }
    } else if (gl < a.ldk) {
        for (int c = 0; c < 4; c++)
            a.Dv[c * a.ldk + gl] = 0.0;
    }
    // fixed-order block reduction of the two sums
    __shared__ double red[2][256];
    red[0][threadIdx.x] = s0;
    red[1][threadIdx.x] = s1;
    __syncthreads();
    for (unsigned h = 128; h > 0; h >>= 1) {
        if (threadIdx.x < h) {
            red[0][threadIdx.x] += red[0][threadIdx.x + h];
            red[1][threadIdx.x] += red[1][threadIdx.x + h];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[blockIdx.x] = red[0][0];
        part[a.nblk + blockIdx.x] = red[1][0];
    }
}

// Per (point, orbital p): A, Rt, the E and F rows (and Gt, Ht for the open
// shell).  Grid (ceil(ldk / 256), nm); threads past the chunk write the zero
// padding of the GEMM operands.
__global__ __launch_bounds__(256) void no_orb_kernel(cutc_no_args a)
{
    const int64_t gl = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int p = blockIdx.y;
    if (gl >= a.ldk)
        return;
    double *E = a.E + (int64_t)p * 5 * a.ldk + gl, *F = a.F + (int64_t)p * 5 * a.ldk + gl;
    double *Gt = a.Gt ? a.Gt + (int64_t)p * 3 * a.ldk + gl : nullptr;
    double *Ht = a.Ht ? a.Ht + (int64_t)p * 3 * a.ldk + gl : nullptr;
    if (gl >= a.ngc) {
        for (int c = 0; c < 5; c++)
            E[c * a.ldk] = F[c * a.ldk] = 0.0;
        if (Gt)
            for (int x = 0; x < 3; x++)
                Gt[x * a.ldk] = Ht[x * a.ldk] = 0.0;
        return;
    }
    const int64_t g = a.g0 + gl;
    const int64_t n1 = 3 * a.ng, n2 = n1 * a.nm;
    const int nb = a.nb, ne = a.ne;
    const double wg = a.w[g], Lp = a.mol[g + p * a.ng], Rp = a.mor[g + p * a.ng];
    double A[3] = {0.0, 0.0, 0.0}, Rt[3] = {0.0, 0.0, 0.0}, ga[3] = {0.0, 0.0, 0.0}, ha[3] = {0.0, 0.0, 0.0};
    double e1 = 0.0, f0 = 0.0;
    for (int e = 0; e < ne; e++) {
        const double c = w1(e, nb), Le = a.mol[g + e * a.ng], Re = a.mor[g + e * a.ng];
        for (int x = 0; x < 3; x++) {
            const double gpe = GIJ(x, p, e), gep = GIJ(x, e, p);
            A[x] += c * Le * gpe;
            Rt[x] += c * Re * gep;
            e1 += gpe * a.TE[(e * 3 + x) * a.ngc + gl];
            f0 += gep * a.TF[(e * 3 + x) * a.ngc + gl];
            if (e >= nb) {
                ga[x] += Le * gpe;
                ha[x] += Re * gep;
            }
        }
    }
    double J[3];
    for (int x = 0; x < 3; x++) {
        A[x] *= wg;
        J[x] = a.J[x * a.ngc + gl];
        a.A[((int64_t)p * 3 + x) * a.ngc + gl] = A[x];
        a.Rt[((int64_t)p * 3 + x) * a.ngc + gl] = Rt[x];
    }
    E[0] = wg * Lp;
    E[a.ldk] = -2.0 * (A[0] * J[0] + A[1] * J[1] + A[2] * J[2]) + wg * e1;
    F[0] = -2.0 * (Rt[0] * J[0] + Rt[1] * J[1] + Rt[2] * J[2]) + Rp * a.S[gl] + f0;
    F[a.ldk] = Rp;
    for (int x = 0; x < 3; x++) {
        E[(2 + x) * a.ldk] = A[x];
        F[(2 + x) * a.ldk] = Rt[x];
        if (Gt) {
            Gt[x * a.ldk] = 0.5 * wg * ga[x];
            Ht[x * a.ldk] = 0.5 * ha[x];
        }
    }
}

// one (point, orbital p) of no_pair_kernel
__device__ static void no_pair_point(const cutc_no_args &a, const int p, const int64_t gl)
{
    if (gl >= a.ldk)
        return;
    const int nm = a.nm, nb = a.nb, ne = a.ne;
    const int64_t nm2 = (int64_t)nm * nm;
    const int64_t L = a.ldk;
    if (gl >= a.ngc) { // zero padding
        for (int s = 0; s < nm; s++) {
            const int64_t col = p + (int64_t)s * nm;
            for (int x = 0; x < 3; x++) {
                a.C1x[(x * nm2 + col) * L + gl] = 0.0;
                a.Gx[(x * nm2 + col) * L + gl] = 0.0;
            }
            a.C2[col * L + gl] = 0.0;
            a.D2[col * L + gl] = 0.0;
        }
        return;
    }
    const int64_t g = a.g0 + gl;
    const int64_t n1 = 3 * a.ng, n2 = n1 * nm;
    const double wg = a.w[g], Lp = a.mol[g + p * a.ng], O = a.O[gl];
    double Ap[3], Jw[3];
    for (int x = 0; x < 3; x++) {
        Ap[x] = a.A[((int64_t)p * 3 + x) * a.ngc + gl];
        Jw[x] = 2.0 * wg * Lp * a.J[x * a.ngc + gl];
    }
    for (int e0 = 0; e0 < (ne > 0 ? ne : 1); e0 += NO_TILE) {
        const int e1 = (e0 + NO_TILE < ne) ? e0 + NO_TILE : ne;
        double u[NO_TILE][3];
#pragma unroll
        for (int t = 0; t < NO_TILE; t++)
#pragma unroll
            for (int x = 0; x < 3; x++)
                u[t][x] = (e0 + t < e1) ? w1(e0 + t, nb) * GIJ(x, p, e0 + t) : 0.0;
        for (int s = 0; s < nm; s++) {
            const int64_t col = p + (int64_t)s * nm;
            double c2 = (e0 == 0) ? 0.0 : a.C2[col * L + gl];
#pragma unroll
            for (int t = 0; t < NO_TILE; t++) {
                if (e0 + t < e1)
#pragma unroll
                    for (int x = 0; x < 3; x++)
                        c2 = fma(u[t][x], GIJ(x, e0 + t, s), c2);
            }
            a.C2[col * L + gl] = c2;
            if (e0 == 0) {
                const double Rs = a.mor[g + s * a.ng];
                a.D2[col * L + gl] = wg * Lp * Rs;
                for (int x = 0; x < 3; x++) {
                    const double gps = GIJ(x, p, s);
                    a.Gx[(x * nm2 + col) * L + gl] = gps;
                    a.C1x[(x * nm2 + col) * L + gl] = Rs * Ap[x] + Lp * wg * a.Rt[((int64_t)s * 3 + x) * a.ngc + gl] -
                                                      wg * O * gps - Rs * Jw[x];
                }
            }
        }
    }
}

// Per (point, orbital p), all s: C1x, the copy Gx of the chunk of G, C2 and
// D2 in the GEMM layouts.  The G_x(p, e) of NO_TILE orbitals e sit in
// registers while s runs; more occupied orbitals take further passes that
// add into C2.  Grid (nm, blocks of 256 points): the nm blocks of one point
// range are adjacent in launch order, so the occupied rows G_x(e, .) they all
// read come from L2 once the first has fetched them.
__global__ __launch_bounds__(256) void no_pair_kernel(cutc_no_args a)
{
    const int p = blockIdx.x;
    for (int64_t gb = blockIdx.y; gb * 256 < a.ldk; gb += gridDim.y)
        no_pair_point(a, p, gb * 256 + threadIdx.x);
}

// no_1e[b] += 2 (sum_x sum_k Gx[x][b][k] Dv[x][k] + sum_k C2[b][k] Dv[3][k])
// (the two DGEMV of src/cutc_no.c:481-506), one block per b, fixed order.
__global__ __launch_bounds__(256) void no_gemv_kernel(const double *__restrict__ Gx, const double *__restrict__ C2,
                                                      const double *__restrict__ Dv, int64_t nm2, int64_t L,
                                                      int64_t ngc, double *no_1e)
{
    const int64_t b = blockIdx.x;
    double s = 0.0;
    for (int x = 0; x < 3; x++) {
        const double *row = Gx + (x * nm2 + b) * L, *d = Dv + x * L;
        for (int64_t k = threadIdx.x; k < ngc; k += 256)
            s = fma(row[k], d[k], s);
    }
    {
        const double *row = C2 + b * L, *d = Dv + 3 * L;
        for (int64_t k = threadIdx.x; k < ngc; k += 256)
            s = fma(row[k], d[k], s);
    }
    __shared__ double red[256];
    red[threadIdx.x] = s;
    __syncthreads();
    for (unsigned h = 128; h > 0; h >>= 1) {
        if (threadIdx.x < h)
            red[threadIdx.x] += red[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        no_1e[b] += 2.0 * red[0];
}

// out[i] += sum_y work[y][i], y = 0 .. ns - 1 in order (split-K partials)
__global__ __launch_bounds__(256) void no_add_slices_kernel(const double *__restrict__ work, int ns, int64_t n,
                                                            double *out)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n)
        return;
    double s = 0.0;
    for (int y = 0; y < ns; y++)
        s += work[(int64_t)y * n + i];
    out[i] += s;
}

// no_2e(p, q, s, t) = N[(p, s), (q, t)] + N[(q, t), (p, s)] with N an
// nm^2 x nm^2 column-major matrix, pair (p, s) = p + s nm: the reference's
// N + N^T (trans_inplace) followed by its (q, s) swap
// (trans_pqst_psqt_inplace.cu:4-31), out of place.  A block takes a 32 x 32
// tile of (p, q) for one (s, t); the transposed term goes through LDS.
__global__ __launch_bounds__(256) void no_sym_swap_kernel(const double *__restrict__ N, int nm, double *out)
{
    __shared__ double tile[32][33];
    const int64_t tiles = (nm + 31) / 32;
    const int64_t bid = blockIdx.x, st = bid / (tiles * tiles), tt = bid - st * tiles * tiles;
    const int64_t tp = tt % tiles, tq = tt / tiles;
    const int64_t s = st % nm, t = st / nm;
    const int64_t nm2 = (int64_t)nm * nm;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    // second term, read with q fastest: N[(q + t nm) + (p + s nm) nm2]
    for (int r = ty; r < 32; r += 8) {
        const int64_t q = (int64_t)tq * 32 + tx, p = (int64_t)tp * 32 + r;
        tile[r][tx] = (p < nm && q < nm) ? N[(q + t * nm) + (p + s * nm) * nm2] : 0.0;
    }
    __syncthreads();
    for (int r = ty; r < 32; r += 8) {
        const int64_t p = (int64_t)tp * 32 + tx, q = (int64_t)tq * 32 + r;
        if (p < nm && q < nm)
            out[p + q * nm + s * nm2 + t * nm2 * nm] = N[(p + s * nm) + (q + t * nm) * nm2] + tile[tx][r];
    }
}

// no_0e = -2 (sum part[0][.]) - 2 (sum part[1][.]), fixed order
__global__ void no_0e_kernel(const double *__restrict__ part, int64_t nblk, double *no_0e)
{
    if (threadIdx.x != 0 || blockIdx.x != 0)
        return;
    double s0 = 0.0, s1 = 0.0;
    for (int64_t b = 0; b < nblk; b++) {
        s0 += part[b];
        s1 += part[nblk + b];
    }
    no_0e[0] = -2.0 * s0 - 2.0 * s1;
}

extern "C" int cutc_no_point(const cutc_no_args *a, double *part, hipStream_t st)
{
    if (a->ngc <= 0)
        return 0;
    hipLaunchKernelGGL(no_point_kernel, dim3((unsigned)((a->ldk + 255) / 256)), dim3(256), 0, st, *a, part);
    return hipGetLastError() == hipSuccess ? 0 : CUTC_ERR_HIP;
}

extern "C" int cutc_no_orb(const cutc_no_args *a, hipStream_t st)
{
    if (a->ngc <= 0 || a->nm <= 0)
        return 0;
    hipLaunchKernelGGL(no_orb_kernel, dim3((unsigned)((a->ldk + 255) / 256), (unsigned)a->nm), dim3(256), 0, st, *a);
    return hipGetLastError() == hipSuccess ? 0 : CUTC_ERR_HIP;
}

extern "C" int cutc_no_pair(const cutc_no_args *a, hipStream_t st)
{
    if (a->ngc <= 0 || a->nm <= 0)
        return 0;
    const int64_t nb = (a->ldk + 255) / 256;
    hipLaunchKernelGGL(no_pair_kernel, dim3((unsigned)a->nm, (unsigned)(nb < 65535 ? nb : 65535)), dim3(256), 0, st,
                       *a);
    return hipGetLastError() == hipSuccess ? 0 : CUTC_ERR_HIP;
}

extern "C" int cutc_no_gemv(const double *Gx, const double *C2, const double *Dv, int nm, int64_t ldk, int64_t ngc,
                            double *no_1e, hipStream_t st)
{
    const int64_t nm2 = (int64_t)nm * nm;
    if (nm2 <= 0 || ngc <= 0)
        return 0;
    if (nm2 > 2147483647)
        return CUTC_ERR_ARG;
    hipLaunchKernelGGL(no_gemv_kernel, dim3((unsigned)nm2), dim3(256), 0, st, Gx, C2, Dv, nm2, ldk, ngc, no_1e);
    return hipGetLastError() == hipSuccess ? 0 : CUTC_ERR_HIP;
}

extern "C" int cutc_no_add_slices(const double *work, int ns, int64_t n, double *out, hipStream_t st)
{
    if (n <= 0 || ns <= 0)
        return 0;
    hipLaunchKernelGGL(no_add_slices_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, work, ns, n, out);
    return hipGetLastError() == hipSuccess ? 0 : CUTC_ERR_HIP;
}

extern "C" int cutc_no_sym_swap(const double *N, int nm, double *out, hipStream_t st)
{
    if (nm <= 0)
        return 0;
    const int64_t tiles = (nm + 31) / 32, nb = tiles * tiles * nm * nm;
    if (nb > 2147483647)
        return CUTC_ERR_ARG;
    hipLaunchKernelGGL(no_sym_swap_kernel, dim3((unsigned)nb), dim3(256), 0, st, N, nm, out);
    return hipGetLastError() == hipSuccess ? 0 : CUTC_ERR_HIP;
}

extern "C" int cutc_no_finish_0e(const double *part, int64_t nblk, double *no_0e, hipStream_t st)
{
    hipLaunchKernelGGL(no_0e_kernel, dim3(1), dim3(64), 0, st, part, nblk, no_0e);
    return hipGetLastError() == hipSuccess ? 0 : CUTC_ERR_HIP;
}
