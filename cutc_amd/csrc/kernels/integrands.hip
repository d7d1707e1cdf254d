// Integrand kernels for gfx950 (bandwidth-bound, FP64).
//
// Reference integrands (same values and multiplication order):
//   long range  A[r2,(i,j)] = (w2 chi_i) chi_j      src/long_range_integ.cu:17-46
//   herm        H[r1,(i,j)] = (w1 chi_i) chi_j      src/short_range_integ_herm.cu:16-45
//   non-herm    NH[(r1,m),(i,j)] = w1 (chi_j d_m chi_i - chi_i d_m chi_j)
//                                                   src/short_range_integ_nonherm.cu:19-63
// with column index j + n_ao i.  The contractions only need the symmetric
// (k <= l) half of each, so the product path writes packed pair layouts:
//   Apk[p][r2]   K-contiguous rows, p = (k, l) packed, zero-padded to n2_fill
//   Hp[r1][q]    q = (a, b) packed, a <= b: (w1 chi_b) chi_a
//   NHp[m][r1][q]  w1 (chi_a d_m chi_b - chi_b d_m chi_a)
// The *_ref launchers write the reference's full layouts for the exported
// wrapper symbols.
#include "common.h"
#include "launch.h"

// Products are rounded one by one (no FMA contraction), as in the reference's
// C expressions, so NH is exactly antisymmetric with an exactly zero diagonal.
#pragma clang fp contract(off)

__global__ void pair_table_kernel(int n_ao, int *pairs)
{
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    int l = blockIdx.y;
    if (k > l || k >= n_ao)
        return;
    int64_t p = cutc_pair_index(k, l);
    pairs[2 * p] = k;
    pairs[2 * p + 1] = l;
}

extern "C" int cutc_launch_pair_table(int n_ao, int *pairs, hipStream_t st)
{
    if (n_ao <= 0)
        return 0;
    hipLaunchKernelGGL(pair_table_kernel, dim3((n_ao + 63) / 64, n_ao), dim3(64), 0, st, n_ao, pairs);
    return hipGetLastError() == hipSuccess ? 0 : CUTC_ERR_HIP;
}

// One block row per l; lanes along r2.  Apk[p(k,l) * lda + i2] = (w chi_l) chi_k.
__global__ __launch_bounds__(256) void pair_weights_kernel(const double *__restrict__ wr2,
                                                           const double *__restrict__ aos2, int64_t n2_ld,
                                                           int64_t jj0, int64_t n2, int64_t n2_fill, double *Apk,
                                                           int64_t lda)
{
    const int64_t i2 = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int l = blockIdx.y;
    if (i2 >= n2_fill)
        return;
    const int64_t p0 = cutc_pair_index(0, l);
    if (i2 >= n2) {
        for (int k = 0; k <= l; k++)
            Apk[(p0 + k) * lda + i2] = 0.0;
        return;
    }
    const int64_t c = jj0 + i2;
    const double wl = wr2[c] * aos2[c + n2_ld * l];
    for (int k = 0; k <= l; k++)
        Apk[(p0 + k) * lda + i2] = wl * aos2[c + n2_ld * k];
}

extern "C" int cutc_launch_pair_weights(const double *wr2, const double *aos2, int64_t n2_ld, int64_t jj0,
                                        int64_t n2, int64_t n2_fill, int n_ao, double *Apk, int64_t lda,
                                        hipStream_t st)
{
    if (n2_fill <= 0 || n_ao <= 0)
        return 0;
    int64_t gx = (n2_fill + 255) / 256;
    if (gx > 2147483647 || n_ao > 65535)
        return CUTC_ERR_ARG;
    hipLaunchKernelGGL(pair_weights_kernel, dim3((unsigned)gx, n_ao), dim3(256), 0, st, wr2, aos2, n2_ld, jj0, n2,
                       n2_fill, Apk, lda);
    return hipGetLastError() == hipSuccess ? 0 : CUTC_ERR_HIP;
}

// R r1 points per block: their 4 AO components are staged in LDS as
// s[(c * n_ao + ao) * R + r]; the lanes then sweep the packed pairs q.
template <int R>
__global__ __launch_bounds__(256) void r1_integrands_kernel(const double *__restrict__ wr1,
                                                            const double *__restrict__ aos1, int64_t n1_ld,
                                                            int64_t ii0, int64_t n1, int n_ao,
                                                            const int *__restrict__ pairs, double *Hp, double *NHp,
                                                            int64_t ldq, int64_t nh_plane,
                                                            const uint32_t *__restrict__ perm)
{
    extern __shared__ double s[];
    __shared__ double sw[R];
    const int64_t r0 = (int64_t)blockIdx.x * R;
    const int nr = (n1 - r0 < R) ? (int)(n1 - r0) : R;
    for (int t = threadIdx.x; t < 4 * n_ao * R; t += 256) {
        int r = t % R, ca = t / R; // ca = c * n_ao + ao
        int ao = ca % n_ao, c = ca / n_ao;
        const int64_t pt = ii0 + (perm && r < nr ? (int64_t)perm[r0 + r] : r0 + r);
        s[t] = (r < nr) ? aos1[pt + n1_ld * (ao + (int64_t)n_ao * c)] : 0.0;
    }
    if (threadIdx.x < R)
        sw[threadIdx.x] =
            ((int)threadIdx.x < nr) ? wr1[ii0 + (perm ? (int64_t)perm[r0 + threadIdx.x] : r0 + threadIdx.x)] : 0.0;
    __syncthreads();
    const int64_t Q = (int64_t)n_ao * (n_ao + 1) / 2;
    const int64_t cs = (int64_t)n_ao * R;
    for (int64_t q = threadIdx.x; q < Q; q += 256) {
        const int a = pairs[2 * q], b = pairs[2 * q + 1];
        for (int r = 0; r < nr; r++) {
            const double w = sw[r];
            const double va = s[a * R + r], vb = s[b * R + r];
            const int64_t row = r0 + r;
            Hp[row * ldq + q] = w * vb * va;
            for (int m = 0; m < 3 && NHp; m++) {
                const double da = s[(m + 1) * cs + a * R + r], db = s[(m + 1) * cs + b * R + r];
                NHp[m * nh_plane + row * ldq + q] = w * (va * db - vb * da);
            }
        }
    }
}

extern "C" int cutc_launch_r1_integrands(const double *wr1, const double *aos1, int64_t n1_ld, int64_t ii0,
                                         int64_t n1, int n_ao, const int *pairs, double *Hp, double *NHp,
                                         int64_t ldq, int64_t nh_plane, const uint32_t *perm, hipStream_t st)
{
    if (n1 <= 0 || n_ao <= 0)
        return 0;
    if ((size_t)4 * n_ao * 4 * sizeof(double) <= 64 * 1024) {
        int64_t nb = (n1 + 3) / 4;
        hipLaunchKernelGGL((r1_integrands_kernel<4>), dim3((unsigned)nb), dim3(256),
                           sizeof(double) * 4 * n_ao * 4, st, wr1, aos1, n1_ld, ii0, n1, n_ao, pairs, Hp, NHp, ldq,
                           nh_plane, perm);
    } else if ((size_t)4 * n_ao * sizeof(double) <= 64 * 1024) {
        hipLaunchKernelGGL((r1_integrands_kernel<1>), dim3((unsigned)n1), dim3(256), sizeof(double) * 4 * n_ao, st,
                           wr1, aos1, n1_ld, ii0, n1, n_ao, pairs, Hp, NHp, ldq, nh_plane, perm);
    } else {
        return CUTC_ERR_UNSUPPORTED;
    }
    return hipGetLastError() == hipSuccess ? 0 : CUTC_ERR_HIP;
}

// ---- reference layouts (exported wrapper symbols) ----

// out[i2 + n2_tot (j + n_ao i)] = (w2 chi_i) chi_j; blockIdx.y = i.
__global__ void long_range_ref_kernel(int64_t jj0, int64_t n2_eff, int64_t n2_tot, int64_t n2, int n_ao,
                                      const double *wr2, const double *aos2, double *out)
{
    const int64_t i2 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int i = blockIdx.y;
    if (i2 >= n2_eff)
        return;
    const int64_t c = jj0 + i2;
    const double wi = wr2[c] * aos2[c + n2 * i];
    for (int j = 0; j < n_ao; j++)
        out[i2 + n2_tot * (j + (int64_t)n_ao * i)] = wi * aos2[c + n2 * j];
}

extern "C" int cutc_launch_long_range_ref(int64_t jj0, int64_t n2_eff, int64_t n2_tot, int64_t n2, int n_ao,
                                          const double *wr2, const double *aos2, double *out, hipStream_t st)
{
    if (n2_eff <= 0 || n_ao <= 0)
        return 0;
    hipLaunchKernelGGL(long_range_ref_kernel, dim3((unsigned)((n2_eff + 255) / 256), n_ao), dim3(256), 0, st, jj0,
                       n2_eff, n2_tot, n2, n_ao, wr2, aos2, out);
    return hipGetLastError() == hipSuccess ? 0 : CUTC_ERR_HIP;
}

// out[i1 + n1 (j + n_ao i)] = (w1 chi_i) chi_j
__global__ void herm_ref_kernel(int64_t n1, int n_ao, const double *wr1, const double *aos1, double *out)
{
    const int64_t i1 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int i = blockIdx.y;
    if (i1 >= n1)
        return;
    const double wi = wr1[i1] * aos1[i1 + n1 * i];
    for (int j = 0; j < n_ao; j++)
        out[i1 + n1 * (j + (int64_t)n_ao * i)] = wi * aos1[i1 + n1 * j];
}

extern "C" int cutc_launch_herm_ref(int64_t n1, int n_ao, const double *wr1, const double *aos1, double *out,
                                    hipStream_t st)
{
    if (n1 <= 0 || n_ao <= 0)
        return 0;
    hipLaunchKernelGGL(herm_ref_kernel, dim3((unsigned)((n1 + 255) / 256), n_ao), dim3(256), 0, st, n1, n_ao, wr1,
                       aos1, out);
    return hipGetLastError() == hipSuccess ? 0 : CUTC_ERR_HIP;
}

// out[(i1 + m n1) + 3 n1 (j + n_ao i)] = w1 (chi_j d_m chi_i - chi_i d_m chi_j)
__global__ void nonherm_ref_kernel(int64_t n1, int n_ao, const double *wr1, const double *aos1, double *out)
{
    const int64_t i1 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int i = blockIdx.y;
    if (i1 >= n1)
        return;
    const int64_t comp = n1 * n_ao;
    const double w = wr1[i1];
    const double vi = aos1[i1 + n1 * i];
    for (int j = 0; j < n_ao; j++) {
        const double vj = aos1[i1 + n1 * j];
        for (int m = 0; m < 3; m++) {
            const double di = aos1[i1 + n1 * i + (m + 1) * comp];
            const double dj = aos1[i1 + n1 * j + (m + 1) * comp];
            out[(i1 + m * n1) + 3 * n1 * (j + (int64_t)n_ao * i)] = w * (vj * di - vi * dj);
        }
    }
}

extern "C" int cutc_launch_nonherm_ref(int64_t n1, int n_ao, const double *wr1, const double *aos1, double *out,
                                       hipStream_t st)
{
    if (n1 <= 0 || n_ao <= 0)
        return 0;
    hipLaunchKernelGGL(nonherm_ref_kernel, dim3((unsigned)((n1 + 255) / 256), n_ao), dim3(256), 0, st, n1, n_ao,
                       wr1, aos1, out);
    return hipGetLastError() == hipSuccess ? 0 : CUTC_ERR_HIP;
}

// dst[0..n) = value (a kernel rather than hipMemsetAsync, so that the call
// stays valid inside captured graphs).
__global__ void fill_kernel(double *dst, int64_t n, double value)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = value;
}

extern "C" int cutc_launch_fill(double *dst, int64_t n, double value, hipStream_t st)
{
    if (n <= 0)
        return 0;
    int64_t nb = (n + 255) / 256;
    if (nb > 65536)
        nb = 65536;
    hipLaunchKernelGGL(fill_kernel, dim3((unsigned)nb), dim3(256), 0, st, dst, n, value);
    return hipGetLastError() == hipSuccess ? 0 : CUTC_ERR_HIP;
}
