// Slice (Ozaki scheme I) split-precision contractions for gfx950: every
// operand row x, with max |x| < 2^e, is cut into S integer digits
//   x 2^(8 - e) = d_0 + d_1 2^-8 + ... + d_{S-1} 2^(-8(S-1)) + rest,
// |d_0| <= 256, |d_s| <= 128 (s > 0), each exact in bf16 and in fp16.  The
// product of two rows is sum_{s+t<S} 2^(-8(s+t)) sum_k d_s(a_k) d_t(b_k),
// scaled by 2^(ea + eb - 16) (the products with s + t >= S are dropped).
// Within a K block of 256 every level L_d = sum_{s+t=d} sum_k d_s d_t is an
// integer below 2^24 in magnitude (level 0: 256 * 2^16; level 1: 256 * 2 *
// 2^15; levels 2, 3: at most 1.5 * 2^23), so the FP32 accumulation of
// v_mfma_f32_32x32x16_{bf16,f16} is exact; after each block the levels are
// combined exactly in FP64 (they span at most 49 bits) and added into FP64
// accumulators.
//
// Kernels:
//   oz_slice_*      digits of the operand rows into S 16-bit planes in the
//                   slice-packed layout below
//   oz_gemm         per block 64 x 128, 4 waves of 32 x 64 (1 x 2 MFMA
//                   32x32x16), S level accumulators in FP32, FP64 sums; both
//                   operands through a 2-stage LDS ring fed by LDS DMA
//   oz_scatter_x, scatter_m   FP64 results scattered to the reference's
//                   layouts (scatter_m also serves the split-K FP64 path)
//
// Slice-packed planes (16-bit elements): rows in blocks of 32, K in steps of
// 32; the 2 KB tile (row block rb, step kt) sits at element (rb KT + kt) 1024
// (KT = kpad / 32) and holds element k = 32 kt + 16 c + 8 h + e of row
// 32 rb + r at element c 512 + (32 h + r) 8 + e: the 16 bytes lane 32 h + r
// feeds to the c-th MFMA of the step (A: row r, k = 8 h + e; the B operand of
// the same layout: column r).
#include "common.h"
#include "launch.h"
#include "split_common.h"

typedef int v4i __attribute__((ext_vector_type(4)));
typedef float v16f __attribute__((ext_vector_type(16)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef _Float16 v8h __attribute__((ext_vector_type(8)));
typedef __attribute__((address_space(3))) int8_t lds8;

#define OZ_MAXS 4

// element offset of (row 32 rb + r, k) in a slice plane
__device__ static inline int64_t oz_offset(int64_t rb, int r, int64_t k, int64_t KT)
{
    const int kk = (int)(k & 31);
    return (rb * KT + (k >> 5)) * 1024 + (kk >> 4) * 512 + (32 * ((kk >> 3) & 1) + r) * 8 + (kk & 7);
}

// 16-bit pattern of the integer-valued float d (|d| <= 256): exact in both
// formats, so the bf16 pattern is the upper half of the float's bits
template <bool F16> __device__ static inline uint16_t oz_bits(float d)
{
    if constexpr (F16) {
        _Float16 h = (_Float16)d;
        return __builtin_bit_cast(uint16_t, h);
    } else {
        return (uint16_t)(__float_as_uint(d) >> 16);
    }
}

// 8 consecutive values of one row -> S digit planes, 16 bytes each
template <int S, bool F16>
__device__ static inline void oz_store_digits(const double (&x)[8], int sh, uint16_t *dst, int64_t plane)
{
    uint32_t w[S][4];
#pragma unroll
    for (int e = 0; e < 8; e += 2) {
        double y0 = ldexp(x[e], sh), y1 = ldexp(x[e + 1], sh);
#pragma unroll
        for (int s = 0; s < S; s++) {
            const double r0 = rint(y0), r1 = rint(y1);
            w[s][e / 2] = (uint32_t)oz_bits<F16>((float)r0) | ((uint32_t)oz_bits<F16>((float)r1) << 16);
            y0 = (y0 - r0) * 256.0;
            y1 = (y1 - r1) * 256.0;
        }
    }
#pragma unroll
    for (int s = 0; s < S; s++)
        *(v4i *)(dst + s * plane) = v4i{(int)w[s][0], (int)w[s][1], (int)w[s][2], (int)w[s][3]};
}

// ---------------------------------------------------------------- slicing

// grid (kpad / 64, ceil(rows / 32)): thread r = tid % 32 takes row 32 rb + r,
// 8 values j = tid / 32 of a 64-wide K range
template <int S, bool F16>
__global__ __launch_bounds__(256) void oz_slice_k_kernel(const double *__restrict__ X, int64_t ld, int64_t rows,
                                                         int64_t K, int64_t KT, int64_t plane,
                                                         const int *__restrict__ rexp, uint16_t *planes)
{
    const int r = threadIdx.x & 31, j = threadIdx.x >> 5;
    const int64_t rb = blockIdx.y, row = rb * 32 + r;
    const int64_t k0 = (int64_t)blockIdx.x * 64 + 8 * j;
    double x[8];
    int sh = 0;
    if (row < rows) {
        sh = 8 - rexp[row];
        const double *src = X + row * ld;
#pragma unroll
        for (int e = 0; e < 8; e++)
            x[e] = (k0 + e < K) ? src[k0 + e] : 0.0;
    } else {
#pragma unroll
        for (int e = 0; e < 8; e++)
            x[e] = 0.0;
    }
    oz_store_digits<S, F16>(x, sh, planes + oz_offset(rb, r, k0, KT), plane);
}

template <int S, bool F16>
__global__ __launch_bounds__(256) void oz_slice_pairw_kernel(const double *__restrict__ wr2,
                                                             const double *__restrict__ aos2, int64_t n2, int n_ao,
                                                             int64_t KT, int64_t plane,
                                                             const int *__restrict__ rexp, uint16_t *planes)
{
    const int r = threadIdx.x & 31, j = threadIdx.x >> 5;
    const int64_t rb = blockIdx.y, p = rb * 32 + r;
    const int64_t k0 = (int64_t)blockIdx.x * 64 + 8 * j;
    const int64_t P = (int64_t)n_ao * (n_ao + 1) / 2;
    double x[8];
    int sh = 0;
    if (p < P) {
        int k, l;
        pair_of(p, k, l);
        sh = 8 - rexp[p];
#pragma unroll
        for (int e = 0; e < 8; e++) {
            const int64_t i = k0 + e;
            x[e] = (i < n2) ? wr2[i] * aos2[i + n2 * l] * aos2[i + n2 * k] : 0.0;
        }
    } else {
#pragma unroll
        for (int e = 0; e < 8; e++)
            x[e] = 0.0;
    }
    oz_store_digits<S, F16>(x, sh, planes + oz_offset(rb, r, k0, KT), plane);
}

// column-major sources: 32 rows x 64 K of one segment per block through LDS.
// grid (ceil(rows / 32), kp / 64, nsrc)
template <int S, bool F16>
__global__ __launch_bounds__(256) void oz_slice_colk_kernel(cutc_colsrc s, int64_t rows, int64_t kp, int64_t KT,
                                                            int64_t plane, const int *__restrict__ rexp,
                                                            uint16_t *planes)
{
    __shared__ double tile[64][33];
    const int64_t rb = blockIdx.x, r0 = rb * 32, k0 = (int64_t)blockIdx.y * 64;
    const int seg = blockIdx.z;
    {
        const int rr = threadIdx.x & 31;
        const int64_t r = r0 + rr;
        const bool rok = r < rows;
        const double *src = s.p[seg] + colsrc_col(s, rok ? r : 0);
        for (int kk = threadIdx.x >> 5; kk < 64; kk += 8) {
            const int64_t k = k0 + kk;
            tile[kk][rr] = (rok && k < s.K) ? src[colsrc_k(s, k) * s.ld] : 0.0;
        }
    }
    __syncthreads();
    const int r = threadIdx.x & 31, j = threadIdx.x >> 5;
    const int64_t row = r0 + r;
    const int sh = (row < rows) ? 8 - rexp[row] : 0;
    double x[8];
#pragma unroll
    for (int e = 0; e < 8; e++)
        x[e] = tile[8 * j + e][r];
    oz_store_digits<S, F16>(x, sh, planes + oz_offset(rb, r, seg * kp + k0 + 8 * j, KT), plane);
}

template <int S, bool F16>
static int slice_k(const double *X, int64_t ld, int64_t rows, int64_t K, int64_t kpad, int64_t plane_rows,
                   const int *rexp, uint16_t *planes, hipStream_t st)
{
    hipLaunchKernelGGL((oz_slice_k_kernel<S, F16>), dim3((unsigned)(kpad / 64), (unsigned)((rows + 31) / 32)),
                       dim3(256), 0, st, X, ld, rows, K, kpad / 32, plane_rows * kpad, rexp, planes);
    return hipGetLastError() == hipSuccess ? 0 : CUTC_ERR_HIP;
}

template <int S, bool F16>
static int slice_pairw(const double *wr2, const double *aos2, int64_t n2, int n_ao, int64_t kpad,
                       int64_t plane_rows, const int *rexp, uint16_t *planes, hipStream_t st)
{
    const int64_t P = (int64_t)n_ao * (n_ao + 1) / 2;
    hipLaunchKernelGGL((oz_slice_pairw_kernel<S, F16>), dim3((unsigned)(kpad / 64), (unsigned)((P + 31) / 32)),
                       dim3(256), 0, st, wr2, aos2, n2, n_ao, kpad / 32, plane_rows * kpad, rexp, planes);
    return hipGetLastError() == hipSuccess ? 0 : CUTC_ERR_HIP;
}

template <int S, bool F16>
static int slice_colk(const cutc_colsrc *src, int64_t rows, int64_t kp, int64_t plane_rows, const int *rexp,
                      uint16_t *planes, hipStream_t st)
{
    const int64_t ktot = kp * src->nsrc;
    hipLaunchKernelGGL((oz_slice_colk_kernel<S, F16>),
                       dim3((unsigned)((rows + 31) / 32), (unsigned)(kp / 64), src->nsrc), dim3(256), 0, st, *src,
                       rows, kp, ktot / 32, plane_rows * ktot, rexp, planes);
    return hipGetLastError() == hipSuccess ? 0 : CUTC_ERR_HIP;
}

#define OZ_DISPATCH(fn, S, F16, ...)                                                                                  \
    ((S) == 2 ? ((F16) ? fn<2, true>(__VA_ARGS__) : fn<2, false>(__VA_ARGS__))                                       \
              : (S) == 3 ? ((F16) ? fn<3, true>(__VA_ARGS__) : fn<3, false>(__VA_ARGS__))                            \
                         : ((F16) ? fn<4, true>(__VA_ARGS__) : fn<4, false>(__VA_ARGS__)))

static int oz_check(int S) { return S >= 2 && S <= OZ_MAXS; }

extern "C" int cutc_oz_slice_k(int S, int f16, const double *X, int64_t ld, int64_t rows, int64_t K, int64_t kpad,
                               int64_t plane_rows, const int *rexp, uint16_t *planes, hipStream_t st)
{
    if (rows <= 0 || kpad <= 0)
        return 0;
    if (!oz_check(S) || kpad % CUTC_PK_KALIGN || plane_rows < rows || (rows + 31) / 32 > 65535)
        return CUTC_ERR_ARG;
    return OZ_DISPATCH(slice_k, S, f16, X, ld, rows, K, kpad, plane_rows, rexp, planes, st);
}

extern "C" int cutc_oz_slice_pairw(int S, int f16, const double *wr2, const double *aos2, int64_t n2, int n_ao,
                                   int64_t kpad, int64_t plane_rows, const int *rexp, uint16_t *planes,
                                   hipStream_t st)
{
    const int64_t P = (int64_t)n_ao * (n_ao + 1) / 2;
    if (P <= 0 || kpad <= 0)
        return 0;
    if (!oz_check(S) || kpad % CUTC_PK_KALIGN || plane_rows < P || (P + 31) / 32 > 65535)
        return CUTC_ERR_ARG;
    return OZ_DISPATCH(slice_pairw, S, f16, wr2, aos2, n2, n_ao, kpad, plane_rows, rexp, planes, st);
}

extern "C" int cutc_oz_slice_colk(int S, int f16, const cutc_colsrc *src, int64_t rows, int64_t kp,
                                  int64_t plane_rows, const int *rexp, uint16_t *planes, hipStream_t st)
{
    if (rows <= 0 || kp <= 0)
        return 0;
    if (!oz_check(S) || kp % 64 || (kp * src->nsrc) % CUTC_PK_KALIGN || src->nsrc < 1 || src->nsrc > 3 ||
        plane_rows < rows)
        return CUTC_ERR_ARG;
    return OZ_DISPATCH(slice_colk, S, f16, src, rows, kp, plane_rows, rexp, planes, st);
}

// ---------------------------------------------------------------- GEMM

#define OZ_WAIT_VM0() __builtin_amdgcn_s_waitcnt((7 << 4) | (15 << 8))

// global -> LDS copy of 16 bytes per lane (lane i lands at M0 + 16 i), as in
// crt.hip: asm so that the compiler's wait insertion leaves it alone; it is
// awaited explicitly (OZ_WAIT_VM0).  m0 has no other use in this kernel.
__device__ static inline void oz_dma16(const void *g, int8_t *l)
{
    const unsigned la = __builtin_amdgcn_readfirstlane((unsigned)(uintptr_t)(lds8 *)l);
    asm volatile("s_mov_b32 m0, %0\n\ts_nop 0\n\tglobal_load_lds_dwordx4 %1, off" ::"s"(la), "v"(g) : "memory");
}

template <bool F16> __device__ static inline v16f oz_mfma(v4i a, v4i b, v16f c)
{
    if constexpr (F16)
        return __builtin_amdgcn_mfma_f32_32x32x16_f16(__builtin_bit_cast(v8h, a), __builtin_bit_cast(v8h, b), c, 0,
                                                      0, 0);
    else
        return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(v8bf, a), __builtin_bit_cast(v8bf, b), c,
                                                       0, 0, 0);
}

// A: S planes of pa elements, B likewise with pb (rows of both padded to
// CUTC_OZ_ROWS); KT steps of 32 K (a multiple of 8).  C[r][c] = the slice
// sum scaled by 2^(eA[r] + eB[c] - 16), for r < M, c < N (row-major, ld N).
// Block tile 64 A rows x 128 B rows, 4 waves of 32 x 64 (1 x 2 MFMA tiles,
// so that the S level accumulators and the FP64 sums of a wave fit its
// registers).  Per step the LDS stage holds A: 2 row blocks x S slices x
// 2 KB, then B: 4 row blocks x S x 2 KB; its 12 S DMA loads are spread over
// the 4 waves and issued between the first MFMAs of the step before.
template <int S, bool F16>
__global__ __launch_bounds__(256, 1) void oz_gemm_kernel(const uint16_t *__restrict__ A, int64_t pa,
                                                         const uint16_t *__restrict__ B, int64_t pb, int64_t KT,
                                                         int tiles_a, int tiles_b, int64_t M, int64_t N,
                                                         const int *__restrict__ eA, const int *__restrict__ eB,
                                                         double *__restrict__ C)
{
    constexpr int AB = 4 * S * 1024; // bytes of the A part of a stage
    constexpr int STAGE = 3 * AB;
    constexpr int NDMA = 3 * S; // DMA loads per wave per step
    __shared__ __attribute__((aligned(16))) int8_t lds[2 * STAGE];

    const int t = cutc_xcd_remap(blockIdx.x, gridDim.x);
    int ta, tb;
    cutc_group_tile<4>(t, tiles_a, tiles_b, ta, tb);
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int wm = wave >> 1, wn = wave & 1;
    const uint32_t lo = (uint32_t)lane * 16;
    const int64_t rbs = KT * 2048; // bytes per row block of a plane

    // the wave's DMA loads: q = NDMA wave + i; q < 4 S: A tile (q / 2) =
    // (row block, slice), half q % 2; then B tiles likewise.  LDS slot q.
    // Wave-uniform bases; the lane offset is added at issue.
    const int8_t *dbase[NDMA];
#pragma unroll
    for (int i = 0; i < NDMA; i++) {
        const int q = NDMA * wave + i;
        const bool isa = q < 4 * S;
        const int rem = isa ? q : q - 4 * S, tile = rem >> 1, half = rem & 1;
        const int rbl = tile / S, sl = tile % S;
        dbase[i] = (isa ? (const int8_t *)A + 2 * sl * pa + ((int64_t)ta * 2 + rbl) * rbs
                        : (const int8_t *)B + 2 * sl * pb + ((int64_t)tb * 4 + rbl) * rbs) +
                   half * 1024;
    }
    auto dma = [&](int i, int64_t kt, int stage) {
        oz_dma16(dbase[i] + kt * 2048 + lo, &lds[stage * STAGE + (NDMA * wave + i) * 1024]);
    };

    v16f acc[S][2];
    double acc64[2][16];
#pragma unroll
    for (int d = 0; d < S; d++)
#pragma unroll
        for (int j = 0; j < 2; j++)
            acc[d][j] = v16f{};
#pragma unroll
    for (int j = 0; j < 2; j++)
#pragma unroll
        for (int r = 0; r < 16; r++)
            acc64[j][r] = 0.0;

#pragma unroll
    for (int i = 0; i < NDMA; i++)
        dma(i, 0, 0);

    for (int64_t kt = 0; kt < KT; kt++) {
        const int cur = (int)(kt & 1);
        OZ_WAIT_VM0();
        __syncthreads();
        const int8_t *sa = &lds[cur * STAGE];
        const int8_t *sb = sa + AB;
        const bool more = kt + 1 < KT;
#pragma unroll
        for (int c = 0; c < 2; c++) {
            v4i fa[S], fb[2][S];
#pragma unroll
            for (int u = 0; u < S; u++) {
                fa[u] = *(const v4i *)(sa + (wm * S + u) * 2048 + c * 1024 + lo);
#pragma unroll
                for (int j = 0; j < 2; j++)
                    fb[j][u] = *(const v4i *)(sb + ((2 * wn + j) * S + u) * 2048 + c * 1024 + lo);
            }
            int q = 0;
#pragma unroll
            for (int s = 0; s < S; s++)
#pragma unroll
                for (int u = 0; u < S - s; u++)
#pragma unroll
                    for (int j = 0; j < 2; j++) {
                        acc[s + u][j] = oz_mfma<F16>(fa[s], fb[j][u], acc[s + u][j]);
                        // the next step's DMA, one load after each of the first MFMAs
                        if (c == 0 && q < NDMA && more)
                            dma(q, kt + 1, cur ^ 1);
                        q++;
                    }
            if (c == 0 && more) {
                // steps with fewer c = 0 MFMAs than loads (S = 2) issue the rest here
#pragma unroll
                for (int i = S * (S + 1); i < NDMA; i++)
                    dma(i, kt + 1, cur ^ 1);
            }
        }
        if ((kt & 7) == 7 || !more) {
            // flush the exact block levels into the FP64 sums
#pragma unroll
            for (int j = 0; j < 2; j++)
#pragma unroll
                for (int r = 0; r < 16; r++) {
                    double v = (double)acc[0][j][r];
#pragma unroll
                    for (int d = 1; d < S; d++)
                        v = fma((double)acc[d][j][r], ldexp(1.0, -8 * d), v);
                    acc64[j][r] += v;
                }
#pragma unroll
            for (int d = 0; d < S; d++)
#pragma unroll
                for (int j = 0; j < 2; j++)
                    acc[d][j] = v16f{};
        }
    }

    // result element r of lane l: row 8 (r/4) + 4 (l/32) + r%4, column l%32
    const int li = lane & 31, lg = lane >> 5;
#pragma unroll
    for (int j = 0; j < 2; j++) {
        const int64_t col = (int64_t)tb * 128 + 64 * wn + 32 * j + li;
        if (col >= N)
            continue;
        const int ec = eB[col] - 16;
#pragma unroll
        for (int r = 0; r < 16; r++) {
            const int64_t row = (int64_t)ta * 64 + 32 * wm + 8 * (r >> 2) + 4 * lg + (r & 3);
            if (row < M)
                C[row * N + col] = ldexp(acc64[j][r], eA[row] + ec);
        }
    }
}

template <int S, bool F16>
static int oz_gemm(const uint16_t *A, int64_t M, const uint16_t *B, int64_t N, int64_t kpad, const int *eA,
                   const int *eB, double *C, hipStream_t st)
{
    const int64_t ta = (M + 63) / 64, tb = (N + 127) / 128;
    hipLaunchKernelGGL((oz_gemm_kernel<S, F16>), dim3((unsigned)(ta * tb)), dim3(256), 0, st, A,
                       CUTC_OZ_ROWS(M) * kpad, B, CUTC_OZ_ROWS(N) * kpad, kpad / 32, (int)ta, (int)tb, M, N, eA, eB,
                       C);
    return hipGetLastError() == hipSuccess ? 0 : CUTC_ERR_HIP;
}

extern "C" int cutc_oz_gemm(int S, int f16, const uint16_t *A, int64_t M, const uint16_t *B, int64_t N, int64_t kpad,
                            const int *eA, const int *eB, double *C, hipStream_t st)
{
    if (M <= 0 || N <= 0)
        return 0;
    if (!oz_check(S) || kpad <= 0 || kpad % CUTC_PK_KALIGN)
        return CUTC_ERR_ARG;
    if (((M + 63) / 64) * ((N + 127) / 128) > 2147483647)
        return CUTC_ERR_ARG;
    return OZ_DISPATCH(oz_gemm, S, f16, A, M, B, N, kpad, eA, eB, C, st);
}

// ---------------------------------------------------------------- scatter

// C[j][p] (j = m n1c + i1, p packed pair) -> X[m][i0 + i1][(k,l) and (l,k)]
__global__ __launch_bounds__(256) void oz_scatter_x_kernel(const double *__restrict__ C, int64_t P, int64_t N,
                                                           const int *__restrict__ pairs, int n_ao, double *X,
                                                           int64_t n1, int64_t n1c, int64_t i0)
{
    const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (p >= P)
        return;
    const int k = pairs[2 * p], l = pairs[2 * p + 1];
    const int64_t n2 = (int64_t)n_ao * n_ao;
    for (int64_t j = blockIdx.y; j < N; j += gridDim.y) {
        const double v = C[j * P + p];
        const int64_t m = j / n1c, i1 = j - m * n1c;
        double *o = X + (m * n1 + i0 + i1) * n2;
        o[k + (int64_t)n_ao * l] = v;
        o[l + (int64_t)n_ao * k] = v;
    }
}

extern "C" int cutc_oz_scatter_x(const double *C, int64_t P, int64_t N, const int *pairs, int n_ao, double *X,
                                 int64_t n1, int64_t n1c, int64_t i0, hipStream_t st)
{
    if (P <= 0 || N <= 0)
        return 0;
    const int64_t gy = N < 65535 ? N : 65535;
    hipLaunchKernelGGL(oz_scatter_x_kernel, dim3((unsigned)((P + 255) / 256), (unsigned)gy), dim3(256), 0, st, C, P, N,
                       pairs, n_ao, X, n1, n1c, i0);
    return hipGetLastError() == hipSuccess ? 0 : CUTC_ERR_HIP;
}

// alpha sum_y C[y][q][p] (y < nsplit, added in order) into M; modes as the
// FP64 step-2 epilogue (0 store, 1 add, 2 add at (kl, ab) and subtract at
// (kl, ba))
__global__ __launch_bounds__(256) void scatter_m_kernel(const double *__restrict__ C, int nsplit, int64_t P,
                                                        double alpha, const int *__restrict__ pairs, int n_ao,
                                                        double *M, int mode)
{
    const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (p >= P)
        return;
    const int k = pairs[2 * p], l = pairs[2 * p + 1];
    const int64_t n2 = (int64_t)n_ao * n_ao;
    const int64_t kl = k + (int64_t)n_ao * l, lk = l + (int64_t)n_ao * k;
    for (int64_t q = blockIdx.y; q < P; q += gridDim.y) {
        double c = C[q * P + p];
        for (int y = 1; y < nsplit; y++)
            c += C[(int64_t)y * P * P + q * P + p];
        const double v = alpha * c;
        const int a = pairs[2 * q], b = pairs[2 * q + 1];
        double *o1 = M + n2 * (a + (int64_t)n_ao * b);
        double *o2 = M + n2 * (b + (int64_t)n_ao * a);
        if (mode == 0) {
            o1[kl] = v;
            o1[lk] = v;
            o2[kl] = v;
            o2[lk] = v;
        } else {
            const double sgn = (mode == 2) ? -v : v;
            o1[kl] += v;
            if (k != l)
                o1[lk] += v;
            if (a != b) {
                o2[kl] += sgn;
                if (k != l)
                    o2[lk] += sgn;
            }
        }
    }
}

extern "C" int cutc_scatter_m(const double *C, int nsplit, int64_t P, double alpha, const int *pairs, int n_ao,
                              double *M, int mode, hipStream_t st)
{
    if (P <= 0)
        return 0;
    if (nsplit < 1)
        return CUTC_ERR_ARG;
    const int64_t gy = P < 65535 ? P : 65535;
    hipLaunchKernelGGL(scatter_m_kernel, dim3((unsigned)((P + 255) / 256), (unsigned)gy), dim3(256), 0, st, C, nsplit,
                       P, alpha, pairs, n_ao, M, mode);
    return hipGetLastError() == hipSuccess ? 0 : CUTC_ERR_HIP;
}
