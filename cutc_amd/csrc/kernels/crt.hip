// Modular (CRT) split-precision contractions for gfx950: operands rounded to
// integers per row, products taken modulo n moduli, rebuilt by CRT into
// FP64.  Scheme: Ozaki, Uchino & Imamura, "Ozaki Scheme II" (constants and
// the rebuild formula in crt_host.c).  Two plane kinds: int8 residues
// (moduli <= 256) with v_mfma_i32_32x32x32_i8 (exact int32 accumulation),
// and residues |r| <= 127 (moduli <= 255) held as bf16 / fp16 values with
// v_mfma_f32_32x32x16_{bf16,f16}, exact because the FP32 accumulators are
// reduced modulo m often enough to stay below 2^24.
//
// Kernels:
//   rowexp_*        row exponents e with ||x_row||_2 < 2^e or max|x_row| < 2^e
//                   (fixed-order reductions)
//   quant_*_pk      x -> q = rint(x 2^(bits - e_row)) -> residues q mod m_i,
//                   int8 or 16-bit planes in the fragment-packed layouts
//                   (launch.h)
//   crt_gemm_pk     per modulus: 512 x 128 block tile, 4 waves of 128 x 128
//                   (4 x 4 MFMA 32x32x32 i8); A fragments streamed into a
//                   register ring, B through an LDS ring fed by LDS DMA;
//                   int8 residues C[i][r][p]
//   crt16_gemm      the same tiling and pipeline with the 16-bit MFMA and
//                   FP32 accumulators reduced modulo m every 960 K; int8
//                   residues
//   rebuild_*       CRT of the n residues per element, scaling, and the
//                   scatter to the reference's full layouts (as the FP64
//                   epilogues)
#include "common.h"
#include "launch.h"
#include "split_common.h"

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));
typedef float v16f __attribute__((ext_vector_type(16)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef _Float16 v8h __attribute__((ext_vector_type(8)));

// residue of the integer-valued double q (|q| < 2^52) modulo m, in
// [-(m / 2), (m - 1) / 2], within [-128, 127] for m <= 256 (the rounded
// quotient is within 1 of q / m, so one correction suffices)
__device__ static inline int residue(double q, int m, double inv_m)
{
    const double t = rint(q * inv_m);
    int r = (int)fma(-t, (double)m, q);
    if (r > ((m - 1) >> 1))
        r -= m;
    if (r < -(m >> 1))
        r += m;
    return r;
}

// Residue of the quantised value q (an integer-valued double, |q| <= 2^50)
// modulo m in [-(m / 2), m / 2], for the operand planes.  The quotient is
// rounded exactly: fl(q inv_m) is within (q / m) 2^-52 < 0.5 / m of q / m,
// closer than any exact quotient of an odd m comes to a half-integer, so
// rint picks the nearest integer and q - m rint(q / m) is exact.  The ends
// +-m/2 are only reached by the even modulus 256 at exact ties, where +128
// and -128 give the same byte; every other residue equals residue()'s.
__device__ static inline int residue_q(double q, int m, double inv_m)
{
    return (int)fma(-rint(q * inv_m), (double)m, q);
}

// Row exponent e with ||x||_2 < 2^e from the computed sum of squares ss:
// the sum carries a relative rounding error far below 2^-29 for the row
// lengths here (at most ~2^20 terms in fixed-order sums), so ss is inflated
// by that much before the root.  A zero row (or one whose squares underflow)
// gets e = 0, which rounds all its entries to 0.
__device__ static inline int norm_exponent(double ss)
{
    if (!(ss > 0.0))
        return 0;
    return ilogb(sqrt(ss * (1.0 + 0x1p-29))) + 1;
}

// Row exponent e with max |x| < 2^e (0 for a zero row)
__device__ static inline int max_exponent(double mx) { return (mx > 0.0) ? ilogb(mx) + 1 : 0; }

// Row statistics: MODE 0 sums squares (2-norm exponents), MODE 1 takes the
// largest magnitude (max exponents); both reductions run in a fixed order.
template <int MODE> __device__ static inline double rs_add(double v, double a)
{
    return MODE == 0 ? fma(a, a, v) : fmax(v, fabs(a));
}
template <int MODE> __device__ static inline double rs_join(double u, double v)
{
    return MODE == 0 ? u + v : fmax(u, v);
}
template <int MODE> __device__ static inline int rs_exponent(double v)
{
    return MODE == 0 ? norm_exponent(v) : max_exponent(v);
}

// block-wide reduction of v (256 threads) in a fixed order, returned to thread 0
template <int MODE> __device__ static inline double block_reduce(double v)
{
    __shared__ double red[4];
    for (int off = 32; off > 0; off >>= 1)
        v = rs_join<MODE>(v, __shfl_xor(v, off, 64));
    if ((threadIdx.x & 63) == 0)
        red[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0)
        v = rs_join<MODE>(rs_join<MODE>(red[0], red[1]), rs_join<MODE>(red[2], red[3]));
    return v;
}

// ---------------------------------------------------------------- row exponents

template <int MODE>
__global__ __launch_bounds__(256) void rowexp_k_kernel(const double *__restrict__ X, int64_t ld, int64_t K, int *rexp)
{
    const double *row = X + (int64_t)blockIdx.x * ld;
    double v = 0.0;
    for (int64_t k = threadIdx.x; k < K; k += 256)
        v = rs_add<MODE>(v, row[k]);
    v = block_reduce<MODE>(v);
    if (threadIdx.x == 0)
        rexp[blockIdx.x] = rs_exponent<MODE>(v);
}

extern "C" int cutc_rowexp_k(const double *X, int64_t ld, int64_t rows, int64_t K, int mode, int *rexp,
                             hipStream_t st)
{
    if (rows <= 0)
        return 0;
    if (rows > 2147483647)
        return CUTC_ERR_ARG;
    if (mode == CUTC_ROWEXP_MAX)
        hipLaunchKernelGGL(rowexp_k_kernel<1>, dim3((unsigned)rows), dim3(256), 0, st, X, ld, K, rexp);
    else
        hipLaunchKernelGGL(rowexp_k_kernel<0>, dim3((unsigned)rows), dim3(256), 0, st, X, ld, K, rexp);
    return hipGetLastError() == hipSuccess ? 0 : CUTC_ERR_HIP;
}

// row p = (k, l) of A[p][r2] = (w chi_l) chi_k
template <int MODE>
__global__ __launch_bounds__(256) void rowexp_pairw_kernel(const double *__restrict__ wr2,
                                                           const double *__restrict__ aos2, int64_t n2, int *rexp)
{
    const int64_t p = blockIdx.x;
    int k, l;
    pair_of(p, k, l);
    double v = 0.0;
    for (int64_t i = threadIdx.x; i < n2; i += 256)
        v = rs_add<MODE>(v, wr2[i] * aos2[i + n2 * l] * aos2[i + n2 * k]);
    v = block_reduce<MODE>(v);
    if (threadIdx.x == 0)
        rexp[p] = rs_exponent<MODE>(v);
}

extern "C" int cutc_rowexp_pairw(const double *wr2, const double *aos2, int64_t n2, int n_ao, int mode, int *rexp,
                                 hipStream_t st)
{
    const int64_t P = (int64_t)n_ao * (n_ao + 1) / 2;
    if (P <= 0)
        return 0;
    if (mode == CUTC_ROWEXP_MAX)
        hipLaunchKernelGGL(rowexp_pairw_kernel<1>, dim3((unsigned)P), dim3(256), 0, st, wr2, aos2, n2, rexp);
    else
        hipLaunchKernelGGL(rowexp_pairw_kernel<0>, dim3((unsigned)P), dim3(256), 0, st, wr2, aos2, n2, rexp);
    return hipGetLastError() == hipSuccess ? 0 : CUTC_ERR_HIP;
}

// grid (ceil(rows/256), ks): block y reduces its K slice of every segment
// into partial[y][r]; the slices are then joined in order
template <int MODE>
__global__ __launch_bounds__(256) void rowexp_colk_kernel(cutc_colsrc s, int64_t rows, double *partial)
{
    const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (r >= rows)
        return;
    const int64_t col = colsrc_col(s, r);
    const int64_t per = (s.K + gridDim.y - 1) / gridDim.y;
    const int64_t k0 = per * blockIdx.y, k1 = (k0 + per < s.K) ? k0 + per : s.K;
    double v = 0.0;
    for (int seg = 0; seg < s.nsrc; seg++) {
        const double *base = s.p[seg] + col;
        for (int64_t k = k0; k < k1; k++)
            v = rs_add<MODE>(v, base[colsrc_k(s, k) * s.ld]);
    }
    partial[(int64_t)blockIdx.y * rows + r] = v;
}

template <int MODE>
__global__ __launch_bounds__(256) void rowexp_finish_kernel(const double *__restrict__ partial, int ks, int64_t rows,
                                                            int *rexp)
{
    const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (r >= rows)
        return;
    double v = 0.0;
    for (int y = 0; y < ks; y++)
        v = rs_join<MODE>(v, partial[(int64_t)y * rows + r]);
    rexp[r] = rs_exponent<MODE>(v);
}

extern "C" int cutc_rowexp_finish(const double *partial, int64_t ks, int64_t rows, int mode, int *rexp,
                                  hipStream_t st)
{
    if (rows <= 0)
        return 0;
    if (ks < 1 || ks > 2147483647LL)
        return CUTC_ERR_ARG;
    const dim3 g2((unsigned)((rows + 255) / 256));
    if (mode == CUTC_ROWEXP_MAX)
        hipLaunchKernelGGL(rowexp_finish_kernel<1>, g2, dim3(256), 0, st, partial, (int)ks, rows, rexp);
    else
        hipLaunchKernelGGL(rowexp_finish_kernel<0>, g2, dim3(256), 0, st, partial, (int)ks, rows, rexp);
    return hipGetLastError() == hipSuccess ? 0 : CUTC_ERR_HIP;
}

extern "C" int cutc_rowexp_colk(const cutc_colsrc *src, int64_t rows, int mode, double *partial, int *rexp,
                                hipStream_t st)
{
    if (rows <= 0)
        return 0;
    int64_t ks = (src->K + 511) / 512;
    if (ks > CUTC_CRT_KSPLIT)
        ks = CUTC_CRT_KSPLIT;
    if (ks < 1)
        ks = 1;
    const dim3 g1((unsigned)((rows + 255) / 256), (unsigned)ks), g2((unsigned)((rows + 255) / 256));
    if (mode == CUTC_ROWEXP_MAX) {
        hipLaunchKernelGGL(rowexp_colk_kernel<1>, g1, dim3(256), 0, st, *src, rows, partial);
        hipLaunchKernelGGL(rowexp_finish_kernel<1>, g2, dim3(256), 0, st, partial, (int)ks, rows, rexp);
    } else {
        hipLaunchKernelGGL(rowexp_colk_kernel<0>, g1, dim3(256), 0, st, *src, rows, partial);
        hipLaunchKernelGGL(rowexp_finish_kernel<0>, g2, dim3(256), 0, st, partial, (int)ks, rows, rexp);
    }
    return hipGetLastError() == hipSuccess ? 0 : CUTC_ERR_HIP;
}

// ---------------------------------------------------------------- residues

// x mod m into [-(m / 2), (m - 1) / 2] for |x| < 2^31
__device__ static inline int reduce_mod(int x, int m, double inv_m) { return residue((double)x, m, inv_m); }

// ------------------------------------------------ fragment-packed planes

// Packed byte offset of (row 32 rb + r, k) in an int8 plane (k a multiple
// of 16): the layout fed to crt_gemm_pk_kernel (launch.h), KT = kpad / 64.
__device__ static inline int64_t pk_offset(int64_t rb, int r, int64_t k, int64_t KT)
{
    const int64_t kt = k >> 6;
    const int kk = (int)(k & 63);
    const int h = kk >> 5, c = (kk >> 4) & 1;
    return (rb * KT + kt) * 2048 + c * 1024 + (32 * h + r) * 16;
}

// the same for a 16-bit plane (k a multiple of 8), KT = kpad / 32
__device__ static inline int64_t pk16_offset(int64_t rb, int r, int64_t k, int64_t KT)
{
    const int64_t kt = k >> 5;
    const int kk = (int)(k & 31);
    const int c = kk >> 4, h = (kk >> 3) & 1;
    return (rb * KT + kt) * 2048 + c * 1024 + (32 * h + r) * 16;
}

// 16-bit pattern of the integer r (|r| <= 256) as a bf16 or fp16 value:
// exact in both, so the bf16 pattern is the upper half of the float's bits
template <int KIND> __device__ static inline unsigned half_bits(int r)
{
    if constexpr (KIND == CUTC_CRT_FP16)
        return __builtin_bit_cast(uint16_t, (_Float16)(float)r);
    else
        return __float_as_uint((float)r) >> 16;
}

// K steps per plane of kpad columns
__host__ __device__ static inline int64_t pk_steps(int kind, int64_t kpad)
{
    return kind == CUTC_CRT_I8 ? kpad / 64 : kpad / 32;
}

// 16 values k0 .. k0 + 15 (k0 a multiple of 16) of row 32 rb + r -> n
// residue planes: one 16-byte int8 chunk, or two 16-byte chunks of 8
// 16-bit residues
template <int KIND>
__device__ static inline void store_residues(const cutc_crt &c, const double (&q)[16], int8_t *planes, int64_t rb,
                                             int r, int64_t k0, int64_t KT, int64_t plane)
{
    if constexpr (KIND == CUTC_CRT_I8) {
        int8_t *dst = planes + pk_offset(rb, r, k0, KT);
        for (int i = 0; i < c.n; i++) {
            int w[4];
#pragma unroll
            for (int g = 0; g < 4; g++) {
                unsigned u = 0;
#pragma unroll
                for (int e = 0; e < 4; e++)
                    u |= ((unsigned)residue_q(q[4 * g + e], c.m[i], c.inv_m[i]) & 0xffu) << (8 * e);
                w[g] = (int)u;
            }
            *(v4i *)(dst + i * plane) = v4i{w[0], w[1], w[2], w[3]};
        }
    } else {
        int8_t *dst0 = planes + pk16_offset(rb, r, k0, KT);
        int8_t *dst1 = planes + pk16_offset(rb, r, k0 + 8, KT);
        for (int i = 0; i < c.n; i++) {
            int w[8];
#pragma unroll
            for (int g = 0; g < 8; g++)
                w[g] = (int)(half_bits<KIND>(residue_q(q[2 * g], c.m[i], c.inv_m[i])) |
                             (half_bits<KIND>(residue_q(q[2 * g + 1], c.m[i], c.inv_m[i])) << 16));
            *(v4i *)(dst0 + i * plane) = v4i{w[0], w[1], w[2], w[3]};
            *(v4i *)(dst1 + i * plane) = v4i{w[4], w[5], w[6], w[7]};
        }
    }
}

// kind -> template instance
#define CRT_KIND_DISPATCH(kind, F)                                                                                     \
    do {                                                                                                               \
        if ((kind) == CUTC_CRT_I8) {                                                                                   \
            F(CUTC_CRT_I8);                                                                                            \
        } else if ((kind) == CUTC_CRT_BF16) {                                                                          \
            F(CUTC_CRT_BF16);                                                                                          \
        } else if ((kind) == CUTC_CRT_FP16) {                                                                          \
            F(CUTC_CRT_FP16);                                                                                          \
        } else {                                                                                                       \
            return CUTC_ERR_ARG;                                                                                       \
        }                                                                                                              \
    } while (0)

// ---------------------------------------------------------------- quantisation

// grid (kpad / 128, ceil(rows / 32)): a block covers one row block and 128
// K.  The 32 x 128 tile of X is first read into LDS with coalesced 16-byte
// loads (8 consecutive lanes cover 128 contiguous bytes of one row); then
// thread r = tid % 32 takes row 32 rb + r, the 16 values from 16 j
// (j = tid / 32), so the 16-byte stores of consecutive threads are
// consecutive.
template <int KIND>
__global__ __launch_bounds__(256) void quant_k_pk_kernel(cutc_crt c, const double *__restrict__ X, int64_t ld,
                                                         int64_t rows, int64_t K, int64_t KT, int64_t plane,
                                                         const int *__restrict__ rexp, int8_t *planes)
{
    __shared__ double tile[32][128 + 1]; // odd row pitch: the column reads below are conflict-free
    const int64_t rb = blockIdx.y, kb = (int64_t)blockIdx.x * 128;
    {
        const int rr = threadIdx.x >> 3, seg = threadIdx.x & 7; // 32 rows x 8 lanes
        const int64_t row = rb * 32 + rr;
        const bool pairs = (((uintptr_t)X | (uintptr_t)(ld * 8)) & 15) == 0;
#pragma unroll
        for (int it = 0; it < 8; it++) {
            const int col = 16 * it + 2 * seg; // lanes of a row: 16 contiguous bytes each, 128 per pass
            const int64_t k = kb + col;
            double v0 = 0.0, v1 = 0.0;
            if (row < rows) {
                const double *src = X + row * ld + k;
                if (k + 1 < K) {
                    if (pairs) {
                        const v2d v = *(const v2d *)src;
                        v0 = v[0];
                        v1 = v[1];
                    } else {
                        v0 = src[0];
                        v1 = src[1];
                    }
                } else if (k < K) {
                    v0 = src[0];
                }
            }
            tile[rr][col] = v0;
            tile[rr][col + 1] = v1;
        }
    }
    __syncthreads();
    const int r = threadIdx.x & 31, j = threadIdx.x >> 5;
    const int64_t row = rb * 32 + r;
    const int64_t k0 = kb + 16 * j;
    double q[16];
    const int sh = (row < rows) ? c.bits - rexp[row] : 0;
#pragma unroll
    for (int e = 0; e < 16; e++)
        q[e] = rint(ldexp(tile[r][16 * j + e], sh));
    store_residues<KIND>(c, q, planes, rb, r, k0, KT, plane);
}

extern "C" int cutc_crt_quant_k_pk(const cutc_crt *c, const double *X, int64_t ld, int64_t rows, int64_t K,
                                   int64_t kpad, const int *rexp, int8_t *planes, hipStream_t st)
{
    if (rows <= 0 || kpad <= 0)
        return 0;
    if (kpad % CUTC_PK_KALIGN || (rows + 31) / 32 > 65535)
        return CUTC_ERR_ARG;
    const int64_t plane = CUTC_PK_ROWS_A(rows) * kpad * CUTC_CRT_RBYTES(c->kind);
    const dim3 grid((unsigned)(kpad / 128), (unsigned)((rows + 31) / 32));
#define QK(K_)                                                                                                         \
    hipLaunchKernelGGL(quant_k_pk_kernel<K_>, grid, dim3(256), 0, st, *c, X, ld, rows, K, pk_steps(K_, kpad), plane,   \
                       rexp, planes)
    CRT_KIND_DISPATCH(c->kind, QK);
#undef QK
    return hipGetLastError() == hipSuccess ? 0 : CUTC_ERR_HIP;
}

// pair rows p = (k, l) of A[p][r2] = (w chi_l) chi_k, computed on the fly;
// (k, l) from the table ptab (2 ints per row) or, without one, the packed
// pair index.  occ (optional): byte [tb][kt] set to 1 when a row of the
// CUTC_SP_ROWS-row B tile tb has a non-zero quantised value in K step kt (64 K for
// int8, 32 for the 16-bit kinds); the caller clears it first.
// one block per (32-row block rb, 128-point range), nrb = ceil(P / 32) row
// blocks fastest; as in quant_k_pk_kernel the 32 x 128 tile of
// products is formed in LDS from coalesced reads (8 consecutive lanes read
// 128 contiguous bytes of one AO row), then thread r = tid % 32 quantises
// row 32 rb + r at the 16 points of j = tid / 32.  A wave's 64 threads then
// cover 32 rows x 32 K inside one step of either size, so one ballot per
// wave decides the flag.
template <int KIND>
__global__ __launch_bounds__(256) void quant_pairw_pk_kernel(cutc_crt c, const double *__restrict__ wr2,
                                                             const double *__restrict__ aos2, int64_t ld2,
                                                             int64_t n2, int n_ao, int64_t KT, int64_t plane,
                                                             const int *__restrict__ rexp,
                                                             const int *__restrict__ ptab, int8_t *planes,
                                                             uint8_t *occ, int nrb)
{
    __shared__ double tile[32][128 + 1]; // odd row pitch: the column reads below are conflict-free
    // row blocks fastest: the blocks in flight share one 128-point range of
    // r2, whose AO values (n_ao x 1 KB) stay in L2 for all of them
    const int64_t rb = blockIdx.x % nrb, kb = (int64_t)(blockIdx.x / nrb) * 128;
    const int64_t P = (int64_t)n_ao * (n_ao + 1) / 2;
    {
        const int rr = threadIdx.x >> 3, seg = threadIdx.x & 7; // 32 rows x 8 lanes
        const int64_t p = rb * 32 + rr;
        int k = 0, l = 0;
        if (p < P) {
            if (ptab) {
                k = ptab[2 * p];
                l = ptab[2 * p + 1];
            } else {
                pair_of(p, k, l);
            }
        }
        const double *al = aos2 + ld2 * l, *ak = aos2 + ld2 * k;
#pragma unroll
        for (int it = 0; it < 8; it++) {
            const int col = 16 * it + 2 * seg;
            const int64_t i = kb + col;
            double v0 = 0.0, v1 = 0.0;
            if (p < P) {
                if (i < n2)
                    v0 = wr2[i] * al[i] * ak[i];
                if (i + 1 < n2)
                    v1 = wr2[i + 1] * al[i + 1] * ak[i + 1];
            }
            tile[rr][col] = v0;
            tile[rr][col + 1] = v1;
        }
    }
    __syncthreads();
    const int r = threadIdx.x & 31, j = threadIdx.x >> 5;
    const int64_t p = rb * 32 + r;
    const int64_t k0 = kb + 16 * j;
    double q[16];
    bool nz = false;
    const int sh = (p < P) ? c.bits - rexp[p] : 0;
#pragma unroll
    for (int e = 0; e < 16; e++) {
        q[e] = rint(ldexp(tile[r][16 * j + e], sh));
        nz |= q[e] != 0.0;
    }
    store_residues<KIND>(c, q, planes, rb, r, k0, KT, plane);
    if (occ) {
        const uint64_t b = __ballot(nz);
        if (b && (threadIdx.x & 63) == 0)
            occ[(rb / (CUTC_SP_ROWS / 32)) * KT + k0 / (KIND == CUTC_CRT_I8 ? 64 : 32)] = 1;
    }
}

extern "C" int cutc_crt_quant_pairw_pk(const cutc_crt *c, const double *wr2, const double *aos2, int64_t n2,
                                       int n_ao, int64_t kpad, const int *rexp, int8_t *planes, hipStream_t st)
{
    return cutc_crt_quant_pairw_pk_range(c, wr2, aos2, n2, n2, n_ao, kpad, rexp, planes, st);
}

/* r2 points [0, n2) of wr2 / aos2 (whose AO columns are ld2 apart): the
 * planes of one K pass when the caller offsets wr2 and aos2 to its start */
extern "C" int cutc_crt_quant_pairw_pk_range(const cutc_crt *c, const double *wr2, const double *aos2, int64_t ld2,
                                             int64_t n2, int n_ao, int64_t kpad, const int *rexp, int8_t *planes,
                                             hipStream_t st)
{
    return cutc_crt_quant_pairw_pk_occ(c, wr2, aos2, ld2, n2, n_ao, kpad, rexp, NULL, planes, NULL, st);
}

extern "C" int cutc_crt_quant_pairw_pk_occ(const cutc_crt *c, const double *wr2, const double *aos2, int64_t ld2,
                                           int64_t n2, int n_ao, int64_t kpad, const int *rexp, const int *ptab,
                                           int8_t *planes, uint8_t *occ, hipStream_t st)
{
    const int64_t P = (int64_t)n_ao * (n_ao + 1) / 2;
    if (P <= 0 || kpad <= 0)
        return 0;
    const int64_t nrb = (P + 31) / 32;
    if (kpad % CUTC_PK_KALIGN || nrb * (kpad / 128) > 2147483647LL)
        return CUTC_ERR_ARG;
    const int64_t plane = CUTC_PK_ROWS_B(P) * kpad * CUTC_CRT_RBYTES(c->kind);
    const dim3 grid((unsigned)(nrb * (kpad / 128)));
#define QP(K_)                                                                                                         \
    hipLaunchKernelGGL(quant_pairw_pk_kernel<K_>, grid, dim3(256), 0, st, *c, wr2, aos2, ld2, n2, n_ao,                \
                       pk_steps(K_, kpad), plane, rexp, ptab, planes, occ, (int)nrb)
    CRT_KIND_DISPATCH(c->kind, QP);
#undef QP
    return hipGetLastError() == hipSuccess ? 0 : CUTC_ERR_HIP;
}

// x - m rint(x / m) in float: keeps x mod m with |result| < 3 m for
// |x| < 2^31 (the float quotient is within 2 of the exact one)
__device__ static inline int shrink_mod(int x, int m, float inv_m)
{
    const int q = (int)__builtin_rintf((float)x * inv_m);
    return x - q * m;
}

// column-major sources (stage 2): a block packs 32 rows x 128 K of one
// segment, read coalesced along the rows into LDS and written as 16-byte
// fragments.  grid (ceil(rows / 32), kp / 128, nsrc); segment s occupies K
// [s kp, (s + 1) kp), zero past the valid s.K.
// occ (optional): the non-zero flags of the (CUTC_SP_ROWS-row tile, K step)
// tiles, one ballot per wave (32 rows x 32 K inside one step), as in
// quant_pairw_pk_kernel.
template <int KIND>
__global__ __launch_bounds__(256) void quant_colk_pk_kernel(cutc_crt c, cutc_colsrc s, int64_t rows, int64_t kp,
                                                            int64_t KT, int64_t plane,
                                                            const int *__restrict__ rexp, int8_t *planes,
                                                            uint8_t *occ)
{
    __shared__ double tile[128][33];
    const int64_t rb = blockIdx.x, r0 = rb * 32, k0 = (int64_t)blockIdx.y * 128;
    const int seg = blockIdx.z;
    {
        const int rr = threadIdx.x & 31;
        const int64_t r = r0 + rr;
        const bool rok = r < rows;
        const int64_t col = colsrc_col(s, rok ? r : 0);
        const double *src = s.p[seg] + col;
        for (int kk = threadIdx.x >> 5; kk < 128; kk += 8) {
            const int64_t k = k0 + kk;
            tile[kk][rr] = (rok && k < s.K) ? src[colsrc_k(s, k) * s.ld] : 0.0;
        }
    }
    __syncthreads();
    const int r = threadIdx.x & 31, j = threadIdx.x >> 5;
    const int64_t row = r0 + r;
    double q[16];
    bool nz = false;
    const int sh = (row < rows) ? c.bits - rexp[row] : 0;
#pragma unroll
    for (int e = 0; e < 16; e++) {
        q[e] = rint(ldexp(tile[16 * j + e][r], sh));
        nz |= q[e] != 0.0;
    }
    store_residues<KIND>(c, q, planes, rb, r, seg * kp + k0 + 16 * j, KT, plane);
    if (occ) {
        const uint64_t b = __ballot(nz);
        if (b && (threadIdx.x & 63) == 0)
            occ[(rb / (CUTC_SP_ROWS / 32)) * KT + (seg * kp + k0 + 16 * j) / (KIND == CUTC_CRT_I8 ? 64 : 32)] = 1;
    }
}

extern "C" int cutc_crt_quant_colk_pk(const cutc_crt *c, const cutc_colsrc *src, int64_t rows, int64_t kp,
                                      int64_t plane_rows, const int *rexp, int8_t *planes, hipStream_t st)
{
    return cutc_crt_quant_colk_pk_occ(c, src, rows, kp, plane_rows, rexp, planes, NULL, st);
}

extern "C" int cutc_crt_quant_colk_pk_occ(const cutc_crt *c, const cutc_colsrc *src, int64_t rows, int64_t kp,
                                          int64_t plane_rows, const int *rexp, int8_t *planes, uint8_t *occ,
                                          hipStream_t st)
{
    if (rows <= 0 || kp <= 0)
        return 0;
    if (kp % 128 || (kp * src->nsrc) % CUTC_PK_KALIGN || src->nsrc < 1 || src->nsrc > 3 || plane_rows < rows ||
        (rows + 31) / 32 > 2147483647)
        return CUTC_ERR_ARG;
    const int64_t ktot = kp * src->nsrc;
    const dim3 grid((unsigned)((rows + 31) / 32), (unsigned)(kp / 128), src->nsrc);
#define QC(K_)                                                                                                         \
    hipLaunchKernelGGL(quant_colk_pk_kernel<K_>, grid, dim3(256), 0, st, *c, *src, rows, kp, pk_steps(K_, ktot),       \
                       plane_rows * ktot * CUTC_CRT_RBYTES(K_), rexp, planes, occ)
    CRT_KIND_DISPATCH(c->kind, QC);
#undef QC
    return hipGetLastError() == hipSuccess ? 0 : CUTC_ERR_HIP;
}

typedef __attribute__((address_space(3))) int8_t lds_i8;
// the moduli of one launch of the GEMM kernels (block z of the grid's
// moduli uses m[z], inv_m[z])
struct gemm_moduli {
    int m[CUTC_CRT_MAX];
    double inv_m[CUTC_CRT_MAX];
};

#define PK_TILE 2048 // bytes of one (32-row block, 64-byte K step) tile

// s_waitcnt with only vmcnt constrained (gfx9 encoding: vmcnt in bits 3:0 and
// 15:14, expcnt 6:4 and lgkmcnt 11:8 left at their maxima)
#define PK_WAIT_VM(n) __builtin_amdgcn_s_waitcnt((((n) & 15) | (((n) >> 4) << 14) | (7 << 4) | (15 << 8)))

// global -> LDS copy of 16 bytes per lane (lane i lands at M0 + 16 i).
// Written as asm so that the compiler's wait insertion, which cannot tell
// the LDS stages apart and would wait for every outstanding load, leaves the
// DMA alone; its completion is awaited explicitly (PK_WAIT_VM).  The
// compiler's own waits for the A registers then only over-count (the DMA
// loads it does not see are younger), which is safe.  m0 has no other use in
// this kernel.
__device__ static inline void pk_dma16(const int8_t *g, int8_t *l)
{
    const unsigned la = __builtin_amdgcn_readfirstlane((unsigned)(uintptr_t)(lds_i8 *)l);
    asm volatile("s_mov_b32 m0, %0\n\ts_nop 0\n\tglobal_load_lds_dwordx4 %1, off" ::"s"(la), "v"(g) : "memory");
}

// Block tile (512 / WB) A rows x 128 WB B rows, 4 waves of 128 x 128 (4 x 4
// MFMA 32x32x32 i8, 256 accumulator registers); WB = 1: the waves split the
// A rows, WB = 2: 2 x 2 waves, each A half read by two waves and each B half
// shared by two.  Each wave streams its own A rows straight from the packed
// planes into a D-slot register ring (one coalesced 1 KB load per
// fragment); the block's B rows go global -> LDS (pk_dma16, WB row blocks
// per wave) into a D-stage ring.  During K step k (64 bytes, 32 MFMAs per
// wave) every wave issues the 2 WB DMA and 8 A loads of step k + L and
// reads the c = 0 fragments of step k + 1.  A barrier opens every BAR-th
// step, after the wait for the DMA of the next BAR steps.  The stage written
// in step k belongs to step k + L - D, which every wave has finished once the
// last barrier (at k - BAR + 1 or later) has passed: D >= L + BAR.  Loads
// past the last step repeat the last step (valid memory, unused).  KT is a
// multiple of D; K runs in segments of SEG steps (a multiple of D,
// SEG 64 2^14 < 2^31), each a full pipeline (prologue, steady state, drain),
// so that the int32 sums can be shrunk modulo m between segments while no
// fragment registers are live.
// SP: the block runs only the kcount[tb] steps listed in klist[tb KT ..]
// (a multiple of D; the steps whose B tile has a non-zero residue, padded
// with a step whose B tile is zero), each segment's byte offsets staged in
// LDS (sK) and read one step ahead, so the offset of step k + L is in a
// register when step k issues its loads.  WB = 2 halves the A bytes per
// MFMA, which bounds the list-driven GEMM: blocks of different B tiles run
// different steps, so their A tiles are not shared through L2.  LA: the
// lists belong to the A tiles (klist[ta KT ..]) instead.
template <int D, int L, int BAR, int SEG, bool SP, int WB, bool LA>
__global__ __launch_bounds__(256, 1) void crt_gemm_pk_kernel(const int8_t *__restrict__ A,
                                                             const int8_t *__restrict__ B, int64_t KT, int tiles_a,
                                                             int tiles_b, int64_t M, int64_t N, gemm_moduli mods,
                                                             int nsplit, int64_t pa, int64_t pb,
                                                             int8_t *__restrict__ C, int accum,
                                                             const uint32_t *__restrict__ klist,
                                                             const int *__restrict__ kcount, int64_t sstride)
{
    static_assert(SEG % D == 0 && D % BAR == 0 && (int64_t)SEG * 64 * 16384 < 2147483648LL, "segment");
    static_assert(D >= L + BAR && L >= BAR + 1 && (8 + 2 * WB) * L <= 63 && (WB == 1 || WB == 2), "pipeline shape");
    __shared__ __attribute__((aligned(16))) int8_t sB[D][4 * WB * PK_TILE];
    __shared__ uint32_t sK[SP ? SEG : 1];

    // one launch runs every modulus z (and every part y of a K split): the
    // grid is z-major, then y, then the tiles; consecutive blocks share the
    // listed tile (4 tiles of the other side each): they run the same list
    // in step and are replaced together, so the listed tile is read once per
    // group through L2
    int t = cutc_xcd_remap(blockIdx.x, gridDim.x);
    const int zy = t / (tiles_a * tiles_b), z = zy / nsplit, y = zy - z * nsplit;
    t -= zy * (tiles_a * tiles_b);
    A += (int64_t)z * pa;
    B += (int64_t)z * pb;
    C += (int64_t)z * M * N;
    const int mod = mods.m[z];
    const double inv_mod = mods.inv_m[z];
    int ta, tb;
    if constexpr (LA)
        cutc_group_tile<4>(t, tiles_b, tiles_a, tb, ta);
    else
        cutc_group_tile<4>(t, tiles_a, tiles_b, ta, tb);
    const int lt = LA ? ta : tb;                       // the tile whose list is run
    int64_t NS = SP ? (int64_t)kcount[lt] : KT;        // steps run
    const uint32_t *kl = SP ? klist + (int64_t)lt * KT : nullptr;
    if constexpr (SP) {
        // K split in nsplit parts: part y runs its share of the list, in
        // whole groups of D steps (the list's padding unit), into its own
        // residues sstride bytes on
        if (nsplit > 1) {
            const int64_t g = NS / D;
            const int64_t s0 = g * y / nsplit * D, s1 = g * (y + 1) / nsplit * D;
            kl += s0;
            NS = s1 - s0;
            C += (int64_t)y * sstride;
        }
    }
    int64_t kbase = 0;                                 // first step of the current segment
    uint32_t so_nx = 0;                                // SP: offset of the step L + 1 ahead
    auto off = [&](int64_t s) -> uint32_t {
        if constexpr (SP)
            return sK[s - kbase];
        else
            return (uint32_t)s * PK_TILE;
    };
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int64_t rbs = KT * PK_TILE; // bytes per row block
    const int wr = WB == 1 ? wave : wave >> 1, wc = WB == 1 ? 0 : wave & 1; // A and B quarter / half
    const int8_t *Aw = A + ((int64_t)ta * (16 / WB) + wr * 4) * rbs;
    const int8_t *Bw = B + ((int64_t)tb * 4 * WB + wave * WB) * rbs; // this wave's DMA row blocks
    const uint32_t lo = (uint32_t)lane * 16;
    const float inv_modf = (float)inv_mod;
    // at a barrier step b the DMA of step b + BAR (issued during step
    // b + BAR - L, loads in issue order: 2 WB DMA then 8 A per step) is
    // complete once no more than its own 8 A loads and the 8 + 2 WB loads of
    // each of the L - BAR - 1 steps in between remain
    constexpr int NLD = 8 + 2 * WB;
    constexpr int WAIT_NEXT = 8 + NLD * (L - BAR - 1);

    v4i ar[D][4][2];
    v4i bf0[4], bf1[4];
    v16i acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++)
            acc[i][j] = v16i{};

    auto issue = [&](auto slot_c, int64_t s) {
        constexpr int slot = decltype(slot_c)::value;
        const uint32_t so = off(s);
#pragma unroll
        for (int d = 0; d < WB; d++) {
            pk_dma16(Bw + ((uint32_t)d * (uint32_t)rbs + so + lo), &sB[slot][(wave * WB + d) * PK_TILE]);
            pk_dma16(Bw + ((uint32_t)d * (uint32_t)rbs + so + 1024 + lo), &sB[slot][(wave * WB + d) * PK_TILE + 1024]);
        }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int i = 0; i < 4; i++) {
            ar[slot][i][0] = *(const v4i *)(Aw + ((uint32_t)i * (uint32_t)rbs + so + lo));
            ar[slot][i][1] = *(const v4i *)(Aw + ((uint32_t)i * (uint32_t)rbs + so + 1024 + lo));
        }
        __builtin_amdgcn_sched_barrier(0);
    };

    // One K step.  With one wave per SIMD every instruction issues in order,
    // so the loads are spread between the MFMAs (one after each of the first
    // 8 + 2 WB c = 0 MFMAs, the next step's fragment reads after the first
    // four c = 1 MFMAs) instead of stalling the matrix pipe behind a burst.
    auto step = [&](auto u_c, int64_t k, int64_t k1) {
        constexpr int u = decltype(u_c)::value;
        constexpr int slot = (u + L) % D;
        if constexpr (u % BAR == 0) {
            PK_WAIT_VM(WAIT_NEXT);
            __syncthreads();
        }
#pragma unroll
        for (int j = 0; j < 4; j++)
            bf1[j] = *(const v4i *)&sB[u][(wc * 4 + j) * PK_TILE + 1024 + lo];
        __builtin_amdgcn_sched_barrier(0);
        uint32_t so;
        if constexpr (SP)
            so = so_nx;
        else
            so = off(k + L < k1 ? k + L : k1 - 1);
        cutc_static_for<0, 16>([&](auto q_c) {
            constexpr int q = decltype(q_c)::value;
            constexpr int i = q / 4, j = q % 4;
            acc[i][j] = __builtin_amdgcn_mfma_i32_32x32x32_i8(ar[u][i][0], bf0[j], acc[i][j], 0, 0, 0);
            if constexpr (q < 2 * WB) {
                constexpr int d = q / 2, h = q % 2;
                pk_dma16(Bw + ((uint32_t)d * (uint32_t)rbs + so + h * 1024 + lo),
                         &sB[slot][(wave * WB + d) * PK_TILE + h * 1024]);
            } else if constexpr (q < NLD) {
                constexpr int ii = (q - 2 * WB) / 2, cc = (q - 2 * WB) % 2;
                ar[slot][ii][cc] = *(const v4i *)(Aw + ((uint32_t)ii * (uint32_t)rbs + so + cc * 1024 + lo));
            }
            __builtin_amdgcn_sched_barrier(0);
        });
        cutc_static_for<0, 16>([&](auto q_c) {
            constexpr int q = decltype(q_c)::value;
            constexpr int i = q / 4, j = q % 4;
            acc[i][j] = __builtin_amdgcn_mfma_i32_32x32x32_i8(ar[u][i][1], bf1[j], acc[i][j], 0, 0, 0);
            if constexpr (SP && q == 0)
                so_nx = off(k + 1 + L < k1 ? k + 1 + L : k1 - 1);
            if constexpr (q < 4)
                bf0[q] = *(const v4i *)&sB[(u + 1) % D][(wc * 4 + q) * PK_TILE + lo];
            __builtin_amdgcn_sched_barrier(0);
        });
    };

    for (int64_t k0 = 0; k0 < NS; k0 += SEG) {
        const int64_t k1 = (NS - k0 < SEG) ? NS : k0 + SEG;
        if constexpr (SP) {
            // the previous segment's last reads of sK are behind its drain barrier
            kbase = k0;
            for (int i = threadIdx.x; i < k1 - k0; i += 256)
                sK[i] = kl[k0 + i] * PK_TILE;
            __syncthreads();
            so_nx = off(k0 + L < k1 ? k0 + L : k1 - 1);
        }
        // prologue: steps k0 .. k0 + L - 1 in flight; DMA of the first BAR + 1 complete
        cutc_static_for<0, L>([&](auto s_c) {
            constexpr int s = decltype(s_c)::value;
            issue(s_c, (k0 + s < k1 ? k0 + s : k1 - 1));
        });
        PK_WAIT_VM(WAIT_NEXT);
        __syncthreads();
#pragma unroll
        for (int j = 0; j < 4; j++)
            bf0[j] = *(const v4i *)&sB[0][(wc * 4 + j) * PK_TILE + lo];

        for (int64_t kb = k0; kb < k1; kb += D)
            cutc_static_for<0, D>([&](auto u_c) { step(u_c, kb + decltype(u_c)::value, k1); });

        // drain: the repeated loads of the last steps land before the next
        // segment's prologue reuses their LDS stages
        PK_WAIT_VM(0);
        __syncthreads();
        if (k1 < NS) {
#pragma unroll
            for (int i = 0; i < 4; i++)
#pragma unroll
                for (int j = 0; j < 4; j++)
#pragma unroll
                    for (int r = 0; r < 16; r++)
                        acc[i][j][r] = shrink_mod(acc[i][j][r], mod, inv_modf);
        }
    }

    // result element r of lane l: A row 8 (r/4) + 4 (l/32) + r%4, B row l%32
    const int li = lane & 31, lg = lane >> 5;
#pragma unroll
    for (int i = 0; i < 4; i++) {
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int64_t col = (int64_t)tb * 128 * WB + wc * 128 + 32 * j + li;
            if (col >= N)
                continue;
#pragma unroll
            for (int r = 0; r < 16; r++) {
                const int64_t row = (int64_t)ta * (512 / WB) + wr * 128 + 32 * i + 8 * (r >> 2) + 4 * lg + (r & 3);
                if (row < M)
                    C[row * N + col] =
                        (int8_t)reduce_mod(acc[i][j][r] + (accum ? (int)C[row * N + col] : 0), mod, inv_mod);
            }
        }
    }
}

// v_mfma_f32_32x32x16_{bf16,f16} on two 16-byte fragments
template <bool F16> __device__ static inline v16f mfma16(v4i a, v4i b, v16f c)
{
    if constexpr (F16)
        return __builtin_amdgcn_mfma_f32_32x32x16_f16(__builtin_bit_cast(v8h, a), __builtin_bit_cast(v8h, b), c, 0,
                                                      0, 0);
    else
        return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(v8bf, a), __builtin_bit_cast(v8bf, b), c, 0,
                                                       0, 0);
}

// x mod m for an integer-valued float |x| < 2^24: x - m rint(x / m) with
// one rounding of an exact result below 1.5 m (the float quotient is within
// 1 of x / m), so the value stays an exact integer
__device__ static inline float fmod_small(float x, float m, float inv_m)
{
    return __builtin_fmaf(-m, __builtin_rintf(x * inv_m), x);
}

// Window positions of the in-loop reductions of the 16-bit GEMM: in a window
// of W steps x 32 MFMA slots, tile q (MFMAs at slots q and 16 + q of every
// step) has its 16 accumulators reduced two per slot at absolute positions
// 33 q + 18 .. 33 q + 25, i.e. from two slots after its second MFMA of step
// q up to at most slot q - 7 of step q + 1, always before its next MFMA.
// Returns the tile reduced at window position pos, or -1.
__host__ __device__ constexpr int crt16_red_tile(int pos)
{
    return (pos >= 18 && (pos - 18) % 33 < 8 && (pos - 18) / 33 < 16) ? (pos - 18) / 33 : -1;
}
__host__ __device__ constexpr int crt16_red_elem(int pos) { return 2 * ((pos - 18) % 33); }

// The 16-bit counterpart of crt_gemm_pk_kernel: the same block tile (512 A
// rows x 128 B rows, 4 waves of 128 x 128), A register ring, B LDS ring and
// load placement, with v_mfma_f32_32x32x16 (a 64-byte step is 32 K, two
// MFMAs per tile) and FP32 accumulators.  Residues are at most 127 in
// magnitude (moduli <= 255), so an accumulator reduced modulo m
// (|x| <= 1.5 m) takes W more steps (32 W K, at most 1024) and stays below
// 1024 * 127^2 + 383 < 2^24: exact.  The K loop runs in windows of W steps,
// fully unrolled, so that the A ring (NA slots, loads L steps ahead), the B
// LDS ring (DB stages) and the reductions all have compile-time slots: once
// per window each tile is reduced in place (x - m rint(x / m), one fma of an
// exact result), two accumulators per MFMA slot while the tile is idle
// (crt16_red_tile), so that at most two values pass through VGPRs at a time
// and nothing goes to scratch.  KT is a multiple of W.
// SP: the kcount[tb] listed steps only (a multiple of W), as in
// crt_gemm_pk_kernel, in segments of SEG steps (a multiple of W), each a
// full pipeline with its byte offsets staged in LDS.  WB, LA: the block
// shape and the side of the lists, as in crt_gemm_pk_kernel.
template <bool F16, int NA, int DB, int L, int BAR, int W, bool SP, int WB, bool LA>
__global__ __launch_bounds__(256, 1) void crt16_gemm_kernel(const int8_t *__restrict__ A,
                                                            const int8_t *__restrict__ B, int64_t KT, int tiles_a,
                                                            int tiles_b, int64_t M, int64_t N, gemm_moduli mods,
                                                            int nsplit, int64_t pa, int64_t pb,
                                                            int8_t *__restrict__ C, int accum,
                                                            const uint32_t *__restrict__ klist,
                                                            const int *__restrict__ kcount, int64_t sstride)
{
    constexpr int SEG = 85 * W;
    static_assert(W * 32 * 127 * 127 + 383 < (1 << 24), "exact FP32 sums between reductions");
    static_assert(W % NA == 0 && W % DB == 0 && W % BAR == 0, "compile-time ring slots in a window");
    static_assert(NA >= L + 1 && DB >= L + BAR && L >= BAR + 1 && (8 + 2 * WB) * L <= 63 && (WB == 1 || WB == 2),
                  "pipeline shape");
    static_assert(33 * 15 + 25 < 32 * W, "every tile reduced in each window");
    __shared__ __attribute__((aligned(16))) int8_t sB[DB][4 * WB * PK_TILE];
    __shared__ uint32_t sK[SP ? SEG : 1];

    // one launch runs every modulus z (and every part y of a K split): the
    // grid is z-major, then y, then the tiles; consecutive blocks share the
    // listed tile (4 tiles of the other side each): they run the same list
    // in step and are replaced together, so the listed tile is read once per
    // group through L2
    int t = cutc_xcd_remap(blockIdx.x, gridDim.x);
    const int zy = t / (tiles_a * tiles_b), z = zy / nsplit, y = zy - z * nsplit;
    t -= zy * (tiles_a * tiles_b);
    A += (int64_t)z * pa;
    B += (int64_t)z * pb;
    C += (int64_t)z * M * N;
    const int mod = mods.m[z];
    const double inv_mod = mods.inv_m[z];
    int ta, tb;
    if constexpr (LA)
        cutc_group_tile<4>(t, tiles_b, tiles_a, tb, ta);
    else
        cutc_group_tile<4>(t, tiles_a, tiles_b, ta, tb);
    const int lt = LA ? ta : tb;                       // the tile whose list is run
    int64_t NS = SP ? (int64_t)kcount[lt] : KT;        // steps run
    const uint32_t *kl = SP ? klist + (int64_t)lt * KT : nullptr;
    if constexpr (SP) {
        // K split in nsplit parts: part y runs its share of the list, in
        // whole windows of W steps (the list's padding unit), into its own
        // residues sstride bytes on
        if (nsplit > 1) {
            const int64_t g = NS / W;
            const int64_t s0 = g * y / nsplit * W, s1 = g * (y + 1) / nsplit * W;
            kl += s0;
            NS = s1 - s0;
            C += (int64_t)y * sstride;
        }
    }
    int64_t kbase = 0;                                 // first step of the current segment
    uint32_t so_nx = 0;                                // SP: offset of the step L + 1 ahead
    auto off = [&](int64_t s) -> uint32_t {
        if constexpr (SP)
            return sK[s - kbase];
        else
            return (uint32_t)s * PK_TILE;
    };
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int64_t rbs = KT * PK_TILE; // bytes per row block
    const int wr = WB == 1 ? wave : wave >> 1, wc = WB == 1 ? 0 : wave & 1; // A and B quarter / half
    const int8_t *Aw = A + ((int64_t)ta * (16 / WB) + wr * 4) * rbs;
    const int8_t *Bw = B + ((int64_t)tb * 4 * WB + wave * WB) * rbs; // this wave's DMA row blocks
    const uint32_t lo = (uint32_t)lane * 16;
    const float mf = (float)mod, imf = (float)inv_mod;
    constexpr int NLD = 8 + 2 * WB; // loads per step: 2 WB DMA, 8 A
    constexpr int WAIT_NEXT = 8 + NLD * (L - BAR - 1);

    v4i ar[NA][4][2];
    v4i bf0[4], bf1[4];
    v16f acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++)
            acc[i][j] = v16f{};

    // loads of segment step s (the last step's again past k1) into A slot
    // s % NA and LDS stage s % DB
    auto issue = [&](auto s_c, int64_t k0, int64_t k1) {
        constexpr int s = decltype(s_c)::value, sa = s % NA, sb = s % DB;
        const uint32_t so = off(k0 + s < k1 ? k0 + s : k1 - 1);
#pragma unroll
        for (int d = 0; d < WB; d++) {
            pk_dma16(Bw + ((uint32_t)d * (uint32_t)rbs + so + lo), &sB[sb][(wave * WB + d) * PK_TILE]);
            pk_dma16(Bw + ((uint32_t)d * (uint32_t)rbs + so + 1024 + lo), &sB[sb][(wave * WB + d) * PK_TILE + 1024]);
        }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int i = 0; i < 4; i++) {
            ar[sa][i][0] = *(const v4i *)(Aw + ((uint32_t)i * (uint32_t)rbs + so + lo));
            ar[sa][i][1] = *(const v4i *)(Aw + ((uint32_t)i * (uint32_t)rbs + so + 1024 + lo));
        }
        __builtin_amdgcn_sched_barrier(0);
    };

    // the reductions placed at window position pos
    auto reduce_at = [&](auto pos_c) {
        constexpr int pos = decltype(pos_c)::value, q = crt16_red_tile(pos);
        if constexpr (q >= 0) {
            constexpr int r = crt16_red_elem(pos);
            acc[q / 4][q % 4][r] = fmod_small(acc[q / 4][q % 4][r], mf, imf);
            acc[q / 4][q % 4][r + 1] = fmod_small(acc[q / 4][q % 4][r + 1], mf, imf);
        }
    };

    // step u of a window, K step k: during it every wave issues the 2 DMA
    // and 8 A loads of step k + L (the last step's again past the end) and
    // reads the c = 0 fragments of step k + 1
    auto step = [&](auto u_c, int64_t k, int64_t k1) {
        constexpr int u = decltype(u_c)::value, sa = u % NA, sb = u % DB;
        constexpr int la = (u + L) % NA, lb = (u + L) % DB, nb = (u + 1) % DB;
        if constexpr (u % BAR == 0) {
            PK_WAIT_VM(WAIT_NEXT);
            __syncthreads();
        }
#pragma unroll
        for (int j = 0; j < 4; j++)
            bf1[j] = *(const v4i *)&sB[sb][(wc * 4 + j) * PK_TILE + 1024 + lo];
        __builtin_amdgcn_sched_barrier(0);
        uint32_t so;
        if constexpr (SP)
            so = so_nx;
        else
            so = off(k + L < k1 ? k + L : k1 - 1);
        cutc_static_for<0, 16>([&](auto q_c) {
            constexpr int q = decltype(q_c)::value;
            constexpr int i = q / 4, j = q % 4;
            acc[i][j] = mfma16<F16>(ar[sa][i][0], bf0[j], acc[i][j]);
            if constexpr (q < 2 * WB) {
                constexpr int d = q / 2, h = q % 2;
                pk_dma16(Bw + ((uint32_t)d * (uint32_t)rbs + so + h * 1024 + lo),
                         &sB[lb][(wave * WB + d) * PK_TILE + h * 1024]);
            } else if constexpr (q < NLD) {
                constexpr int ii = (q - 2 * WB) / 2, cc = (q - 2 * WB) % 2;
                ar[la][ii][cc] = *(const v4i *)(Aw + ((uint32_t)ii * (uint32_t)rbs + so + cc * 1024 + lo));
            }
            reduce_at(std::integral_constant<int, 32 * u + q>{});
            __builtin_amdgcn_sched_barrier(0);
        });
        cutc_static_for<0, 16>([&](auto q_c) {
            constexpr int q = decltype(q_c)::value;
            constexpr int i = q / 4, j = q % 4;
            acc[i][j] = mfma16<F16>(ar[sa][i][1], bf1[j], acc[i][j]);
            if constexpr (SP && q == 0)
                so_nx = off(k + 1 + L < k1 ? k + 1 + L : k1 - 1);
            if constexpr (q < 4)
                bf0[q] = *(const v4i *)&sB[nb][(wc * 4 + q) * PK_TILE + lo];
            reduce_at(std::integral_constant<int, 32 * u + 16 + q>{});
            __builtin_amdgcn_sched_barrier(0);
        });
    };

    // dense: one segment of KT steps; SP: segments of SEG listed steps
    const int64_t segl = SP ? SEG : NS;
    for (int64_t k0 = 0; k0 < NS; k0 += segl) {
        const int64_t k1 = (NS - k0 < segl) ? NS : k0 + segl;
        if constexpr (SP) {
            // the previous segment's last reads of sK are behind its drain barrier
            kbase = k0;
            for (int i = threadIdx.x; i < k1 - k0; i += 256)
                sK[i] = kl[k0 + i] * PK_TILE;
            __syncthreads();
            so_nx = off(k0 + L < k1 ? k0 + L : k1 - 1);
        }
        // prologue: steps k0 .. k0 + L - 1 in flight; DMA of the first BAR + 1 complete
        cutc_static_for<0, L>([&](auto s_c) { issue(s_c, k0, k1); });
        PK_WAIT_VM(WAIT_NEXT);
        __syncthreads();
#pragma unroll
        for (int j = 0; j < 4; j++)
            bf0[j] = *(const v4i *)&sB[0][(wc * 4 + j) * PK_TILE + lo];

        for (int64_t kw = k0; kw < k1; kw += W)
            cutc_static_for<0, W>([&](auto u_c) { step(u_c, kw + decltype(u_c)::value, k1); });
        PK_WAIT_VM(0); // the repeated loads past the last step
        if constexpr (SP)
            __syncthreads(); // before the next segment's offsets and LDS stages
    }

    // result element r of lane l: A row 8 (r/4) + 4 (l/32) + r%4, B row l%32
    const int li = lane & 31, lg = lane >> 5;
#pragma unroll
    for (int i = 0; i < 4; i++) {
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int64_t col = (int64_t)tb * 128 * WB + wc * 128 + 32 * j + li;
            if (col >= N)
                continue;
#pragma unroll
            for (int r = 0; r < 16; r++) {
                const int64_t row = (int64_t)ta * (512 / WB) + wr * 128 + 32 * i + 8 * (r >> 2) + 4 * lg + (r & 3);
                if (row < M)
                    C[row * N + col] = (int8_t)reduce_mod(
                        (int)fmod_small(acc[i][j][r], mf, imf) + (accum ? (int)C[row * N + col] : 0), mod, inv_mod);
            }
        }
    }
}

// 16-bit GEMM pipeline: A ring of 4 slots, B LDS ring of 6 stages, loads 3
// steps ahead, a barrier every second step, windows of 24 steps (768 K; a
// 16-bit plane's KT = kpad / 32 is a multiple of 24)
#define CRT16_SHAPE 4, 6, 3, 2, 24
// (SP, WB, LA) of the three GEMM forms: dense, lists per B tile, lists per A tile
#define GEMM_FALSE_1_FALSE false, 1, false
#define GEMM_TRUE_2_FALSE true, 2, false
#define GEMM_TRUE_2_TRUE true, 2, true
static_assert(CUTC_PK_KALIGN / 32 % 24 == 0, "16-bit K steps per window");

extern "C" int cutc_crt_gemm_pk(const cutc_crt *c, const int8_t *A, int64_t M, const int8_t *B, int64_t N,
                                int64_t kpad, int8_t *C, int accumulate, hipStream_t st)
{
    return cutc_crt_gemm_pk_list(c, A, M, B, N, kpad, NULL, NULL, 0, 1, C, accumulate, st);
}

extern "C" int cutc_crt_gemm_pk_list(const cutc_crt *c, const int8_t *A, int64_t M, const int8_t *B, int64_t N,
                                     int64_t kpad, const uint32_t *klist, const int *kcount, int list_a,
                                     int nsplit, int8_t *C, int accumulate, hipStream_t st)
{
    if (M <= 0 || N <= 0)
        return 0;
    if (kpad <= 0 || kpad % CUTC_PK_KALIGN || c->kind < CUTC_CRT_I8 || c->kind > CUTC_CRT_FP16 ||
        (!klist) != (!kcount) || nsplit < 1 || nsplit > 65535 || (nsplit > 1 && !klist))
        return CUTC_ERR_ARG;
    const int64_t KT = pk_steps(c->kind, kpad), rb = CUTC_CRT_RBYTES(c->kind);
    if (4 * KT * PK_TILE > 4294967295LL) /* 32-bit offsets inside a wave's four row blocks */
        return CUTC_ERR_ARG;
    const int sp = klist != NULL, la = sp && list_a;
    /* list-driven: 256 x 256 blocks (CUTC_SP_ROWS rows per listed tile); dense: 512 x 128 */
    static_assert(CUTC_SP_ROWS == 256, "list tiles of the 256 x 256 block");
    const int64_t ta = CUTC_PK_ROWS_A(M) / (sp ? 256 : 512), tb = CUTC_PK_ROWS_B(N) / (sp ? 256 : 128);
    if (ta * tb > 2147483647)
        return CUTC_ERR_ARG;
    const int64_t pa = CUTC_PK_ROWS_A(M) * kpad * rb, pb = CUTC_PK_ROWS_B(N) * kpad * rb;
    /* One launch covers every part of a K split and every tile of `per`
     * moduli: all of them, except for the contraction with B-side lists over
     * at least 1024 tiles in int8, or in 16-bit arithmetic as one of several
     * K passes, where a launch per modulus keeps one modulus' residues in L2
     * at a time (int8: 1% faster; the 16-bit K passes of the 40-nucleus
     * build: 2-3% faster).  The 16-bit single pass and the small grids gain
     * 4-5% from the single launch. */
    const int kpass = (accumulate & CUTC_GEMM_KPASS) != 0;
    accumulate &= 1;
    const int per = (sp && !la && ta * tb * nsplit >= 1024 && (c->kind == CUTC_CRT_I8 || kpass)) ? 1 : c->n;
    if (ta * tb * nsplit * per > 2147483647)
        return CUTC_ERR_ARG;
    const int64_t sstride = (int64_t)c->n * M * N; /* bytes between the parts of a K split */
    for (int z0 = 0; z0 < c->n; z0 += per) {
        const int nz = c->n - z0 < per ? c->n - z0 : per;
        const dim3 grid((unsigned)(ta * tb * nsplit * nz));
        gemm_moduli mods;
        for (int i = 0; i < CUTC_CRT_MAX; i++) {
            mods.m[i] = i < nz ? c->m[z0 + i] : 1;
            mods.inv_m[i] = i < nz ? c->inv_m[z0 + i] : 1.0;
        }
        const int8_t *Az = A + z0 * pa, *Bz = B + z0 * pb;
        int8_t *Cz = C + (int64_t)z0 * M * N;
        /* int8: 6-stage rings, loads 4 steps ahead, a barrier every second
         * step; 16-bit: CRT16_SHAPE */
#define GEMM_ARGS                                                                                                      \
    grid, dim3(256), 0, st, Az, Bz, KT, (int)ta, (int)tb, M, N, mods, nsplit, pa, pb, Cz, accumulate, klist, kcount,   \
        sstride
#define GEMM_SP(K8, K16)                                                                                               \
    do {                                                                                                               \
        if (c->kind == CUTC_CRT_I8)                                                                                    \
            hipLaunchKernelGGL((crt_gemm_pk_kernel<6, 4, 2, 2040, K8>), GEMM_ARGS);                                    \
        else if (c->kind == CUTC_CRT_BF16)                                                                             \
            hipLaunchKernelGGL((crt16_gemm_kernel<false, CRT16_SHAPE, K16>), GEMM_ARGS);                               \
        else                                                                                                           \
            hipLaunchKernelGGL((crt16_gemm_kernel<true, CRT16_SHAPE, K16>), GEMM_ARGS);                                \
    } while (0)
        if (la)
            GEMM_SP(GEMM_TRUE_2_TRUE, GEMM_TRUE_2_TRUE);
        else if (sp)
            GEMM_SP(GEMM_TRUE_2_FALSE, GEMM_TRUE_2_FALSE);
        else
            GEMM_SP(GEMM_FALSE_1_FALSE, GEMM_FALSE_1_FALSE);
#undef GEMM_SP
#undef GEMM_ARGS
    }
    return hipGetLastError() == hipSuccess ? 0 : CUTC_ERR_HIP;
}

// ---------------------------------------------------------------- rebuild

// The rebuilt values V of W consecutive elements C[0 .. W) (W-byte aligned;
// the residue planes of the moduli `stride` bytes apart, the parts of a
// K-split contraction `sstride` bytes apart, their residues summed first),
// as doubles (see crt_host.c).  u_i = (c_i y_i) mod m_i in [0, m_i) for the
// summed residue c_i: |c_i y_i| < 2^15 nsplit, so for nsplit < 512 the float
// quotient is within one of the exact one, and the two corrections undo
// that (checked against the integer remainder for every modulus, y and
// residue of one part).  One load of W bytes per modulus and part.
template <int W> struct wbytes;
template <> struct wbytes<1> { typedef uint8_t T; };
template <> struct wbytes<2> { typedef uint16_t T; };
template <> struct wbytes<4> { typedef uint32_t T; };

template <int W>
__device__ static inline void crt_values(const cutc_crt &c, const int8_t *C, int64_t stride, int nsplit,
                                         int64_t sstride, double (&v)[W])
{
    double sh[W], sm[W], sl[W];
#pragma unroll
    for (int q = 0; q < W; q++)
        sh[q] = sm[q] = sl[q] = 0.0;
    for (int i = 0; i < c.n; i++) {
        int r[W];
#pragma unroll
        for (int q = 0; q < W; q++)
            r[q] = 0;
        for (int s = 0; s < nsplit; s++) {
            const uint32_t w = *(const typename wbytes<W>::T *)(C + (int64_t)s * sstride + (int64_t)i * stride);
#pragma unroll
            for (int q = 0; q < W; q++)
                r[q] += (int)(int8_t)(w >> (8 * q));
        }
#pragma unroll
        for (int q = 0; q < W; q++) {
            const int x = r[q] * c.y[i];
            int u = x - c.m[i] * (int)__builtin_floorf((float)x * (float)c.inv_m[i]);
            if (u >= c.m[i])
                u -= c.m[i];
            if (u < 0)
                u += c.m[i];
            const double du = (double)u;
            sh[q] = fma(du, c.Mi[i][0], sh[q]);
            sm[q] = fma(du, c.Mi[i][1], sm[q]);
            sl[q] = fma(du, c.Mi[i][2], sl[q]);
        }
    }
    const double two40 = 1099511627776.0, two80 = 1208925819614629174706176.0;
#pragma unroll
    for (int q = 0; q < W; q++) {
        const double s = (sh[q] * two80 + sm[q] * two40) + sl[q];
        const double t = rint(s * c.Minv);
        const double vh = sh[q] - t * c.M[0], vm = sm[q] - t * c.M[1], vl = sl[q] - t * c.M[2];
        v[q] = (vh * two80 + vm * two40) + vl;
    }
}

// Block -> (row of C, block of 256 W columns) of the rebuild kernels: one
// row's npb blocks are consecutive in each XCD's contiguous range of the
// grid (cutc_xcd_remap), so the scattered stores of a row (its whole output
// line of n_ao^2 doubles) meet in one L2 and leave it as full lines; rows
// past the grid repeat with stride gridDim.x / npb.  Thread: columns
// p .. p + W - 1.
template <int W> __device__ static inline void rebuild_rowcol(int npb, int64_t &row, int64_t &p)
{
    const int t = cutc_xcd_remap(blockIdx.x, gridDim.x);
    row = t / npb;
    p = (int64_t)(t % npb) * 256 * W + W * threadIdx.x;
}

// thread per (p .. p + W - 1, j), P a multiple of W
template <int W>
__global__ __launch_bounds__(256) void rebuild_x_kernel(cutc_crt c, const int8_t *__restrict__ C, int64_t P,
                                                        int64_t N, const int *__restrict__ rexpA,
                                                        const int *__restrict__ rexpB,
                                                        const int *__restrict__ pairs, int n_ao, double *X,
                                                        int64_t n1, int64_t n1c, int64_t i0, int npb)
{
    int64_t j0, p;
    rebuild_rowcol<W>(npb, j0, p);
    if (p >= P)
        return;
    int k[W], l[W], ea[W];
#pragma unroll
    for (int q = 0; q < W; q++) {
        k[q] = pairs[2 * (p + q)];
        l[q] = pairs[2 * (p + q) + 1];
        ea[q] = rexpA[p + q] - 2 * c.bits;
    }
    const int64_t n2 = (int64_t)n_ao * n_ao;
    for (int64_t j = j0; j < N; j += gridDim.x / npb) {
        double V[W];
        crt_values<W>(c, C + j * P + p, P * N, 1, 0, V);
        const int eb = rexpB[j];
        const int64_t m = j / n1c, i1 = j - m * n1c;
        double *o = X + (m * n1 + i0 + i1) * n2;
#pragma unroll
        for (int q = 0; q < W; q++) {
            const double v = ldexp(V[q], ea[q] + eb);
            o[k[q] + (int64_t)n_ao * l[q]] = v;
            o[l[q] + (int64_t)n_ao * k[q]] = v;
        }
    }
}

// the widest W in {4, 2, 1} that divides P with C aligned to it
static int rebuild_width(int64_t P, const int8_t *C)
{
    if (P % 4 == 0 && (uintptr_t)C % 4 == 0)
        return 4;
    if (P % 2 == 0 && (uintptr_t)C % 2 == 0)
        return 2;
    return 1;
}

extern "C" int cutc_crt_rebuild_x(const cutc_crt *c, const int8_t *C, int64_t P, int64_t N,
                                  const int *rexpA, const int *rexpB, const int *pairs, int n_ao, double *X, int64_t n1,
                                  int64_t n1c, int64_t i0, hipStream_t st)
{
    if (P <= 0 || N <= 0)
        return 0;
    const int W = rebuild_width(P, C);
    const int64_t npb = (P + 256 * W - 1) / (256 * W), rows = N < 1048576 / npb + 1 ? N : 1048576 / npb + 1;
    if (npb > 1048576)
        return CUTC_ERR_ARG;
    const dim3 grid((unsigned)(npb * rows));
    if (W == 4)
        hipLaunchKernelGGL(rebuild_x_kernel<4>, grid, dim3(256), 0, st, *c, C, P, N, rexpA, rexpB, pairs, n_ao, X,
                           n1, n1c, i0, (int)npb);
    else if (W == 2)
        hipLaunchKernelGGL(rebuild_x_kernel<2>, grid, dim3(256), 0, st, *c, C, P, N, rexpA, rexpB, pairs, n_ao, X,
                           n1, n1c, i0, (int)npb);
    else
        hipLaunchKernelGGL(rebuild_x_kernel<1>, grid, dim3(256), 0, st, *c, C, P, N, rexpA, rexpB, pairs, n_ao, X,
                           n1, n1c, i0, (int)npb);
    return hipGetLastError() == hipSuccess ? 0 : CUTC_ERR_HIP;
}

// thread per (p .. p + W - 1, q), p fastest, P a multiple of W; the rows q
// of the A side have their own pair table pairsB.
// The stage-2 results of the modular contraction go into M in halves, so
// that every store is coalesced: thread p (X-side pairs in packed order,
// (k, l) = pairs[p], k <= l, so that consecutive p are consecutive rows
// k + n_ao l) and row q of C (H-side pair (a, b) = pairsB[q], a <= b).
// part 0 (X^3 H, symmetric in (a, b)): the S slot, column a + n_ao b;
// part 1 (X^{0..2} NH, antisymmetric in (a, b)): the A slot, column
// b + n_ao a for a < b (for a = b it is zero: NH has a zero diagonal).  Only
// the rows k <= l are written; mode 0 stores, mode 1 adds.
// finalize_m_kernel then forms M[:, ab] = S + A, M[:, ba] = S - A and the
// rows k > l from the rows l < k, which is the full layout the reference's
// GEMMs write (before the symmetrisation M + M^T).
template <int W>
__global__ __launch_bounds__(256) void rebuild_m_kernel(cutc_crt c, const int8_t *__restrict__ C, int64_t P,
                                                        const int *__restrict__ rexpA, const int *__restrict__ rexpB,
                                                        double alpha, const int *__restrict__ pairs,
                                                        const int *__restrict__ pairsB, int n_ao, double *M, int part,
                                                        int mode, int npb, int nsplit, int64_t sstride)
{
    int64_t q0, p;
    rebuild_rowcol<W>(npb, q0, p);
    if (p >= P)
        return;
    int64_t kl[W];
    int ea[W];
#pragma unroll
    for (int e = 0; e < W; e++) {
        kl[e] = pairs[2 * (p + e)] + (int64_t)n_ao * pairs[2 * (p + e) + 1];
        ea[e] = rexpA[p + e] - 2 * c.bits;
    }
    const int64_t n2 = (int64_t)n_ao * n_ao;
    for (int64_t q = q0; q < P; q += gridDim.x / npb) {
        const int a = pairsB[2 * q], b = pairsB[2 * q + 1];
        if (part == 1 && a == b)
            continue;
        double V[W];
        crt_values<W>(c, C + q * P + p, P * P, nsplit, sstride, V);
        const int eb = rexpB[q];
        double *o = M + n2 * (part == 0 ? a + (int64_t)n_ao * b : b + (int64_t)n_ao * a);
#pragma unroll
        for (int e = 0; e < W; e++) {
            const double v = alpha * ldexp(V[e], ea[e] + eb);
            if (mode == 0)
                o[kl[e]] = v;
            else
                o[kl[e]] += v;
        }
    }
}

// one block per column pair (a, b), a <= b (blockIdx.x = the packed pair
// index): the n_ao x n_ao matrices Mc[k][l] = M[k + n_ao l][col] of the
// columns ab and ba in 32 x 32 tiles (i, j), i <= j; the rows k <= l of S
// (column ab) and A (column ba, a < b) give
//   M_ab[k][l] = M_ab[l][k] = S[k][l] + A[k][l],
//   M_ba[k][l] = M_ba[l][k] = S[k][l] - A[k][l],
// the lower tiles written through LDS so that both stores are coalesced.
__global__ __launch_bounds__(256) void finalize_m_kernel(double *M, int n_ao)
{
    __shared__ double ts[32][33], ta[32][33];
    int a, b;
    pair_of(blockIdx.x, a, b);
    const int64_t n2 = (int64_t)n_ao * n_ao;
    double *Mab = M + n2 * (a + (int64_t)n_ao * b), *Mba = M + n2 * (b + (int64_t)n_ao * a);
    const bool anti = a < b;
    const int nt = (n_ao + 31) / 32;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5; // 32 x 8
    for (int ti = 0; ti < nt; ti++)
        for (int tj = ti; tj < nt; tj++) {
            // upper tile rows k in tile ti, columns l in tile tj (k <= l wherever ti < tj)
            __syncthreads();
            for (int yy = ty; yy < 32; yy += 8) {
                const int k = 32 * ti + tx, l = 32 * tj + yy;
                double sv = 0.0, av = 0.0;
                if (k < n_ao && l < n_ao && k <= l) {
                    const int64_t o = k + (int64_t)n_ao * l;
                    sv = Mab[o];
                    av = anti ? Mba[o] : 0.0;
                    Mab[o] = sv + av;
                    if (anti)
                        Mba[o] = sv - av;
                }
                ts[yy][tx] = sv;
                ta[yy][tx] = av;
            }
            __syncthreads();
            // lower tile: rows l (tile tj), columns k (tile ti), k < l
            for (int yy = ty; yy < 32; yy += 8) {
                const int l = 32 * tj + tx, k = 32 * ti + yy;
                if (k < n_ao && l < n_ao && k < l) {
                    const int64_t o = l + (int64_t)n_ao * k;
                    const double sv = ts[tx][yy], av = ta[tx][yy];
                    Mab[o] = sv + av;
                    if (anti)
                        Mba[o] = sv - av;
                }
            }
        }
}

extern "C" int cutc_crt_rebuild_m(const cutc_crt *c, const int8_t *C, int64_t P, const int *rexpA,
                                  const int *rexpB, double alpha, const int *pairs, const int *pairsB, int n_ao,
                                  double *M, int part, int mode, int nsplit, hipStream_t st)
{
    if (P <= 0)
        return 0;
    if ((part != 0 && part != 1) || (mode != 0 && mode != 1) || nsplit < 1 || nsplit > 256)
        return CUTC_ERR_ARG;
    const int W = rebuild_width(P, C);
    const int64_t npb = (P + 256 * W - 1) / (256 * W), rows = P < 1048576 / npb + 1 ? P : 1048576 / npb + 1;
    if (npb > 1048576)
        return CUTC_ERR_ARG;
    const dim3 grid((unsigned)(npb * rows));
    const int64_t sstride = (int64_t)c->n * P * P;
    const int *pb = pairsB ? pairsB : pairs;
    if (W == 4)
        hipLaunchKernelGGL(rebuild_m_kernel<4>, grid, dim3(256), 0, st, *c, C, P, rexpA, rexpB, alpha, pairs, pb, n_ao,
                           M, part, mode, (int)npb, nsplit, sstride);
    else if (W == 2)
        hipLaunchKernelGGL(rebuild_m_kernel<2>, grid, dim3(256), 0, st, *c, C, P, rexpA, rexpB, alpha, pairs, pb, n_ao,
                           M, part, mode, (int)npb, nsplit, sstride);
    else
        hipLaunchKernelGGL(rebuild_m_kernel<1>, grid, dim3(256), 0, st, *c, C, P, rexpA, rexpB, alpha, pairs, pb, n_ao,
                           M, part, mode, (int)npb, nsplit, sstride);
    return hipGetLastError() == hipSuccess ? 0 : CUTC_ERR_HIP;
}

extern "C" int cutc_crt_finalize_m(double *M, int n_ao, hipStream_t st)
{
    const int64_t P = (int64_t)n_ao * (n_ao + 1) / 2;
    if (P <= 0)
        return 0;
    if (P > 2147483647LL)
        return CUTC_ERR_ARG;
    hipLaunchKernelGGL(finalize_m_kernel, dim3((unsigned)P), dim3(256), 0, st, M, n_ao);
    return hipGetLastError() == hipSuccess ? 0 : CUTC_ERR_HIP;
}
