// Orderings and K-step lists of the block-sparse stage-1 modular contraction
// (gfx950).
//
// The quantised pair operand w2 chi_k chi_l is mostly exact zeros: a pair of
// localised AOs is non-negligible (relative to its row norm, at the bits of
// the modular scheme) only near its own centres.  Two orderings turn that
// into whole zero tiles of the packed planes, which the GEMM then skips:
//   - r2 points sorted by nearest nucleus, then by distance to it, so that
//     a pair's support is a few contiguous runs of K;
//   - pair rows sorted by the set of r2 segments where their residues are
//     non-zero, so that the pairs of a B tile share their supports.
// A (pair tile, K step) with no non-zero residue contributes 0 to every
// modular sum, so skipping it changes no result bit.  The quantiser flags the
// non-zero tiles (quant_pairw_pk_kernel), klist_kernel compacts the flags
// into per-tile step lists, and the GEMM kernels walk those lists.
#include "split_common.h"

// ------------------------------------------------------------ bitonic sort

// (key, value) pairs, n a power of two >= SORT_BLK.  Every compare-exchange
// swaps only on a strict inequality, so equal keys keep a deterministic order
// (the network is fixed); the callers' keys are unique anyway.
#define SORT_BLK 2048

__device__ static inline void sort_cx(uint64_t *k, uint32_t *v, int i, int j, bool asc)
{
    const uint64_t a = k[i], b = k[j];
    if (a != b && (a > b) == asc) {
        k[i] = b;
        k[j] = a;
        const uint32_t t = v[i];
        v[i] = v[j];
        v[j] = t;
    }
}

// size == 0: sorts each SORT_BLK block (ascending and descending blocks
// alternating, the input of the first global merge); size > 0: the strides
// SORT_BLK / 2 .. 1 of the merge of bitonic runs of that size
__global__ __launch_bounds__(256) void sort_lds_kernel(uint64_t *keys, uint32_t *vals, int64_t size)
{
    __shared__ uint64_t sk[SORT_BLK];
    __shared__ uint32_t sv[SORT_BLK];
    const int64_t base = (int64_t)blockIdx.x * SORT_BLK;
    for (int i = threadIdx.x; i < SORT_BLK; i += 256) {
        sk[i] = keys[base + i];
        sv[i] = vals[base + i];
    }
    __syncthreads();
    auto pass = [&](int64_t sz, int stride) {
        for (int t = threadIdx.x; t < SORT_BLK / 2; t += 256) {
            const int i = (t / stride) * 2 * stride + t % stride;
            sort_cx(sk, sv, i, i + stride, ((base + i) & sz) == 0);
        }
        __syncthreads();
    };
    if (size == 0) {
        for (int sz = 2; sz <= SORT_BLK; sz *= 2)
            for (int stride = sz / 2; stride > 0; stride /= 2)
                pass(sz, stride);
    } else {
        for (int stride = SORT_BLK / 2; stride > 0; stride /= 2)
            pass(size, stride);
    }
    for (int i = threadIdx.x; i < SORT_BLK; i += 256) {
        keys[base + i] = sk[i];
        vals[base + i] = sv[i];
    }
}

// one global stride (>= SORT_BLK) of the merge of runs of the given size
__global__ __launch_bounds__(256) void sort_global_kernel(uint64_t *keys, uint32_t *vals, int64_t n, int64_t size,
                                                          int64_t stride)
{
    const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (t >= n / 2)
        return;
    const int64_t i = (t / stride) * 2 * stride + t % stride, j = i + stride;
    const uint64_t a = keys[i], b = keys[j];
    const bool asc = (i & size) == 0;
    if (a != b && (a > b) == asc) {
        keys[i] = b;
        keys[j] = a;
        const uint32_t v = vals[i];
        vals[i] = vals[j];
        vals[j] = v;
    }
}

extern "C" int64_t cutc_sort_len(int64_t n)
{
    int64_t m = SORT_BLK;
    while (m < n)
        m *= 2;
    return m;
}

extern "C" int cutc_sort_u64(uint64_t *keys, uint32_t *vals, int64_t n, hipStream_t st)
{
    if (n < SORT_BLK || (n & (n - 1)))
        return CUTC_ERR_ARG;
    hipLaunchKernelGGL(sort_lds_kernel, dim3((unsigned)(n / SORT_BLK)), dim3(256), 0, st, keys, vals, (int64_t)0);
    for (int64_t size = 2 * SORT_BLK; size <= n; size *= 2) {
        for (int64_t stride = size / 2; stride >= SORT_BLK; stride /= 2)
            hipLaunchKernelGGL(sort_global_kernel, dim3((unsigned)((n / 2 + 255) / 256)), dim3(256), 0, st, keys,
                               vals, n, size, stride);
        hipLaunchKernelGGL(sort_lds_kernel, dim3((unsigned)(n / SORT_BLK)), dim3(256), 0, st, keys, vals, size);
    }
    return hipGetLastError() == hipSuccess ? 0 : CUTC_ERR_HIP;
}

// ------------------------------------------------------------ r2 order

// key of r2 point i: its nearest nucleus (the lowest index on a tie) in the
// high 32 bits, the float32 pattern of the distance to it in the low 32
// (non-negative floats order as their patterns); with no nuclei, the
// distance to the origin.  Entries [n2, n) sort last.
__global__ __launch_bounds__(256) void r2_key_kernel(const double *__restrict__ r2, int64_t n2,
                                                     const double *__restrict__ rn, int n_nuc, int64_t n,
                                                     uint64_t *keys, uint32_t *vals)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n)
        return;
    vals[i] = (uint32_t)i;
    if (i >= n2) {
        keys[i] = ~(uint64_t)0;
        return;
    }
    const double x = r2[3 * i], y = r2[3 * i + 1], z = r2[3 * i + 2];
    double best = x * x + y * y + z * z;
    int a = 0;
    for (int A = 0; A < n_nuc; A++) {
        const double dx = x - rn[3 * A], dy = y - rn[3 * A + 1], dz = z - rn[3 * A + 2];
        const double d2 = dx * dx + dy * dy + dz * dz;
        if (A == 0 || d2 < best) {
            best = d2;
            a = A;
        }
    }
    keys[i] = ((uint64_t)(uint32_t)a << 32) | __float_as_uint((float)sqrt(best));
}

// r2p[j] = r2[perm[j]] (3 coordinates), wr2p[j] = wr2[perm[j]] (grid row
// y = n_ao) and aos2p[y n2 + j] = aos2[y ld2 + perm[j]] (rows y < n_ao)
__global__ __launch_bounds__(256) void r2_gather_kernel(const uint32_t *__restrict__ perm, int64_t n2,
                                                        const double *__restrict__ r2,
                                                        const double *__restrict__ wr2,
                                                        const double *__restrict__ aos2, int64_t ld2, int n_ao,
                                                        double *r2p, double *wr2p, double *aos2p)
{
    const int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (j >= n2)
        return;
    const int64_t i = perm[j];
    const int y = blockIdx.y;
    if (y < n_ao) {
        aos2p[(int64_t)y * n2 + j] = aos2[(int64_t)y * ld2 + i];
    } else {
        r2p[3 * j] = r2[3 * i];
        r2p[3 * j + 1] = r2[3 * i + 1];
        r2p[3 * j + 2] = r2[3 * i + 2];
        wr2p[j] = wr2[i];
    }
}

extern "C" int cutc_order_points(const double *r, int64_t n, const double *rn, int n_nuc, uint64_t *keys,
                                 uint32_t *vals, int64_t nsort, hipStream_t st)
{
    if (n <= 0)
        return 0;
    if (nsort < n || nsort != cutc_sort_len(nsort) || n > 4294967295LL)
        return CUTC_ERR_ARG;
    hipLaunchKernelGGL(r2_key_kernel, dim3((unsigned)((nsort + 255) / 256)), dim3(256), 0, st, r, n, rn, n_nuc,
                       nsort, keys, vals);
    return cutc_sort_u64(keys, vals, nsort, st);
}

extern "C" int cutc_order_r2(const double *r2, const double *wr2, const double *aos2, int64_t n2, int n_ao,
                             const double *rn, int n_nuc, uint64_t *keys, uint32_t *vals, int64_t nsort,
                             double *r2p, double *wr2p, double *aos2p, hipStream_t st)
{
    if (n2 <= 0)
        return 0;
    if (n_ao < 0 || n_ao > 65534)
        return CUTC_ERR_ARG;
    int rc = cutc_order_points(r2, n2, rn, n_nuc, keys, vals, nsort, st);
    if (rc)
        return rc;
    hipLaunchKernelGGL(r2_gather_kernel, dim3((unsigned)((n2 + 255) / 256), (unsigned)(n_ao + 1)), dim3(256), 0, st,
                       vals, n2, r2, wr2, aos2, n2, n_ao, r2p, wr2p, aos2p);
    return hipGetLastError() == hipSuccess ? 0 : CUTC_ERR_HIP;
}

// ------------------------------------------------------------ pair order

// Support key of natural pair p = (k, l): bit s of the high 32 bits is set
// when a quantised entry rint(w chi_k chi_l 2^(bits - rexp[p])) of r2
// segment s (CUTC_PAIR_SEGS equal runs of the ordered points) is non-zero;
// p itself in the low 32 bits, so the keys are unique.  Entries [P, n) sort
// last.
__global__ __launch_bounds__(256) void pair_key_kernel(const double *__restrict__ wr2,
                                                       const double *__restrict__ aos2, int64_t n2, int bits,
                                                       const int *__restrict__ rexp, int64_t P, uint64_t *keys,
                                                       uint32_t *vals)
{
    __shared__ unsigned smask;
    const int64_t p = blockIdx.x;
    if (threadIdx.x == 0)
        smask = 0;
    __syncthreads();
    if (p >= P) {
        if (threadIdx.x == 0) {
            keys[p] = ~(uint64_t)0;
            vals[p] = (uint32_t)p;
        }
        return;
    }
    int k, l;
    pair_of(p, k, l);
    const int sh = bits - rexp[p];
    unsigned m = 0;
    for (int64_t i = threadIdx.x; i < n2; i += 256) {
        const double q = rint(ldexp(wr2[i] * aos2[i + n2 * l] * aos2[i + n2 * k], sh));
        if (q != 0.0)
            m |= 1u << (int)(i * CUTC_PAIR_SEGS / n2);
    }
    if (m)
        atomicOr(&smask, m);
    __syncthreads();
    if (threadIdx.x == 0) {
        keys[p] = ((uint64_t)smask << 32) | (uint32_t)p;
        vals[p] = (uint32_t)p;
    }
}

// sorted position s <- natural pair order[s]: (k, l) and its exponent
__global__ __launch_bounds__(256) void pair_perm_kernel(const uint32_t *__restrict__ order, int64_t P,
                                                        const int *__restrict__ rexp_nat, int *pairs, int *rexp)
{
    const int64_t s = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (s >= P)
        return;
    const int64_t p = order[s];
    int k, l;
    pair_of(p, k, l);
    pairs[2 * s] = k;
    pairs[2 * s + 1] = l;
    rexp[s] = rexp_nat[p];
}

extern "C" int cutc_order_pairs(const double *wr2, const double *aos2, int64_t n2, int n_ao, int bits,
                                const int *rexp_nat, uint64_t *keys, uint32_t *vals, int64_t nsort, int *pairs,
                                int *rexp, hipStream_t st)
{
    const int64_t P = (int64_t)n_ao * (n_ao + 1) / 2;
    if (P <= 0)
        return 0;
    if (nsort < P || nsort != cutc_sort_len(nsort) || nsort > 2147483647LL)
        return CUTC_ERR_ARG;
    hipLaunchKernelGGL(pair_key_kernel, dim3((unsigned)nsort), dim3(256), 0, st, wr2, aos2, n2, bits, rexp_nat, P,
                       keys, vals);
    int rc = cutc_sort_u64(keys, vals, nsort, st);
    if (rc)
        return rc;
    hipLaunchKernelGGL(pair_perm_kernel, dim3((unsigned)((P + 255) / 256)), dim3(256), 0, st, vals, P, rexp_nat,
                       pairs, rexp);
    return hipGetLastError() == hipSuccess ? 0 : CUTC_ERR_HIP;
}

// Support keys of the rows of a one-segment column source (stage 2: rows =
// packed pairs, K = the chunk's r1 points), as pair_key_kernel: block of 64
// rows x 4 K lanes, coalesced along the rows; rows [rows, n) sort last.
__global__ __launch_bounds__(256) void colk_key_kernel(cutc_colsrc s, int64_t rows, int bits,
                                                       const int *__restrict__ rexp, int64_t n, uint64_t *keys,
                                                       uint32_t *vals)
{
    __shared__ unsigned smask[64];
    const int rl = threadIdx.x & 63, kk = threadIdx.x >> 6;
    const int64_t r = (int64_t)blockIdx.x * 64 + rl;
    if (threadIdx.x < 64)
        smask[threadIdx.x] = 0;
    __syncthreads();
    if (r < rows) {
        const double *src = s.p[0] + colsrc_col(s, r);
        const int sh = bits - rexp[r];
        unsigned m = 0;
        for (int64_t k = kk; k < s.K; k += 4) {
            const double q = rint(ldexp(src[colsrc_k(s, k) * s.ld], sh));
            if (q != 0.0)
                m |= 1u << (int)(k * CUTC_PAIR_SEGS / s.K);
        }
        if (m)
            atomicOr(&smask[rl], m);
    }
    __syncthreads();
    if (kk == 0 && r < n) {
        keys[r] = r < rows ? (((uint64_t)smask[rl] << 32) | (uint32_t)r) : ~(uint64_t)0;
        vals[r] = (uint32_t)r;
    }
}

// sorted position s <- natural pair order[s]: (k, l)
__global__ __launch_bounds__(256) void pair_table_perm_kernel(const uint32_t *__restrict__ order, int64_t P,
                                                              int *pairs)
{
    const int64_t s = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (s >= P)
        return;
    int k, l;
    pair_of(order[s], k, l);
    pairs[2 * s] = k;
    pairs[2 * s + 1] = l;
}

extern "C" int cutc_order_pairs_colk(const cutc_colsrc *src, int n_ao, int bits, const int *rexp, uint64_t *keys,
                                     uint32_t *vals, int64_t nsort, int *pairs, hipStream_t st)
{
    const int64_t P = (int64_t)n_ao * (n_ao + 1) / 2;
    if (P <= 0)
        return 0;
    if (nsort < P || nsort != cutc_sort_len(nsort) || nsort > 2147483647LL || src->nsrc != 1 || src->K <= 0)
        return CUTC_ERR_ARG;
    hipLaunchKernelGGL(colk_key_kernel, dim3((unsigned)(nsort / 64)), dim3(256), 0, st, *src, P, bits, rexp, nsort,
                       keys, vals);
    int rc = cutc_sort_u64(keys, vals, nsort, st);
    if (rc)
        return rc;
    hipLaunchKernelGGL(pair_table_perm_kernel, dim3((unsigned)((P + 255) / 256)), dim3(256), 0, st, vals, P, pairs);
    return hipGetLastError() == hipSuccess ? 0 : CUTC_ERR_HIP;
}

// ------------------------------------------------------------ K-step lists

__global__ __launch_bounds__(256) void u8_zero_kernel(uint8_t *p, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n)
        p[i] = 0;
}

extern "C" int cutc_zero_u8(uint8_t *p, int64_t n, hipStream_t st)
{
    if (n <= 0)
        return 0;
    hipLaunchKernelGGL(u8_zero_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, p, n);
    return hipGetLastError() == hipSuccess ? 0 : CUTC_ERR_HIP;
}

// one block per B tile tb: the steps k with occ[tb][k] set, ascending, then
// copies of the lowest step whose flag is clear up to a multiple of align
// (there is one whenever the count is short of KT, a multiple of align);
// count[tb] = the padded length
__global__ __launch_bounds__(256) void klist_kernel(const uint8_t *__restrict__ occ, int64_t KT, int align,
                                                    uint32_t *list, int *count)
{
    __shared__ int wsum[4];
    __shared__ int base, zstep;
    const int64_t tb = blockIdx.x;
    const uint8_t *o = occ + tb * KT;
    uint32_t *out = list + tb * KT;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (threadIdx.x == 0) {
        base = 0;
        zstep = 0x7fffffff;
    }
    __syncthreads();
    for (int64_t c0 = 0; c0 < KT; c0 += 256) {
        const int64_t k = c0 + threadIdx.x;
        const bool in = k < KT;
        const bool f = in && o[k] != 0;
        if (in && !f)
            atomicMin(&zstep, (int)k);
        const uint64_t b = __ballot(f);
        const int before = __popcll(b & ((lane ? (~0ull >> (64 - lane)) : 0ull)));
        if (lane == 0)
            wsum[wave] = __popcll(b);
        __syncthreads();
        int off = base;
        for (int w = 0; w < wave; w++)
            off += wsum[w];
        if (f)
            out[off + before] = (uint32_t)k;
        __syncthreads();
        if (threadIdx.x == 0)
            base += wsum[0] + wsum[1] + wsum[2] + wsum[3];
        __syncthreads();
    }
    const int n = base, npad = (n + align - 1) / align * align;
    for (int i = n + threadIdx.x; i < npad; i += 256)
        out[i] = (uint32_t)zstep;
    if (threadIdx.x == 0)
        count[tb] = npad;
}

extern "C" int cutc_klists(const uint8_t *occ, int64_t tiles, int64_t KT, int align, uint32_t *list, int *count,
                           hipStream_t st)
{
    if (tiles <= 0)
        return 0;
    if (KT <= 0 || align <= 0 || KT % align || tiles > 2147483647LL || KT > 2147483647LL)
        return CUTC_ERR_ARG;
    hipLaunchKernelGGL(klist_kernel, dim3((unsigned)tiles), dim3(256), 0, st, occ, KT, align, list, count);
    return hipGetLastError() == hipSuccess ? 0 : CUTC_ERR_HIP;
}
