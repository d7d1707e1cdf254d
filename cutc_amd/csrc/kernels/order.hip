// Orderings and K-step lists of the block-sparse stage-1 modular contraction
// (gfx950).
//
// The quantised pair operand w2 chi_k chi_l is mostly exact zeros: a pair of
// localised AOs is non-negligible (relative to its row norm, at the bits of
// the modular scheme) only near its own centres.  Two orderings turn that
// into whole zero tiles of the packed planes, which the GEMM then skips:
//   - r2 points sorted along a Z-order (Morton) curve through their bounding
//     box, so that a pair's support (two balls around its AO centres) is a
//     few contiguous runs of K;
//   - pair rows clustered greedily into tiles by the Hamming distance of
//     their support bitmaps (the K steps where their residues are
//     non-zero), so that the pairs of a B tile share their supports.
// On the 20-nucleus workload (200000 r2 points, 256-pair tiles, 64-point
// steps) the fraction of non-zero (tile, step) tiles is 0.17 this way,
// against 0.25 for points by nearest nucleus and radius with pairs sorted by
// a 16-segment support mask, and 0.105 for single pairs.
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

// ------------------------------------------------------------ point order

// bounding box of the n points r (n x 3): box[0..2] = min, box[3..5] = max
// (one block, fixed-order reductions)
__global__ __launch_bounds__(256) void bbox_kernel(const double *__restrict__ r, int64_t n, double *box)
{
    __shared__ double red[6][256];
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int64_t i = threadIdx.x; i < n; i += 256)
        for (int d = 0; d < 3; d++) {
            lo[d] = fmin(lo[d], r[3 * i + d]);
            hi[d] = fmax(hi[d], r[3 * i + d]);
        }
    for (int d = 0; d < 3; d++) {
        red[d][threadIdx.x] = lo[d];
        red[3 + d][threadIdx.x] = hi[d];
    }
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off)
            for (int d = 0; d < 3; d++) {
                red[d][threadIdx.x] = fmin(red[d][threadIdx.x], red[d][threadIdx.x + off]);
                red[3 + d][threadIdx.x] = fmax(red[3 + d][threadIdx.x], red[3 + d][threadIdx.x + off]);
            }
        __syncthreads();
    }
    if (threadIdx.x < 6)
        box[threadIdx.x] = red[threadIdx.x][0];
}

// the 21 low bits of v spread to every third bit
__device__ static inline uint64_t spread3(uint32_t v)
{
    uint64_t x = v & 0x1fffffu;
    x = (x | x << 32) & 0x1f00000000ffffull;
    x = (x | x << 16) & 0x1f0000ff0000ffull;
    x = (x | x << 8) & 0x100f00f00f00f00full;
    x = (x | x << 4) & 0x10c30c30c30c30c3ull;
    x = (x | x << 2) & 0x1249249249249249ull;
    return x;
}

// key of point i: its Morton (Z-order) code, 21 bits per axis of the
// bounding box (x in bits 3b, y in 3b + 1, z in 3b + 2).  Points close in
// space get close keys, so the support of a pair of localised AOs (two
// balls) becomes a few runs of the ordered points.  Entries [n, nsort) sort
// last.
__global__ __launch_bounds__(256) void morton_key_kernel(const double *__restrict__ r, int64_t n,
                                                         const double *__restrict__ box, int64_t nsort,
                                                         uint64_t *keys, uint32_t *vals)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= nsort)
        return;
    vals[i] = (uint32_t)i;
    if (i >= n) {
        keys[i] = ~(uint64_t)0;
        return;
    }
    uint64_t key = 0;
    for (int d = 0; d < 3; d++) {
        const double ext = box[3 + d] - box[d];
        const double t = ext > 0.0 ? (r[3 * i + d] - box[d]) / ext * 2097152.0 : 0.0;
        const uint32_t q = t <= 0.0 ? 0u : (t >= 2097151.0 ? 2097151u : (uint32_t)t);
        key |= spread3(q) << d;
    }
    keys[i] = key;
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

extern "C" int cutc_order_points(const double *r, int64_t n, double *box, uint64_t *keys, uint32_t *vals,
                                 int64_t nsort, hipStream_t st)
{
    if (n <= 0)
        return 0;
    if (nsort < n || nsort != cutc_sort_len(nsort) || n > 4294967295LL)
        return CUTC_ERR_ARG;
    hipLaunchKernelGGL(bbox_kernel, dim3(1), dim3(256), 0, st, r, n, box);
    hipLaunchKernelGGL(morton_key_kernel, dim3((unsigned)((nsort + 255) / 256)), dim3(256), 0, st, r, n, box, nsort,
                       keys, vals);
    return cutc_sort_u64(keys, vals, nsort, st);
}

extern "C" int cutc_order_r2(const double *r2, const double *wr2, const double *aos2, int64_t n2, int n_ao,
                             double *box, uint64_t *keys, uint32_t *vals, int64_t nsort, double *r2p, double *wr2p,
                             double *aos2p, hipStream_t st)
{
    if (n2 <= 0)
        return 0;
    if (n_ao < 0 || n_ao > 65534)
        return CUTC_ERR_ARG;
    int rc = cutc_order_points(r2, n2, box, keys, vals, nsort, st);
    if (rc)
        return rc;
    hipLaunchKernelGGL(r2_gather_kernel, dim3((unsigned)((n2 + 255) / 256), (unsigned)(n_ao + 1)), dim3(256), 0, st,
                       vals, n2, r2, wr2, aos2, n2, n_ao, r2p, wr2p, aos2p);
    return hipGetLastError() == hipSuccess ? 0 : CUTC_ERR_HIP;
}

// ------------------------------------------------------------ row supports

// Support bitmap of natural pair p = (k, l) (one block per pair): bit s of
// row p (words of 64 steps) is set when a quantised value
// rint(w chi_k chi_l 2^(bits - rexp[p])) of K step s (KS points) is
// non-zero, the values and rounding of quant_pairw_pk_kernel; cnt[p] = the
// number of set bits.  One ballot per wave covers 64 consecutive points.
template <int KS>
__global__ __launch_bounds__(256) void pair_support_kernel(const double *__restrict__ wr2,
                                                           const double *__restrict__ aos2, int64_t n2, int bits,
                                                           const int *__restrict__ rexp, int64_t W, uint64_t *bm,
                                                           int *cnt)
{
    extern __shared__ unsigned long long srow[]; // W words
    __shared__ int scnt;
    const int64_t p = blockIdx.x;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int64_t w = threadIdx.x; w < W; w += 256)
        srow[w] = 0;
    if (threadIdx.x == 0)
        scnt = 0;
    __syncthreads();
    int k, l;
    pair_of(p, k, l);
    const int sh = bits - rexp[p];
    for (int64_t i0 = 0; i0 < n2; i0 += 256) {
        const int64_t i = i0 + threadIdx.x;
        const bool nz = i < n2 && rint(ldexp(wr2[i] * aos2[i + n2 * l] * aos2[i + n2 * k], sh)) != 0.0;
        const uint64_t b = __ballot(nz);
        if (lane == 0 && b) {
            const int64_t s0 = (i0 + 64 * wave) / KS; // first step of the wave's 64 points
            if constexpr (KS == 64) {
                atomicOr(&srow[s0 >> 6], 1ull << (s0 & 63));
            } else {
                if ((uint32_t)b)
                    atomicOr(&srow[s0 >> 6], 1ull << (s0 & 63));
                if (b >> 32)
                    atomicOr(&srow[(s0 + 1) >> 6], 1ull << ((s0 + 1) & 63));
            }
        }
    }
    __syncthreads();
    int c = 0;
    for (int64_t w = threadIdx.x; w < W; w += 256) {
        bm[p * W + w] = srow[w];
        c += __popcll(srow[w]);
    }
    if (c)
        atomicAdd(&scnt, c);
    __syncthreads();
    if (threadIdx.x == 0)
        cnt[p] = scnt;
}

// Support bitmaps of the rows of a one-segment column source (rows = packed
// pairs, K = the chunk's points), the values and rounding of
// quant_colk_pk_kernel: grid (ceil(rows / 64), W), a block covers 64 rows x
// one word (64 steps of KS points), coalesced along the rows; cnt (zeroed
// by the caller) gathers the set bits.
template <int KS>
__global__ __launch_bounds__(256) void colk_support_kernel(cutc_colsrc s, int64_t rows, int bits,
                                                           const int *__restrict__ rexp, int64_t W, uint64_t *bm,
                                                           int *cnt)
{
    __shared__ unsigned long long sw[64];
    const int rl = threadIdx.x & 63, kk = threadIdx.x >> 6;
    const int64_t r = (int64_t)blockIdx.x * 64 + rl, w = blockIdx.y;
    if (threadIdx.x < 64)
        sw[threadIdx.x] = 0;
    __syncthreads();
    if (r < rows) {
        const double *src = s.p[0] + colsrc_col(s, r);
        const int sh = bits - rexp[r];
        const int64_t k0 = w * 64 * KS, k1 = (k0 + 64 * KS < s.K) ? k0 + 64 * KS : s.K;
        uint64_t m = 0;
        for (int64_t k = k0 + kk; k < k1; k += 4)
            if (rint(ldexp(src[colsrc_k(s, k) * s.ld], sh)) != 0.0)
                m |= 1ull << ((k - k0) / KS);
        if (m)
            atomicOr(&sw[rl], m);
    }
    __syncthreads();
    if (kk == 0 && r < rows) {
        bm[r * W + w] = sw[rl];
        const int c = __popcll(sw[rl]);
        if (c)
            atomicAdd(&cnt[r], c);
    }
}

// ------------------------------------------------------------ row clustering

#define CL_THREADS 1024
#define CL_BINS 8192

// exclusive prefix of v over the block (CL_THREADS threads); *total = the sum
__device__ static inline int cl_exscan(int v, int *tmp, int *total)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int x = v;
    for (int off = 1; off < 64; off <<= 1) {
        const int y = __shfl_up(x, off, 64);
        if (lane >= off)
            x += y;
    }
    if (lane == 63)
        tmp[wave] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
        int a = 0;
        for (int i = 0; i < CL_THREADS / 64; i++) {
            const int t = tmp[i];
            tmp[i] = a;
            a += t;
        }
        tmp[CL_THREADS / 64] = a;
    }
    __syncthreads();
    const int ex = tmp[wave] + x - v;
    *total = tmp[CL_THREADS / 64];
    __syncthreads();
    return ex;
}

// Greedy clustering of the P rows into tiles of T rows with similar
// supports (bitmaps bm, W words per row, cnt set bits): repeatedly take the
// remaining row with the largest support (the lowest index on ties) as the
// seed and give its tile the T remaining rows nearest to it in Hamming
// distance (the seed itself at 0; distances clamped to CL_BINS - 1, ties in
// ascending row order), listed in ascending row order; order[0 .. P) is the
// resulting permutation.  Per tile: cl_dist_kernel (all rows, the whole
// device) and cl_select_kernel (one block, which also picks the next seed
// into state[0]).  In the one-block kernels row p is always handled by
// thread p % CL_THREADS.

// the remaining row with the largest cnt, the lowest index on ties, to
// state[0] (one block)
__device__ static inline void cl_pick_seed(const int *__restrict__ cnt, int64_t P, const uint8_t *taken,
                                           unsigned long long *sbest, int *state)
{
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    unsigned long long best = 0;
#pragma unroll 8
    for (int64_t p = tid; p < P; p += CL_THREADS) {
        const int c = cnt[p];
        const unsigned long long key =
            ((unsigned long long)(unsigned)c << 32) | (unsigned long long)(0xffffffffu - (uint32_t)p);
        best = (!taken[p] && key > best) ? key : best;
    }
    for (int off = 32; off > 0; off >>= 1) {
        const unsigned long long o = __shfl_xor(best, off, 64);
        best = o > best ? o : best;
    }
    if (lane == 0)
        sbest[wave] = best;
    __syncthreads();
    if (tid == 0) {
        unsigned long long b = 0;
        for (int i = 0; i < CL_THREADS / 64; i++)
            b = sbest[i] > b ? sbest[i] : b;
        state[0] = (int)(0xffffffffu - (uint32_t)(b & 0xffffffffu));
    }
}

__global__ __launch_bounds__(CL_THREADS) void cl_init_kernel(const int *__restrict__ cnt, int64_t P, uint8_t *taken,
                                                              int *state)
{
    __shared__ unsigned long long sbest[CL_THREADS / 64];
    for (int64_t p = threadIdx.x; p < P; p += CL_THREADS)
        taken[p] = 0;
    cl_pick_seed(cnt, P, taken, sbest, state);
}

__global__ __launch_bounds__(256) void cl_dist_kernel(const uint64_t *__restrict__ bm, int64_t W, int64_t P,
                                                      const uint8_t *__restrict__ taken,
                                                      const int *__restrict__ state, uint32_t *dist)
{
    const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (p >= P || taken[p])
        return;
    const uint64_t *sr = bm + (int64_t)state[0] * W, *pr = bm + p * W;
    int d = 0;
    for (int64_t w = 0; w < W; w++)
        d += __popcll(pr[w] ^ sr[w]);
    dist[p] = (uint32_t)(d < CL_BINS - 1 ? d : CL_BINS - 1);
}

// tile [base, base + need): every remaining row below the threshold
// distance and the first rows at it, then the next seed
__global__ __launch_bounds__(CL_THREADS) void cl_select_kernel(const int *__restrict__ cnt, int64_t P, int64_t base,
                                                                int need, const uint32_t *__restrict__ dist,
                                                                uint8_t *taken, uint32_t *order, int *state)
{
    __shared__ int hist[CL_BINS];
    __shared__ int tmp[CL_THREADS / 64 + 1];
    __shared__ unsigned long long sbest[CL_THREADS / 64];
    __shared__ int s_thr, s_quota;
    const int tid = threadIdx.x;
    for (int i = tid; i < CL_BINS; i += CL_THREADS)
        hist[i] = 0;
    __syncthreads();
#pragma unroll 8
    for (int64_t p = tid; p < P; p += CL_THREADS) {
        const uint32_t d = dist[p];
        if (!taken[p])
            atomicAdd(&hist[d < CL_BINS ? d : CL_BINS - 1], 1);
    }
    __syncthreads();
    // the smallest d* with need rows at distance <= d*
    constexpr int BPT = CL_BINS / CL_THREADS;
    int loc = 0;
    for (int j = 0; j < BPT; j++)
        loc += hist[BPT * tid + j];
    int tot;
    const int ex = cl_exscan(loc, tmp, &tot);
    if (ex < need && need <= ex + loc) {
        int run = ex;
        for (int j = 0; j < BPT; j++) {
            const int h = hist[BPT * tid + j];
            if (run + h >= need) {
                s_thr = BPT * tid + j;
                s_quota = need - run;
                break;
            }
            run += h;
        }
    }
    __syncthreads();
    const uint32_t thr = (uint32_t)s_thr;
    const int quota = s_quota;
    // ordered compaction: thread tid owns the contiguous rows [r0, r1); it
    // counts its rows below the threshold and at it, one scan of each gives
    // its place in the tile and the rank of its first row at the threshold
    // (the first `quota` of those, in row order, are taken)
    const int64_t run = (P + CL_THREADS - 1) / CL_THREADS;
    const int64_t r0 = (int64_t)tid * run < P ? (int64_t)tid * run : P, r1 = r0 + run < P ? r0 + run : P;
    int nlt = 0, neq = 0;
#pragma unroll 8
    for (int64_t p = r0; p < r1; p++) {
        const uint32_t d = dist[p];
        const bool rem = !taken[p];
        nlt += rem && d < thr;
        neq += rem && d == thr;
    }
    int teq, tlt;
    int eqr = cl_exscan(neq, tmp, &teq);
    const int ltr = cl_exscan(nlt, tmp, &tlt);
    // rows placed before this thread's: all below the threshold in earlier
    // runs and those at it among the first quota
    int pos = ltr + (eqr < quota ? eqr : quota);
#pragma unroll 8
    for (int64_t p = r0; p < r1; p++) {
        const uint32_t d = dist[p];
        if (!taken[p]) {
            const int sel = d < thr || (d == thr && eqr < quota);
            eqr += d == thr;
            if (sel) {
                order[base + pos++] = (uint32_t)p;
                taken[p] = 1;
            }
        }
    }
    __syncthreads();
    if (base + need < P)
        cl_pick_seed(cnt, P, taken, sbest, state);
}

// the launches of the clustering (tiles of T rows)
static int cluster_rows(const uint64_t *bm, int64_t W, const int *cnt, int64_t P, int T, uint32_t *dist,
                        uint8_t *taken, uint32_t *order, int *state, hipStream_t st)
{
    hipLaunchKernelGGL(cl_init_kernel, dim3(1), dim3(CL_THREADS), 0, st, cnt, P, taken, state);
    for (int64_t base = 0; base < P; base += T) {
        const int need = (int)((P - base < T) ? P - base : T);
        hipLaunchKernelGGL(cl_dist_kernel, dim3((unsigned)((P + 255) / 256)), dim3(256), 0, st, bm, W, P, taken,
                           state, dist);
        hipLaunchKernelGGL(cl_select_kernel, dim3(1), dim3(CL_THREADS), 0, st, cnt, P, base, need, dist, taken,
                           order, state);
    }
    return hipGetLastError() == hipSuccess ? 0 : CUTC_ERR_HIP;
}

extern "C" size_t cutc_cluster_bytes(int64_t P, int64_t nsteps)
{
    const int64_t W = (nsteps + 63) / 64;
    return (size_t)(8 * P * W + 4 * P + 4 * P + P + 4 * P + 4 + 6 * 256);
}

// scratch regions of cutc_cluster_bytes: bitmaps, counts, distances, flags, order, seed
struct cl_ws {
    uint64_t *bm;
    int *cnt;
    uint32_t *dist;
    uint8_t *taken;
    uint32_t *order;
    int *state;
};

static cl_ws cl_carve(void *ws, int64_t P, int64_t W)
{
    char *b = (char *)ws;
    auto take = [&](size_t n) {
        char *r = b;
        b += (n + 255) / 256 * 256;
        return r;
    };
    cl_ws c;
    c.bm = (uint64_t *)take(8 * P * W);
    c.cnt = (int *)take(4 * P);
    c.dist = (uint32_t *)take(4 * P);
    c.taken = (uint8_t *)take(P);
    c.order = (uint32_t *)take(4 * P);
    c.state = (int *)take(4);
    return c;
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
    if (rexp)
        rexp[s] = rexp_nat[p];
}

static int check_cluster(int64_t P, int64_t nsteps, int kstep)
{
    const int64_t W = (nsteps + 63) / 64;
    if (P <= 0 || P > 2147483647LL || nsteps <= 0 || W > 8192 || (kstep != 32 && kstep != 64))
        return CUTC_ERR_ARG;
    return 0;
}

extern "C" int cutc_order_pairs(const double *wr2, const double *aos2, int64_t n2, int n_ao, int bits,
                                const int *rexp_nat, int kstep, int64_t nsteps, void *ws, int *pairs, int *rexp,
                                hipStream_t st)
{
    const int64_t P = (int64_t)n_ao * (n_ao + 1) / 2;
    if (P <= 0)
        return 0;
    if (check_cluster(P, nsteps, kstep) || (n2 + kstep - 1) / kstep > nsteps)
        return CUTC_ERR_ARG;
    const int64_t W = (nsteps + 63) / 64;
    const cl_ws c = cl_carve(ws, P, W);
    if (kstep == 64)
        hipLaunchKernelGGL(pair_support_kernel<64>, dim3((unsigned)P), dim3(256), 8 * W, st, wr2, aos2, n2, bits,
                           rexp_nat, W, c.bm, c.cnt);
    else
        hipLaunchKernelGGL(pair_support_kernel<32>, dim3((unsigned)P), dim3(256), 8 * W, st, wr2, aos2, n2, bits,
                           rexp_nat, W, c.bm, c.cnt);
    if (cluster_rows(c.bm, W, c.cnt, P, CUTC_SP_ROWS, c.dist, c.taken, c.order, c.state, st))
        return CUTC_ERR_HIP;
    hipLaunchKernelGGL(pair_perm_kernel, dim3((unsigned)((P + 255) / 256)), dim3(256), 0, st, c.order, P, rexp_nat,
                       pairs, rexp);
    return hipGetLastError() == hipSuccess ? 0 : CUTC_ERR_HIP;
}

__global__ __launch_bounds__(256) void i32_zero_kernel(int *p, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n)
        p[i] = 0;
}

extern "C" int cutc_order_pairs_colk(const cutc_colsrc *src, int n_ao, int bits, const int *rexp, int kstep,
                                     int64_t nsteps, void *ws, int *pairs, hipStream_t st)
{
    const int64_t P = (int64_t)n_ao * (n_ao + 1) / 2;
    if (P <= 0)
        return 0;
    if (check_cluster(P, nsteps, kstep) || src->nsrc != 1 || src->K <= 0 || (src->K + kstep - 1) / kstep > nsteps ||
        (P + 63) / 64 > 2147483647LL)
        return CUTC_ERR_ARG;
    const int64_t W = (nsteps + 63) / 64;
    const cl_ws c = cl_carve(ws, P, W);
    hipLaunchKernelGGL(i32_zero_kernel, dim3((unsigned)((P + 255) / 256)), dim3(256), 0, st, c.cnt, P);
    const dim3 grid((unsigned)((P + 63) / 64), (unsigned)W);
    if (kstep == 64)
        hipLaunchKernelGGL(colk_support_kernel<64>, grid, dim3(256), 0, st, *src, P, bits, rexp, W, c.bm, c.cnt);
    else
        hipLaunchKernelGGL(colk_support_kernel<32>, grid, dim3(256), 0, st, *src, P, bits, rexp, W, c.bm, c.cnt);
    if (cluster_rows(c.bm, W, c.cnt, P, CUTC_SP_ROWS, c.dist, c.taken, c.order, c.state, st))
        return CUTC_ERR_HIP;
    hipLaunchKernelGGL(pair_perm_kernel, dim3((unsigned)((P + 255) / 256)), dim3(256), 0, st, c.order, P, nullptr,
                       pairs, nullptr);
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
