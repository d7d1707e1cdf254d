// FP64 MFMA contractions of the TC integral path for gfx950.
//
// Reference: the six cublasDgemm calls of the path --
//   step 1  X^m(n_ao^2 x n1c) (+)= A^T G^m   src/build_int2_grad1_u12_ao.cu:152-159 (+205,264,315)
//   step 2  M = X^3 H - 1/2 X^{0..2} NH      src/build_int_2e_ao.cu:60-67, 97-104
// Both contractions are symmetric in one AO pair (A and H in (i,j), NH
// antisymmetric), so only packed pairs k <= l (rows) and a <= b (columns of
// step 2) are computed and the epilogues scatter each result to every
// (k,l)/(l,k) x (a,b)/(b,a) position of the reference's full layouts.
//
// Kernel: 128 x 128 block tile, BK = 16, 4 waves (2 x 2) of 64 x 64, each a
// 4 x 4 grid of v_mfma_f64_16x16x4f64.  Operands are staged through LDS with
// a 144-byte row pitch (conflict-free 16-byte reads), double-buffered, one
// barrier per K step.  K is permuted inside a 16-wide step so that every lane
// reads its 4 operand values of a step with two 16-byte LDS reads: MFMA q of
// the step consumes k = 4 g + q from lane group g = lane / 16 (A and B use the
// same permutation, so the sum is unchanged).  Tiles are remapped so that
// each XCD runs a contiguous, 8-row-grouped range of tiles.
#include "common.h"
#include "launch.h"

#define BM 128
#define BN 128
#define BK 16
#define LDSS 18

// ROWK_ANY: K-contiguous rows with any leading dimension and K (8-byte loads,
// k >= Ka read as zero).  ROWK_PAIR: even leading dimension and Ka, 16-byte
// aligned rows: 16-byte loads, pairs at k >= Ka read as zero.
enum { OPA_ROWK = 0, OPA_COLK_GATHER = 1, OPA_ROWK_ANY = 2, OPA_ROWK_PAIR = 3 };
enum { OPB_ROWK = 0, OPB_COLK = 1, OPB_ROWK_ANY = 2, OPB_ROWK_PAIR = 3 };
// EPI_DENSE_T: the block computes C^T (operands exchanged by the caller) and
// stores element (row, col) at out[row ldc + col], so that the 16 lanes of
// an MFMA column write 16 consecutive doubles of C
enum { EPI_X = 0, EPI_X_ADD = 1, EPI_M_STORE = 2, EPI_M_ADD = 3, EPI_M_ADDSUB = 4, EPI_RAW = 5, EPI_DENSE = 6,
       EPI_DENSE_T = 7 };

struct GemmArgs {
    const double *A;
    int64_t lda;
    const double *B;
    int64_t ldb;
    int64_t M, N, K; // K: loop length, a multiple of BK
    int64_t Ka;      // COLK A operand: rows >= Ka are clamped (B is zero there)
    const int *pairs;
    int n_ao;
    double *out;
    int64_t n1c, n1, i0;
    double alpha;
    int tiles_m, tiles_n;
    int64_t kslice; // split K (gridDim.y > 1): block y takes K rows [y kslice, (y + 1) kslice)
    int64_t ldc;    // EPI_DENSE: out[col ldc + row] = alpha acc + beta out[...]
    double beta;
};

// Fetch this thread's 8 operand values of one K step into registers.
// ROWK:   thread t -> row t/2, k = 8 (t%2) .. +7   (K-contiguous rows)
// ROWK_ANY: the same elements one by one, k >= klim read as zero
// ROWK_PAIR: the 16-byte loads of ROWK, pairs at k >= klim read as zero
// COLK:   thread t -> row t%128, k = 8 (t/128) .. +7 (rows of K at stride ld),
//         rows >= klim read row klim - 1
template <int MODE>
__device__ static inline void fetch(const double *__restrict__ P, int64_t ld, int64_t nrows, int64_t klim,
                                    int64_t row0, int64_t k0, int64_t gidx, double (&r)[8])
{
    const int t = threadIdx.x;
    if constexpr (MODE == 0 || MODE == 3) {
        int64_t row = row0 + (t >> 1);
        if (row >= nrows)
            row = nrows - 1;
        const int64_t kb = k0 + 8 * (t & 1);
        const v2d *src = (const v2d *)(P + row * ld + kb);
#pragma unroll
        for (int i = 0; i < 4; i++) {
            v2d v = (MODE == 0 || kb + 2 * i < klim) ? src[i] : v2d{0.0, 0.0};
            r[2 * i] = v[0];
            r[2 * i + 1] = v[1];
        }
    } else if constexpr (MODE == 2) {
        int64_t row = row0 + (t >> 1);
        if (row >= nrows)
            row = nrows - 1;
        const int64_t kb = k0 + 8 * (t & 1);
        const double *src = P + row * ld;
#pragma unroll
        for (int i = 0; i < 8; i++)
            r[i] = (kb + i < klim) ? src[kb + i] : 0.0;
    } else {
        const int64_t kb = k0 + 8 * (t >> 7);
        const double *base = P + gidx;
#pragma unroll
        for (int i = 0; i < 8; i++) {
            const int64_t k = (kb + i < klim) ? kb + i : klim - 1;
            r[i] = base[k * ld];
        }
    }
}

template <int MODE>
__device__ static inline void stash(double *__restrict__ S, const double (&r)[8])
{
    const int t = threadIdx.x;
    double *d = (MODE != 1) ? S + (t >> 1) * LDSS + 8 * (t & 1) : S + (t & 127) * LDSS + 8 * (t >> 7);
#pragma unroll
    for (int i = 0; i < 4; i++)
        *(v2d *)(d + 2 * i) = v2d{r[2 * i], r[2 * i + 1]};
}

template <int AM, int BMODE, int EPI>
__global__ __launch_bounds__(256, 2) void gemm_f64_kernel(GemmArgs g)
{
    __shared__ __attribute__((aligned(16))) double As[2][BM * LDSS];
    __shared__ __attribute__((aligned(16))) double Bs[2][BN * LDSS];

    const int t = cutc_xcd_remap(blockIdx.x, gridDim.x);
    int tm, tn;
    cutc_group_tile<8>(t, g.tiles_m, g.tiles_n, tm, tn);
    const int64_t row0 = (int64_t)tm * BM, col0 = (int64_t)tn * BN;
    if (gridDim.y > 1) {
        // this block's K slice of the (K-row) operands; slices start at
        // multiples of BK below K = roundup(Ka, BK), so Ka - start >= 1
        const int64_t ks = (int64_t)blockIdx.y * g.kslice;
        // K runs along a row of a ROWK operand and down the columns of a COLK one
        g.A += (AM != OPA_COLK_GATHER) ? ks : ks * g.lda;
        g.B += (BMODE != OPB_COLK) ? ks : ks * g.ldb;
        g.K = (g.K - ks < g.kslice) ? g.K - ks : g.kslice;
        g.Ka = (g.Ka - ks < g.kslice) ? g.Ka - ks : g.kslice;
    }

    // per-thread operand row (fixed over the K loop)
    int64_t aidx = 0, bidx = 0;
    if constexpr (AM == OPA_COLK_GATHER) {
        int64_t p = row0 + (threadIdx.x & 127);
        if (p >= g.M)
            p = g.M - 1;
        const int k = g.pairs[2 * p], l = g.pairs[2 * p + 1];
        aidx = k + (int64_t)g.n_ao * l;
    }
    if constexpr (BMODE == OPB_COLK) {
        int64_t j = col0 + (threadIdx.x & 127);
        bidx = (j < g.N) ? j : g.N - 1;
    }

    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wm = (wave >> 1) * 64, wn = (wave & 1) * 64;
    const int li = lane & 15, lg = lane >> 4;

    v4d acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++)
            acc[i][j] = v4d{0.0, 0.0, 0.0, 0.0};

    const int64_t nk = (g.K + BK - 1) / BK;
    double ra[8], rb[8];
    fetch<AM>(g.A, g.lda, g.M, g.Ka, row0, 0, aidx, ra);
    const int64_t kb_lim = (BMODE == OPB_ROWK_ANY || BMODE == OPB_ROWK_PAIR) ? g.Ka : g.K;
    fetch<BMODE>(g.B, g.ldb, g.N, kb_lim, col0, 0, bidx, rb);
    stash<AM>(As[0], ra);
    stash<BMODE>(Bs[0], rb);
    __syncthreads();

    for (int64_t kt = 0; kt < nk; kt++) {
        const int cur = kt & 1;
        const bool more = kt + 1 < nk;
        if (more) {
            fetch<AM>(g.A, g.lda, g.M, g.Ka, row0, (kt + 1) * BK, aidx, ra);
            fetch<BMODE>(g.B, g.ldb, g.N, kb_lim, col0, (kt + 1) * BK, bidx, rb);
        }
        v4d a[4], b[4];
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const double *pa = &As[cur][(wm + 16 * i + li) * LDSS + 4 * lg];
            const double *pb = &Bs[cur][(wn + 16 * i + li) * LDSS + 4 * lg];
            v2d a0 = *(const v2d *)pa, a1 = *(const v2d *)(pa + 2);
            v2d b0 = *(const v2d *)pb, b1 = *(const v2d *)(pb + 2);
            a[i] = v4d{a0[0], a0[1], a1[0], a1[1]};
            b[i] = v4d{b0[0], b0[1], b1[0], b1[1]};
        }
#pragma unroll
        for (int q = 0; q < 4; q++)
#pragma unroll
            for (int i = 0; i < 4; i++)
#pragma unroll
                for (int j = 0; j < 4; j++)
                    acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i][q], b[j][q], acc[i][j], 0, 0, 0);
        if (more) {
            stash<AM>(As[cur ^ 1], ra);
            stash<BMODE>(Bs[cur ^ 1], rb);
        }
        __syncthreads();
    }

    // Epilogue.  v_mfma_f64_16x16x4f64 result layout (measured on gfx950):
    // lane l, element r holds row 4 r + l / 16, column l % 16 of the block.
    const int64_t n2 = (int64_t)g.n_ao * g.n_ao;
#pragma unroll
    for (int i = 0; i < 4; i++) {
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int64_t col = col0 + wn + 16 * j + li;
            if (col >= g.N)
                continue;
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int64_t row = row0 + wm + 16 * i + 4 * r + lg;
                if (row >= g.M)
                    continue;
                if constexpr (EPI == EPI_RAW) {
                    // unscaled partial sum of this K slice: out[y][col][row]
                    g.out[((int64_t)blockIdx.y * g.N + col) * g.M + row] = acc[i][j][r];
                    continue;
                }
                if constexpr (EPI == EPI_DENSE || EPI == EPI_DENSE_T) {
                    double *o = (EPI == EPI_DENSE) ? g.out + col * g.ldc + row : g.out + row * g.ldc + col;
                    *o = (g.beta == 0.0) ? g.alpha * acc[i][j][r] : fma(g.alpha, acc[i][j][r], g.beta * *o);
                    continue;
                }
                const double v = g.alpha * acc[i][j][r];
                const int k = g.pairs[2 * row], l = g.pairs[2 * row + 1];
                const int64_t kl = k + (int64_t)g.n_ao * l, lk = l + (int64_t)g.n_ao * k;
                if constexpr (EPI == EPI_X || EPI == EPI_X_ADD) {
                    const int64_t m = col / g.n1c, i1 = col - m * g.n1c;
                    double *o = g.out + (m * g.n1 + g.i0 + i1) * n2;
                    if constexpr (EPI == EPI_X) {
                        o[kl] = v;
                        o[lk] = v;
                    } else {
                        o[kl] += v;
                        if (k != l)
                            o[lk] += v;
                    }
                } else {
                    const int a = g.pairs[2 * col], b = g.pairs[2 * col + 1];
                    double *o1 = g.out + n2 * (a + (int64_t)g.n_ao * b);
                    double *o2 = g.out + n2 * (b + (int64_t)g.n_ao * a);
                    if constexpr (EPI == EPI_M_STORE) {
                        o1[kl] = v;
                        o1[lk] = v;
                        o2[kl] = v;
                        o2[lk] = v;
                    } else {
                        const double s = (EPI == EPI_M_ADDSUB) ? -v : v;
                        o1[kl] += v;
                        if (k != l)
                            o1[lk] += v;
                        if (a != b) {
                            o2[kl] += s;
                            if (k != l)
                                o2[lk] += s;
                        }
                    }
                }
            }
        }
    }
}

// nsplit > 1: K slices of g.kslice rows (a multiple of BK, set by the
// caller with nsplit = ceil(K / kslice), so every slice starts below K)
template <int AM, int BMODE, int EPI>
static int launch_gemm(const GemmArgs &g0, hipStream_t st, int nsplit = 1)
{
    GemmArgs g = g0;
    if (g.M <= 0 || g.N <= 0)
        return 0;
    int64_t tm = (g.M + BM - 1) / BM, tn = (g.N + BN - 1) / BN;
    if (tm * tn > 2147483647 || nsplit < 1 || nsplit > 65535)
        return CUTC_ERR_ARG;
    if (nsplit == 1)
        g.kslice = g.K;
    else if (g.kslice <= 0 || g.kslice % BK || (nsplit - 1) * g.kslice >= g.K || nsplit * g.kslice < g.K)
        return CUTC_ERR_ARG;
    if (AM == OPA_COLK_GATHER && !g.pairs)
        return CUTC_ERR_ARG;
    g.tiles_m = (int)tm;
    g.tiles_n = (int)tn;
    hipLaunchKernelGGL((gemm_f64_kernel<AM, BMODE, EPI>), dim3((unsigned)(tm * tn), (unsigned)nsplit), dim3(256), 0,
                       st, g);
    return hipGetLastError() == hipSuccess ? 0 : CUTC_ERR_HIP;
}

extern "C" int cutc_gemm1_f64(const double *Apk, int64_t lda, const double *G, int64_t ldg, int64_t P, int64_t n1c,
                              int64_t K, const int *pairs, int n_ao, double *X, int64_t n1, int64_t i0,
                              int accumulate, hipStream_t st)
{
    if (K % BK != 0 || lda % 2 != 0 || ldg % 2 != 0)
        return CUTC_ERR_ARG;
    GemmArgs g = {};
    g.A = Apk;
    g.lda = lda;
    g.B = G;
    g.ldb = ldg;
    g.M = P;
    g.N = 4 * n1c;
    g.K = K;
    g.Ka = K;
    g.pairs = pairs;
    g.n_ao = n_ao;
    g.out = X;
    g.n1c = n1c;
    g.n1 = n1;
    g.i0 = i0;
    g.alpha = 1.0;
    if (accumulate)
        return launch_gemm<OPA_ROWK, OPB_ROWK, EPI_X_ADD>(g, st);
    return launch_gemm<OPA_ROWK, OPB_ROWK, EPI_X>(g, st);
}

extern "C" int cutc_gemm2_f64(const double *X, int64_t xld, const double *B, int64_t ldb, int64_t P, int64_t K,
                              double alpha, const int *pairs, int n_ao, double *M, int mode, hipStream_t st)
{
    GemmArgs g = {};
    g.A = X;
    g.lda = xld;
    g.B = B;
    g.ldb = ldb;
    g.M = P;
    g.N = P;
    g.K = (K + BK - 1) / BK * BK;
    g.Ka = K;
    g.pairs = pairs;
    g.n_ao = n_ao;
    g.out = M;
    g.alpha = alpha;
    if (K <= 0) {
        return (mode == 0) ? CUTC_ERR_ARG : 0;
    }
    switch (mode) {
    case 0: return launch_gemm<OPA_COLK_GATHER, OPB_COLK, EPI_M_STORE>(g, st);
    case 1: return launch_gemm<OPA_COLK_GATHER, OPB_COLK, EPI_M_ADD>(g, st);
    case 2: return launch_gemm<OPA_COLK_GATHER, OPB_COLK, EPI_M_ADDSUB>(g, st);
    }
    return CUTC_ERR_ARG;
}

/* Split-K form of the step-2 contraction: the K rows run in nsplit slices
 * (blocks y = 0 .. nsplit - 1), each writing its unscaled partial sums
 * work[y][q][p] (q: B row, p: A row); cutc_scatter_m adds the slices in
 * order and scatters them.  nsplit is reduced so that every slice has rows. */
extern "C" int cutc_gemm2_f64_split(const double *X, int64_t xld, const double *B, int64_t ldb, int64_t P, int64_t K,
                                    const int *pairs, int n_ao, int nsplit, double *work, hipStream_t st)
{
    if (K <= 0 || nsplit < 1 || !pairs)
        return CUTC_ERR_ARG;
    GemmArgs g = {};
    g.A = X;
    g.lda = xld;
    g.B = B;
    g.ldb = ldb;
    g.M = P;
    g.N = P;
    g.K = (K + BK - 1) / BK * BK;
    g.Ka = K;
    g.pairs = pairs; /* the X operand is gathered through the pair table */
    g.n_ao = n_ao;
    g.out = work;
    g.alpha = 1.0;
    g.kslice = ((g.K + nsplit - 1) / nsplit + BK - 1) / BK * BK;
    const int ns = (int)((g.K + g.kslice - 1) / g.kslice);
    return launch_gemm<OPA_COLK_GATHER, OPB_COLK, EPI_RAW>(g, st, ns);
}

extern "C" int cutc_gemm2_split_count(int64_t K, int nsplit)
{
    if (K <= 0 || nsplit < 1)
        return 1;
    const int64_t Kp = (K + BK - 1) / BK * BK;
    const int64_t kslice = ((Kp + nsplit - 1) / nsplit + BK - 1) / BK * BK;
    return (int)((Kp + kslice - 1) / kslice);
}

/* Dense C (M x N, column-major, ld ldc) = alpha A^T B + beta C for
 * K-contiguous operands: A row i at A + i lda, B row j at B + j ldb.
 * Operand loads: K a multiple of BK with even leading dimensions and 16-byte
 * aligned operands reads whole 16-wide K steps; an even K with the same
 * alignment reads 16-byte pairs up to K; anything else reads 8-byte elements
 * up to K.  nsplit == 1 runs the block on C^T (operands exchanged) so that
 * the stores of C run along its columns.  nsplit > 1: the K rows run in
 * slices whose unscaled partial sums go to work[y][j][i] instead (alpha, beta
 * and C are then unused); the caller adds the slices. */
extern "C" int cutc_gemm_tn_f64(const double *A, int64_t lda, const double *B, int64_t ldb, int64_t M, int64_t N,
                                int64_t K, double alpha, double beta, double *C, int64_t ldc, int nsplit,
                                double *work, hipStream_t st)
{
    if (M <= 0 || N <= 0)
        return 0;
    if (K <= 0 || lda < K || ldb < K || nsplit < 1 || (nsplit == 1 && ldc < M) || (nsplit > 1 && !work))
        return CUTC_ERR_ARG;
    const bool wide = lda % 2 == 0 && ldb % 2 == 0 && ((uintptr_t)A % 16) == 0 && ((uintptr_t)B % 16) == 0;
    const int form = (wide && K % BK == 0) ? 0 : (wide && K % 2 == 0) ? 1 : 2;
    GemmArgs g = {};
    g.K = (K + BK - 1) / BK * BK;
    g.Ka = K;
    g.alpha = alpha;
    g.beta = beta;
    g.ldc = ldc;
    if (nsplit == 1) {
        g.A = B; // the block computes C^T = B^T A
        g.lda = ldb;
        g.B = A;
        g.ldb = lda;
        g.M = N;
        g.N = M;
        g.out = C;
        switch (form) {
        case 0: return launch_gemm<OPA_ROWK, OPB_ROWK, EPI_DENSE_T>(g, st);
        case 1: return launch_gemm<OPA_ROWK_PAIR, OPB_ROWK_PAIR, EPI_DENSE_T>(g, st);
        default: return launch_gemm<OPA_ROWK_ANY, OPB_ROWK_ANY, EPI_DENSE_T>(g, st);
        }
    }
    g.A = A;
    g.lda = lda;
    g.B = B;
    g.ldb = ldb;
    g.M = M;
    g.N = N;
    g.out = work;
    g.kslice = ((g.K + nsplit - 1) / nsplit + BK - 1) / BK * BK;
    const int ns = (int)((g.K + g.kslice - 1) / g.kslice);
    switch (form) {
    case 0: return launch_gemm<OPA_ROWK, OPB_ROWK, EPI_RAW>(g, st, ns);
    case 1: return launch_gemm<OPA_ROWK_PAIR, OPB_ROWK_PAIR, EPI_RAW>(g, st, ns);
    default: return launch_gemm<OPA_ROWK_ANY, OPB_ROWK_ANY, EPI_RAW>(g, st, ns);
    }
}

extern "C" int cutc_gemm_tn_split_count(int64_t K, int nsplit)
{
    return cutc_gemm2_split_count(K, nsplit);
}
