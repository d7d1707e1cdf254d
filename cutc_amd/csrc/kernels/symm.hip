// M <- M + M^T for an n x n column-major FP64 matrix, in place (gfx950).
//
// Reference: src/add_trans_inplace.cu:2-21, where thread (i, j <= i) updates
// one pair through strided global accesses and half the threads idle.  Here
// each block owns a pair of 32 x 32 tiles (I, J) and (J, I), I <= J, staged
// through LDS so that every global read and write is a coalesced 256-byte
// column segment; HBM-bound, 16 n^2 bytes moved in total.  The small tiles
// (17 KB of LDS per block) let many blocks share a CU, so that enough loads
// are in flight to cover the memory latency.
#include "common.h"
#include "launch.h"

#define ST 32
#define SR (256 / ST) // tile columns handled per pass of the block

__global__ __launch_bounds__(256) void trans_inplace_kernel(double *M, int64_t n)
{
    const int I = blockIdx.x, J = blockIdx.y;
    if (I > J)
        return;
    __shared__ double a[ST][ST + 1]; // a[c][r] = M[I*ST + r, J*ST + c]
    __shared__ double b[ST][ST + 1]; // b[c][r] = M[J*ST + r, I*ST + c]
    const int tx = threadIdx.x % ST, ty = threadIdx.x / ST;
    const int64_t i0 = (int64_t)I * ST, j0 = (int64_t)J * ST;
    for (int c = ty; c < ST; c += SR) {
        const int64_t gi = i0 + tx, gj = j0 + c;
        a[c][tx] = (gi < n && gj < n) ? M[gi + n * gj] : 0.0;
        if (I != J) {
            const int64_t hi = j0 + tx, hj = i0 + c;
            b[c][tx] = (hi < n && hj < n) ? M[hi + n * hj] : 0.0;
        }
    }
    __syncthreads();
    if (I == J) {
        for (int c = ty; c < ST; c += SR) {
            const int64_t gi = i0 + tx, gj = j0 + c;
            if (gi < n && gj < n)
                M[gi + n * gj] = a[c][tx] + a[tx][c];
        }
        return;
    }
    for (int c = ty; c < ST; c += SR) {
        // M[I*ST+tx, J*ST+c] += M[J*ST+c, I*ST+tx] = b[tx][c]
        const int64_t gi = i0 + tx, gj = j0 + c;
        if (gi < n && gj < n)
            M[gi + n * gj] = a[c][tx] + b[tx][c];
        // M[J*ST+tx, I*ST+c] += M[I*ST+c, J*ST+tx] = a[tx][c]
        const int64_t hi = j0 + tx, hj = i0 + c;
        if (hi < n && hj < n)
            M[hi + n * hj] = b[c][tx] + a[tx][c];
    }
}

extern "C" int cutc_launch_trans_inplace(double *M, int64_t n, hipStream_t st)
{
    if (n <= 0)
        return 0;
    int64_t nt = (n + ST - 1) / ST;
    if (nt > 65535)
        return CUTC_ERR_UNSUPPORTED;
    hipLaunchKernelGGL(trans_inplace_kernel, dim3((unsigned)nt, (unsigned)nt), dim3(256), 0, st, M, n);
    return hipGetLastError() == hipSuccess ? 0 : CUTC_ERR_HIP;
}
