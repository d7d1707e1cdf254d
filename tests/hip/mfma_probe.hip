// One MFMA per call on known operands, raw per-lane results out: lets the
// tests check the lane layouts the contraction kernels assume on gfx950.
// in:  a[64 * na], b[64 * nb] per-lane operand words as the builtin takes them
// out: d[64 * nd] per-lane result elements (lane-major)
#include "common.h"
#include "launch.h"

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));
typedef float v16f __attribute__((ext_vector_type(16)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef _Float16 v8h __attribute__((ext_vector_type(8)));

__global__ void probe_f64(const double *a, const double *b, double *d)
{
    const int l = threadIdx.x;
    v4d c = {0.0, 0.0, 0.0, 0.0};
    c = __builtin_amdgcn_mfma_f64_16x16x4f64(a[l], b[l], c, 0, 0, 0);
    for (int r = 0; r < 4; r++)
        d[4 * l + r] = c[r];
}

__global__ void probe_i8(const v4i *a, const v4i *b, int *d)
{
    const int l = threadIdx.x;
    v16i c = {};
    c = __builtin_amdgcn_mfma_i32_32x32x32_i8(a[l], b[l], c, 0, 0, 0);
    for (int r = 0; r < 16; r++)
        d[16 * l + r] = c[r];
}

__global__ void probe_bf16(const v8bf *a, const v8bf *b, float *d)
{
    const int l = threadIdx.x;
    v16f c = {};
    c = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[l], b[l], c, 0, 0, 0);
    for (int r = 0; r < 16; r++)
        d[16 * l + r] = c[r];
}

__global__ void probe_f16(const v8h *a, const v8h *b, float *d)
{
    const int l = threadIdx.x;
    v16f c = {};
    c = __builtin_amdgcn_mfma_f32_32x32x16_f16(a[l], b[l], c, 0, 0, 0);
    for (int r = 0; r < 16; r++)
        d[16 * l + r] = c[r];
}

// kind: 0 f64 16x16x4, 1 i8 32x32x32, 2 bf16 32x32x16, 3 f16 32x32x16.
// Buffers are device pointers sized for one wave (64 lanes).
extern "C" int cutc_mfma_probe(int kind, const void *a, const void *b, void *d)
{
    switch (kind) {
    case 0: hipLaunchKernelGGL(probe_f64, dim3(1), dim3(64), 0, 0, (const double *)a, (const double *)b, (double *)d); break;
    case 1: hipLaunchKernelGGL(probe_i8, dim3(1), dim3(64), 0, 0, (const v4i *)a, (const v4i *)b, (int *)d); break;
    case 2: hipLaunchKernelGGL(probe_bf16, dim3(1), dim3(64), 0, 0, (const v8bf *)a, (const v8bf *)b, (float *)d); break;
    case 3: hipLaunchKernelGGL(probe_f16, dim3(1), dim3(64), 0, 0, (const v8h *)a, (const v8h *)b, (float *)d); break;
    default: return CUTC_ERR_ARG;
    }
    if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess)
        return CUTC_ERR_HIP;
    return 0;
}
