// Contracted Cartesian Gaussian AOs and their gradients on a grid (gfx950):
// the aos_data(n_grid, n_ao, 4) input of cutc_int, which CuTC's callers
// evaluate themselves (read by src/long_range_integ.cu:27-39 and
// src/short_range_integ_nonherm.cu:45-55).
//
// One thread per (grid point, AO): block x covers 256 consecutive points,
// block y one AO, so the AO's centre, powers and primitives are uniform in
// the block (scalar loads) and each of the 4 output planes is written as
// one contiguous 2 KB run per block.  HBM-bound: 32 bytes written per
// (point, AO) against 24 read (the point, shared by all AOs through L2).
// Table entries out of range (centre, powers, primitive count) give NaN for
// that AO instead of reading through them.
#include "common.h"
#include "launch.h"

__device__ static inline double ipow_small(double x, int e)
{
    double r = 1.0;
    for (int k = 0; k < e; k++)
        r *= x;
    return r;
}

__global__ __launch_bounds__(256) void aos_kernel(int64_t n, const double *__restrict__ r,
                                                  const double *__restrict__ rn, int n_nuc, int n_ao,
                                                  const int *__restrict__ nucl, const int *__restrict__ power,
                                                  const int *__restrict__ nprim, int pmax,
                                                  const double *__restrict__ expo, const double *__restrict__ coef,
                                                  double *__restrict__ out)
{
    const int64_t g = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int i = blockIdx.y;
    if (g >= n)
        return;
    const int64_t ld = n * n_ao, o = g + n * i;
    const int a = nucl[i] - 1;
    const int ea = power[i], eb = power[i + n_ao], ec = power[i + 2 * n_ao];
    const int np = nprim[i];
    if (a < 0 || a >= n_nuc || ea < 0 || eb < 0 || ec < 0 || np < 0 || np > pmax) {
        // a table entry out of range: NaN for this AO, nothing read through it
        for (int c = 0; c < 4; c++)
            out[o + c * ld] = __builtin_nan("");
        return;
    }
    const double x = r[3 * g] - rn[3 * a], y = r[3 * g + 1] - rn[3 * a + 1], z = r[3 * g + 2] - rn[3 * a + 2];
    const double r2 = x * x + y * y + z * z;
    const double *al = expo + (int64_t)pmax * i, *cf = coef + (int64_t)pmax * i;
    // radial part R = sum_p c_p exp(-alpha_p r2) and D = dR/d(r2) * 2 = sum_p -2 alpha_p c_p exp(-alpha_p r2)
    double R = 0.0, D = 0.0;
    for (int p = 0; p < np; p++) {
        const double e = cf[p] * exp(-al[p] * r2);
        R += e;
        D -= 2.0 * al[p] * e;
    }
    const double px = ipow_small(x, ea), py = ipow_small(y, eb), pz = ipow_small(z, ec);
    // d/dx (x^a R) = a x^(a-1) R + x^(a+1) D
    const double qx = (ea ? ea * ipow_small(x, ea - 1) * R : 0.0) + px * x * D;
    const double qy = (eb ? eb * ipow_small(y, eb - 1) * R : 0.0) + py * y * D;
    const double qz = (ec ? ec * ipow_small(z, ec - 1) * R : 0.0) + pz * z * D;
    out[o] = px * py * pz * R;
    out[o + ld] = qx * py * pz;
    out[o + 2 * ld] = px * qy * pz;
    out[o + 3 * ld] = px * py * qz;
}

extern "C" int cutc_launch_aos(int64_t n, const double *r, const double *rn, int n_nuc, int n_ao, const int *nucl,
                               const int *power, const int *nprim, int pmax, const double *expo, const double *coef,
                               double *out, hipStream_t st)
{
    if (n <= 0 || n_ao <= 0)
        return 0;
    const int64_t nb = (n + 255) / 256;
    if (nb > 2147483647 || n_ao > 65535)
        return CUTC_ERR_ARG;
    hipLaunchKernelGGL(aos_kernel, dim3((unsigned)nb, (unsigned)n_ao), dim3(256), 0, st, n, r, rn, n_nuc, n_ao, nucl,
                       power, nprim, pmax, expo, coef, out);
    return hipGetLastError() == hipSuccess ? 0 : CUTC_ERR_HIP;
}
