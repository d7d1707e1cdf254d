// Device-side helpers shared by the gfx950 kernels of libcutcint.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#define CUTC_WAVE 64

// f(std::integral_constant<int, i>{}) for i = B .. E - 1, unrolled at compile time
template <int B, int E, typename F>
__device__ __forceinline__ void cutc_static_for(F &&f)
{
    if constexpr (B < E) {
        f(std::integral_constant<int, B>{});
        cutc_static_for<B + 1, E>(f);
    }
}

typedef double v4d __attribute__((ext_vector_type(4)));
typedef double v2d __attribute__((ext_vector_type(2)));

// Packed AO pair index for k <= l: p = l (l + 1) / 2 + k.  For a fixed l the
// pairs k = 0..l are consecutive, and so are the full-matrix entries
// k + n_ao l, which keeps the scatter to (n_ao, n_ao) layouts coalesced.
__host__ __device__ static inline int64_t cutc_pair_index(int k, int l)
{
    return (int64_t)l * (l + 1) / 2 + k;
}

// Launch-side XCD remap: the dispatcher hands workgroup b to XCD (b % 8).
// Give every XCD a contiguous range of tile ids so that the tiles resident on
// one XCD share operand panels in that XCD's L2.
__device__ static inline int cutc_xcd_remap(int b, int nblk)
{
    const int nx = 8;
    int per = nblk / nx, rem = nblk % nx;
    int x = b % nx, r = b / nx;
    return (x < rem) ? x * (per + 1) + r : rem * (per + 1) + (x - rem) * per + r;
}

// Grouped rasterisation: walk GROUP tile-rows down before moving right.
template <int GROUP>
__device__ static inline void cutc_group_tile(int t, int tiles_m, int tiles_n, int &tm, int &tn)
{
    int group_size = GROUP * tiles_n;
    int gid = t / group_size;
    int first_m = gid * GROUP;
    int gm = tiles_m - first_m;
    if (gm > GROUP)
        gm = GROUP;
    int in = t - gid * group_size;
    tm = first_m + in % gm;
    tn = in / gm;
}

// rbar(d) = d / (1 + d) and its gradient wrt the first point,
// (dx,dy,dz) / (d (1 + d)^2); squared distances <= 1e-15 give zeros
// (reference: src/jast_bh.cu:139-156).
__device__ static inline void cutc_rbar_grad(double dx, double dy, double dz, double &f, double &gx, double &gy,
                                             double &gz)
{
    double d2 = dx * dx + dy * dy + dz * dz;
    if (d2 > 1e-15) {
        double d = sqrt(d2);
        double t1 = 1.0 / (1.0 + d);
        double t2 = t1 * t1 / d;
        f = d * t1;
        gx = t2 * dx;
        gy = t2 * dy;
        gz = t2 * dz;
    } else {
        f = gx = gy = gz = 0.0;
    }
}

__device__ static inline double cutc_rbar(double dx, double dy, double dz)
{
    double d2 = dx * dx + dy * dy + dz * dz;
    if (d2 > 1e-15) {
        double d = sqrt(d2);
        return d / (1.0 + d);
    }
    return 0.0;
}

// a^b for the small exponents of the Boys-Handy table, same multiplication
// order as the reference for b <= 6 (src/jast_bh.cu:3-28), repeated products
// above.  b is wave-uniform at every call site.
__device__ static inline double cutc_powi(double a, int b)
{
    double a2;
    switch (b) {
    case 0: return 1.0;
    case 1: return a;
    case 2: return a * a;
    case 3: return a * a * a;
    case 4: a2 = a * a; return a2 * a2;
    case 5: a2 = a * a; return a * a2 * a2;
    case 6: a2 = a * a; return a2 * a2 * a2;
    default: {
        double r = 1.0;
        for (int i = 0; i < b; i++)
            r *= a;
        return r;
    }
    }
}
