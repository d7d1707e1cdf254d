// Device helpers shared by the split-precision kernels (crt.hip, ozaki1.hip).
#pragma once
#include "common.h"
#include "launch.h"

// packed pair index p -> (k, l), k <= l
__device__ static inline void pair_of(int64_t p, int &k, int &l)
{
    l = (int)((sqrt(8.0 * (double)p + 1.0) - 1.0) * 0.5);
    while ((int64_t)(l + 1) * (l + 2) / 2 <= p)
        l++;
    while ((int64_t)l * (l + 1) / 2 > p)
        l--;
    k = (int)(p - (int64_t)l * (l + 1) / 2);
}

// column of row r in a column-major stage-2 source
__device__ static inline int64_t colsrc_col(const cutc_colsrc &s, int64_t r)
{
    if (s.pairs)
        return s.pairs[2 * r] + (int64_t)s.n_ao * s.pairs[2 * r + 1];
    return r;
}

// source row of K index k in a segment of a column-major stage-2 source
__device__ static inline int64_t colsrc_k(const cutc_colsrc &s, int64_t k)
{
    return s.kperm ? (int64_t)s.kperm[k] : k;
}
