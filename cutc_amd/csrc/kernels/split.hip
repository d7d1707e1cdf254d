// Split-precision (Ozaki) contractions -- not built yet.
#include "common.h"
#include "launch.h"

extern "C" int cutc_split_supported(int precision, int slices)
{
    (void)precision;
    (void)slices;
    return 0;
}

extern "C" size_t cutc_split_work_bytes(int precision, int slices, int64_t P, int64_t n1c, int64_t K)
{
    (void)precision;
    (void)slices;
    (void)P;
    (void)n1c;
    (void)K;
    return 0;
}

extern "C" int cutc_stage1_split(int precision, int slices, const double *Apk, int64_t lda, int64_t P, const double *G,
                                 int64_t ldg, int64_t n1c, int64_t K, const int *pairs, int n_ao, double *X, int64_t n1,
                                 int64_t i0, void *work, size_t work_bytes, hipStream_t st, double *ms_split,
                                 int *n_launch)
{
    return CUTC_ERR_UNSUPPORTED;
}

extern "C" size_t cutc_split2_work_bytes(int precision, int slices, int64_t P, int64_t K) { return 0; }

extern "C" int cutc_stage2_split(int precision, int slices, const double *X, int64_t xld, const double *B, int64_t ldb,
                                 int64_t P, int64_t K, double alpha, const int *pairs, int n_ao, double *M, int mode,
                                 void *work, size_t work_bytes, hipStream_t st)
{
    return CUTC_ERR_UNSUPPORTED;
}
