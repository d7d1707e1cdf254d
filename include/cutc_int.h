/*
 * libcutcint -- MI355X (gfx950) transcorrelated-integral path, C ABI.
 *
 * Drop-in for the reference library's cutc_int path.  Every symbol in the
 * first section keeps the reference's name, argument list, argument meaning,
 * array layout and return convention; each cites the reference definition it
 * replaces.  Self-contained: needs no CUDA or HIP header.
 *
 * Array layouts are Fortran (column-major), as in the reference's binding
 * src/cutc_module.f90:23-35:
 *   r1(3,n_grid1)  wr1(n_grid1)  r2(3,n_grid2)  wr2(n_grid2)  rn(3,n_nuc)
 *   aos_data1(n_grid1,n_ao,4)  aos_data2(n_grid2,n_ao,4)  (value, d/dx, d/dy, d/dz)
 *   c_bh(size_bh,n_nuc) double,  m_bh/n_bh/o_bh(size_bh,n_nuc) int
 *   int2_grad1_u12_ao(n_ao,n_ao,n_grid1,3) on the host, (n_ao,n_ao,n_grid1,4) on the device
 *   int_2e_ao(n_ao,n_ao,n_ao,n_ao)
 */
#ifndef CUTC_INT_H
#define CUTC_INT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Launch geometry passed by value; same layout as CUDA's/HIP's dim3
 * (three 32-bit unsigned integers). */
typedef struct cutc_dim3 {
    unsigned int x, y, z;
} cutc_dim3;

/* Return codes.  0 and -1 are the reference's (src/cutc_int.c:79-86,242);
 * the others replace its exit(0) on runtime errors (src/utils.cu:12,50). */
#define CUTC_OK 0
#define CUTC_ERR_GEOMETRY (-1) /* launch geometry exceeds the device limits   */
#define CUTC_ERR_ARG (-2)      /* negative size, NULL pointer, bad exponent    */
#define CUTC_ERR_NOMEM (-3)    /* the working set does not fit in device memory */
#define CUTC_ERR_HIP (-4)      /* HIP runtime error                             */
#define CUTC_ERR_UNSUPPORTED (-5)

/* ---------------------------------------------------------------------
 * Reference interface
 * ------------------------------------------------------------------- */

/* Full path on host arrays: uploads inputs, builds int2_grad1_u12_ao (all 4
 * components on the device) and int_2e_ao, downloads components 0..2 of
 * int2_grad1_u12_ao (n_ao*n_ao*n_grid1*3 doubles) and int_2e_ao (n_ao^4).
 * Returns 0, or -1 when blockx*blocky*blockz exceeds the device's
 * maxThreadsPerBlock or a grid dimension exceeds maxGridSize, or another
 * negative CUTC_ERR_* code.  Results do not depend on the geometry.
 * Replaces: cutc_int, src/cutc_int.c:15-243 (Fortran binding
 * src/cutc_module.f90:10-37). */
int cutc_int(int nxBlocks, int nyBlocks, int nzBlocks, int blockxSize, int blockySize, int blockzSize, int n_grid1,
             int n_grid2, int n_ao, int n_nuc, int size_bh, double *h_r1, double *h_wr1, double *h_r2, double *h_wr2,
             double *h_rn, double *h_aos_data1, double *h_aos_data2, double *h_c_bh, int *h_m_bh, int *h_n_bh,
             int *h_o_bh, double *h_int2_grad1_u12_ao, double *h_int_2e_ao);

/* Stage 1 on device pointers: int2_grad1_u12_ao(n_ao,n_ao,n_grid1,4), with
 * X^m_{kl}(r1) = sum_r2 wr2 chi_k chi_l G^m(r1,r2).  Void, as in the
 * reference; the status is in cutc_get_last_error().
 * Replaces: get_int2_grad1_u12_ao, include/cutc_int.h:8-11,
 * src/build_int2_grad1_u12_ao.cu:9-347. */
void get_int2_grad1_u12_ao(cutc_dim3 dimGrid, cutc_dim3 dimBlock, int n_grid1, int n_grid2, int n_ao, int n_nuc,
                           int size_bh, double *r1, double *r2, double *wr2, double *rn, double *aos_data2,
                           double *c_bh, int *m_bh, int *n_bh, int *o_bh, double *int2_grad1_u12_ao);

/* Stage 2 on device pointers: tc_int_2e_ao = M + M^T with
 * M = X^3 H - 1/2 X^{0..2} NH (reads all 4 components of
 * int2_grad1_u12_ao).  Status in cutc_get_last_error().
 * Replaces: get_int_2e_ao, include/cutc_int.h:13-14,
 * src/build_int_2e_ao.cu:11-145. */
void get_int_2e_ao(int n_grid1, int n_ao, double *wr1, double *aos_data1, double *int2_grad1_u12_ao,
                   double *tc_int_2e_ao);

/* Error helpers of the reference header: print the message and terminate
 * the process when err != 0 (the reference exits with status 0; these exit
 * with status 1).  Replace: include/cutc_int.h:4-5, src/utils.cu:8-14,43-51. */
void checkCudaErrors(int err, const char *msg, const char *file, int line);
void checkCublasErrors(int status, const char *msg, const char *file, int line);

/* Launch wrappers the reference library exports (not in its header), on
 * device pointers, with the reference's output layouts.  Geometry arguments
 * are accepted and ignored (results never depended on them).
 * Replace: src/jast_bh.cu:268-303, src/long_range_integ.cu:52-66,
 * src/short_range_integ_herm.cu:51-63, src/short_range_integ_nonherm.cu:69-81,
 * src/add_trans_inplace.cu:24-34. */
void cutc_int_bh(cutc_dim3 dimGrid, cutc_dim3 dimBlock, int ii0, int n_grid1_eff, int n_grid1_tot, int jj0,
                 int n_grid2_eff, int n_grid2_tot, int n_nuc, int size_bh, double *r1, double *r2, double *rn,
                 double *c_bh, int *m_bh, int *n_bh, int *o_bh, double *grad1_u12);
void int_long_range(int jj0, int n_grid2_eff, int n_grid2_tot, int n_grid2, int n_ao, double *wr2, double *aos_data2,
                    double *int_fct_long_range);
void int_short_range_herm(int n_grid1, int n_ao, double *wr1, double *aos_data1, double *int_fct_short_range_herm);
void int_short_range_nonherm(int n_grid1, int n_ao, double *wr1, double *aos_data1,
                             double *int_fct_short_range_nonherm);
void trans_inplace(double *data, int size);

/* ---------------------------------------------------------------------
 * Additions (no counterpart in the reference)
 * ------------------------------------------------------------------- */

/* Status of the last call made by this host thread (0 or CUTC_ERR_*). */
int cutc_get_last_error(void);
const char *cutc_error_string(int code);

/* Arithmetic of the two dense contractions. */
#define CUTC_PREC_FP64 0 /* FP64 MFMA                                                 */
#define CUTC_PREC_BF16 1 /* Ozaki slices in bf16, FP32 MFMA accumulate -> FP64         */
#define CUTC_PREC_FP16 2 /* Ozaki slices in fp16, FP32 MFMA accumulate -> FP64         */
#define CUTC_PREC_INT8 3 /* operands rounded to integers, products modulo `slices`
                            moduli on int8 MFMA (exact int32), CRT rebuild in FP64  */
#define CUTC_PREC_BF16_CRT 4 /* as INT8 with moduli <= 255, residues as bf16 values,
                                FP32 MFMA accumulate kept exact, CRT rebuild in FP64 */
#define CUTC_PREC_FP16_CRT 5 /* the same with fp16 residues                          */

typedef struct cutc_options {
    int precision;    /* CUTC_PREC_*                                             */
    int slices;       /* bf16/fp16: slices per operand (2..4, 0 = 4); int8: moduli
                         (6..16, 0 = 11); bf16/fp16 CRT: moduli (6..16, 0 = 10) */
    int verbose;      /* 1: print the reference's stage-time lines               */
    int n1_chunk;     /* r1 rows per G chunk, 0 = planner                        */
    size_t mem_limit; /* device bytes the workspaces may use, 0 = free memory    */
    int host_copy;    /* host entry points (cutc_int, cutc_int_ex): 0 = download
                         each r1 chunk of X while the build runs (pinned staging,
                         copy stream, host copy threads); 1 = blocking copies
                         after the build, as the reference                      */
    int streams;      /* cutc_int_batch_device: concurrent branches the systems
                         are recorded on (1..4; 0 = 4 for the modular
                         precisions, 1 for the others; fewer when the device
                         cannot hold a workspace and chunk buffer per branch) */
    int dense;        /* modular precisions: 1 = the stage-1 contraction runs
                         every K step; 0 = it skips the (256-pair, K-step)
                         tiles whose residues are all zero.  Same results, bit
                         for bit; 1 is for checking and timing. */
} cutc_options;

typedef struct cutc_stats {
    double ms_total;       /* whole call, stream-ordered hipEvents      */
    double ms_jastrow;     /* Jastrow pair kernels                      */
    double ms_integrands;  /* pair-weight and r1 integrand kernels      */
    double ms_gemm1;       /* step-1 contraction kernels                */
    double ms_gemm2;       /* step-2 contraction kernels                */
    double ms_symm;        /* M + M^T                                   */
    double ms_split;       /* slicing kernels of the split modes        */
    int n_gemm1;           /* step-1 contraction calls (one per r1 chunk
                              and K pass; modular: every modulus, or for
                              int8 chunks of >= 1024 tiles one launch per
                              modulus)                                   */
    int n_gemm2;           /* step-2 contraction launches               */
    int n_jastrow;         /* Jastrow launches                          */
    int64_t n1_chunk;      /* r1 rows per chunk used                    */
    int n1_chunks;         /* number of r1 chunks                       */
    double flops_exec;     /* FP64-equivalent multiply-adds x 2 executed by the contractions */
    size_t workspace_bytes;/* device workspace allocated by the call    */
    double ms_host_in;     /* host entry points: device allocation and upload before the build (wall clock) */
    double ms_host_out;    /* host entry points: end of the build to the last byte in host memory (wall clock) */
    double ms_long_range;  /* of ms_integrands: stage-1 pair weights (the long-range integrand) */
    double ms_short_range; /* of ms_integrands: stage-2 r1 integrands (herm and non-herm, one kernel) */
    double ms_stage1;      /* stage 1 as a whole (Jastrow, integrand, contraction, split kernels) */
    double ms_stage2;      /* stage 2 as a whole, with the symmetrisation */
    double k_density;      /* modular stage 1: fraction of the (256-pair tile, K step)
                              tiles the contraction ran (1 = every one; 0 when not
                              measured: FP64, slices, graph capture) */
    double k_density_2e;   /* the same for the stage-2 contraction (256-pair H tiles) */
} cutc_stats;

/* FP64, slices = 0 (each precision's own default), everything else 0. */
void cutc_default_options(cutc_options *opt);

/* Whole path on device pointers (inputs already in HBM), on `stream`
 * (a hipStream_t, NULL = default stream); returns when the work is
 * complete.  X: (n_ao,n_ao,n_grid1,4) device buffer.  int_2e_ao: n_ao^4.
 * stats may be NULL.  The split precisions run the Jastrow of the next r1
 * chunk on a second stream of the library's (one per device, non-blocking,
 * default priority), ordered against `stream` by events. */
int cutc_int_device(const cutc_options *opt, int n_grid1, int n_grid2, int n_ao, int n_nuc, int size_bh,
                    const double *r1, const double *wr1, const double *r2, const double *wr2, const double *rn,
                    const double *aos_data1, const double *aos_data2, const double *c_bh, const int *m_bh,
                    const int *n_bh, const int *o_bh, double *int2_grad1_u12_ao, double *int_2e_ao, void *stream,
                    cutc_stats *stats);

/* cutc_int_device without any device allocation: no counterpart in the
 * reference.  For callers that cannot allocate (code under stream capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Same arguments,
 * layouts and results as cutc_int_device, minus the statistics.  All scratch
 * (the Jastrow program and every stage's workspace) lives in a fixed device
 * array of CUTC_NOALLOC_SCRATCH_BYTES (256 MiB) per GPU, held from the
 * library's first use of that GPU to the end of the process; the r1 chunks
 * are planned to fit it.  Returns CUTC_ERR_NOMEM when a single r1 row of a
 * stage does not fit (e.g. n_ao = 24 with 16000 r2 points fits in every
 * precision; n_ao = 200 with 200000 r2 points does not), CUTC_ERR_ARG beyond
 * 4096 Jastrow terms (n_nuc x size_bh).  Calls on one GPU share the array:
 * the library serialises its own calls and each returns after its work is
 * complete, so calls must not be issued so that they overlap on one GPU by
 * other means.  Nothing in the array needs to start at zero. */
#define CUTC_NOALLOC_SCRATCH_BYTES ((size_t)256 << 20)
int cutc_int_device_noalloc(const cutc_options *opt, int n_grid1, int n_grid2, int n_ao, int n_nuc, int size_bh,
                            const double *r1, const double *wr1, const double *r2, const double *wr2,
                            const double *rn, const double *aos_data1, const double *aos_data2, const double *c_bh,
                            const int *m_bh, const int *n_bh, const int *o_bh, double *int2_grad1_u12_ao,
                            double *int_2e_ao, void *stream);

/* cutc_int_device and cutc_int_ex keep their device workspace (one arena
 * per device, grown on demand) between calls, so repeated builds allocate
 * nothing; this frees it.  cutc_int, the get_* functions and the host-array
 * cutc_no and cutc_ao_to_mo release it before returning, as the reference
 * frees its buffers.  Calls are
 * serialised process-wide (they share the arenas). */
void cutc_release_workspace(void);

/* cutc_int with options and statistics (host arrays, like cutc_int). */
int cutc_int_ex(const cutc_options *opt, int n_grid1, int n_grid2, int n_ao, int n_nuc, int size_bh,
                const double *h_r1, const double *h_wr1, const double *h_r2, const double *h_wr2, const double *h_rn,
                const double *h_aos_data1, const double *h_aos_data2, const double *h_c_bh, const int *h_m_bh,
                const int *h_n_bh, const int *h_o_bh, double *h_int2_grad1_u12_ao, double *h_int_2e_ao,
                cutc_stats *stats);

/* Streamed whole path on device pointers: the same inputs and int_2e_ao as
 * cutc_int_device, without int2_grad1_u12_ao.  X is built and contracted r1
 * chunk by r1 chunk (one chunk buffer per call), so only int_2e_ao (n_ao^4
 * doubles) has to fit beside the working set.  No reference counterpart: it
 * serves workloads whose X (4 n_ao^2 n_grid1 doubles) does not fit in HBM. */
int cutc_int_streamed_device(const cutc_options *opt, int n_grid1, int n_grid2, int n_ao, int n_nuc, int size_bh,
                             const double *r1, const double *wr1, const double *r2, const double *wr2,
                             const double *rn, const double *aos_data1, const double *aos_data2, const double *c_bh,
                             const int *m_bh, const int *n_bh, const int *o_bh, double *int_2e_ao, void *stream,
                             cutc_stats *stats);

/* One system of a batch: the inputs of cutc_int_device on the device and
 * the system's own int_2e_ao (n_ao^4 doubles). */
typedef struct cutc_system {
    int n_grid1, n_grid2, n_ao, n_nuc, size_bh;
    const double *r1, *wr1, *r2, *wr2, *rn, *aos_data1, *aos_data2, *c_bh;
    const int *m_bh, *n_bh, *o_bh;
    double *int_2e_ao;
} cutc_system;

typedef struct cutc_batch_stats {
    double ms_setup;  /* plans, programs, autotuning, workspace, graph recording and instantiation */
    double ms_tune;   /* of which the autotuner (timed trial contractions) */
    double ms_launch; /* the one graph launch, stream-ordered hipEvents */
    int64_t n_nodes;  /* nodes recorded in the graph */
    int n_systems;
    int n_split;      /* systems whose step-2 contraction is split along K (FP64: by
                         the autotuner; block-sparse modular: by tile count) */
    int n_branches;   /* concurrent branches the systems were recorded on */
} cutc_batch_stats;

/* Batched build of independent systems of any sizes: each is built streamed
 * (X in r1 chunks of opt->n1_chunk points, by default whole systems as far
 * as a quarter of the device memory allows) into its own
 * int_2e_ao; the launches of the whole batch are recorded into one hipGraph
 * and launched once on `stream`.  opt->streams > 1 spreads the systems over
 * that many concurrent branches of the graph, each with its own workspace
 * and chunk buffer, every system on the branch with the least work so far
 * (for batches of small systems, whose launches leave the device idle; the
 * default is one branch).  Before recording, an autotuner times the
 * K splits of the FP64 step-2 contraction for every small pair count (cached
 * per shape for the process).  No reference counterpart (the reference
 * builds one system per call). */
int cutc_int_batch_device(const cutc_options *opt, int n_systems, const cutc_system *systems, void *stream,
                          cutc_batch_stats *stats);

/* Memory plan (pure host function): r1 chunk for the given device budget.
 * slices as in cutc_options (0: the precision's default).  Returns 0,
 * CUTC_ERR_ARG or CUTC_ERR_NOMEM. */
typedef struct cutc_plan {
    int64_t n1_chunk;      /* r1 rows per G chunk                   */
    int64_t ldk;           /* padded r2 length of the K-contiguous operands */
    int64_t n_pairs;       /* n_ao (n_ao + 1) / 2                   */
    int64_t n1_chunk_2e;   /* r1 rows per integrand chunk in stage 2 */
    size_t bytes_stage1;   /* workspace of stage 1                  */
    size_t bytes_stage2;   /* workspace of stage 2                  */
    int64_t k_chunk;       /* r2 points per K pass of the stage-1 contraction (ldk: one pass; modular
                              precisions take several when the pair residue planes do not fit) */
} cutc_plan;

int cutc_make_plan(int64_t n_grid1, int64_t n_grid2, int n_ao, int n_nuc, int precision, int slices,
                   size_t budget_bytes, int64_t n1_chunk_request, cutc_plan *plan);

#ifdef __cplusplus
}
#endif
#endif
