/*
 * AO -> MO transform of int2_grad1_u12 (cutc_ao_to_mo): host driver.
 *
 * CuTC leaves this step to its callers (cutc_int returns the AO-basis
 * int2_grad1_u12_ao(n_ao, n_ao, n_grid1, 3), src/cutc_int.c:210-211;
 * cutc_no reads the MO-basis int2_grad1_u12(n_grid1, 3, n_mo, n_mo),
 * src/no_tmpJ_cs_kernel.cu:18).  For each grid point g and component x the
 * result is C_L^T A_gx C_R.  With R = 3 n_grid1 rows gx = g + n_grid1 x:
 *
 *   pass 1  Z[l + n_ao (gx + R i')] = sum_k A[k + n_ao (l + n_ao gx)] C_L[k + n_ao i]
 *           one TN GEMM: rows (l, gx) of A at stride n_ao, columns i of the
 *           pass, K = k;
 *   pass 2  out[gx + R i + R n_mo j] = sum_l Z[l + n_ao (gx + R i')] C_R[l + n_ao j]
 *           one TN GEMM: rows (gx, i') of Z at stride n_ao, columns j, K = l,
 *           written straight into the (n_grid1, 3, n_mo, n_mo) layout.
 *
 * Both are K-contiguous on both operands, so the library's FP64 MFMA GEMM
 * runs them without any transpose.  The MO index i runs in passes sized to
 * the free device memory (Z holds n_ao R doubles per index i); each pass
 * reads all of A once.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <hip/hip_runtime_api.h>

#include "cutc_mo.h"
#include "internal.h"
#include "kernels/launch.h"
#include "xfer.h"

static int hrc(hipError_t e)
{
    if (e == hipSuccess)
        return CUTC_OK;
    return e == hipErrorOutOfMemory ? CUTC_ERR_NOMEM : CUTC_ERR_HIP;
}

#define TRY(x)                                                                                                       \
    do {                                                                                                             \
        rc = (x);                                                                                                    \
        if (rc != CUTC_OK)                                                                                           \
            goto done;                                                                                               \
    } while (0)

/* events around each pass of each chunk: 2 per pass, 2 passes per chunk */
typedef struct {
    hipEvent_t *ev;
    int n, cap;
} evlist;

static void mark(evlist *l, hipStream_t st)
{
    if (l->ev && l->n < l->cap && hipEventCreate(&l->ev[l->n]) == hipSuccess)
        hipEventRecord(l->ev[l->n++], st);
}

static int mo_path(int64_t n1, int64_t na, int64_t nm, const double *CL, const double *CR, const double *A,
                   double *out, size_t mem_limit, int64_t req_chunk, hipStream_t st, cutc_mo_stats *stats)
{
    if (n1 < 0 || na < 0 || nm < 0 || req_chunk < 0)
        return CUTC_ERR_ARG;
    const int64_t R = 3 * n1;
    if (stats)
        memset(stats, 0, sizeof *stats);
    if (R == 0 || nm == 0)
        return CUTC_OK;
    if (!out || (na > 0 && (!CL || !CR || !A)))
        return CUTC_ERR_ARG;
    int rc = CUTC_OK;
    char *base = NULL;
    evlist ev = {0};
    hipEvent_t t0 = NULL, t1 = NULL;
    if (na == 0) { /* empty sums */
        TRY(cutc_launch_fill(out, R * nm * nm, 0.0, st));
        TRY(hrc(hipStreamSynchronize(st)));
        goto done;
    }

    /* MO indices per pass: what the budget holds, in whole 128-column GEMM
     * tiles when it is not all of them */
    const size_t per_mo = (size_t)8 * (size_t)na * (size_t)R;
    const size_t budget = cutc_internal_budget(mem_limit);
    int64_t ci = (int64_t)(budget / per_mo);
    if (ci > nm)
        ci = nm;
    if (ci < nm && ci >= 128)
        ci -= ci % 128;
    if (req_chunk > 0 && req_chunk < ci)
        ci = req_chunk;
    if (ci < 1)
        return CUTC_ERR_NOMEM;
    const int64_t nchunk = (nm + ci - 1) / ci;
    const size_t need = per_mo * (size_t)ci;
    TRY(cutc_internal_workspace(need, &base));
    double *Z = (double *)base;

    if (stats) {
        ev.cap = (int)(4 * nchunk);
        ev.ev = (hipEvent_t *)calloc((size_t)ev.cap, sizeof(hipEvent_t));
        if (hipEventCreate(&t0) == hipSuccess && hipEventCreate(&t1) == hipSuccess)
            hipEventRecord(t0, st);
    }
    for (int64_t i0 = 0; i0 < nm; i0 += ci) {
        const int64_t c = (nm - i0 < ci) ? nm - i0 : ci;
        mark(&ev, st);
        TRY(cutc_gemm_tn_f64(A, na, CL + na * i0, na, na * R, c, na, 1.0, 0.0, Z, na * R, 1, NULL, st));
        mark(&ev, st);
        mark(&ev, st);
        TRY(cutc_gemm_tn_f64(Z, na, CR, na, R * c, nm, na, 1.0, 0.0, out + R * i0, R * nm, 1, NULL, st));
        mark(&ev, st);
    }
    if (t1)
        hipEventRecord(t1, st);
    TRY(hrc(hipStreamSynchronize(st)));
    if (stats) {
        float ms = 0.0f;
        if (t1 && hipEventElapsedTime(&ms, t0, t1) == hipSuccess)
            stats->ms_total = ms;
        for (int i = 0; i + 1 < ev.n; i += 2)
            if (hipEventElapsedTime(&ms, ev.ev[i], ev.ev[i + 1]) == hipSuccess)
                *((i / 2) % 2 ? &stats->ms_right : &stats->ms_left) += ms;
        stats->flops = 2.0 * (double)R * (double)na * (double)nm * (double)(na + nm);
        stats->mo_chunk = ci;
        stats->n_chunks = nchunk;
        stats->workspace_bytes = need;
    }
done:
    if (rc != CUTC_OK)
        hipStreamSynchronize(st);
    for (int i = 0; i < ev.n; i++)
        hipEventDestroy(ev.ev[i]);
    free(ev.ev);
    if (t0)
        hipEventDestroy(t0);
    if (t1)
        hipEventDestroy(t1);
    return rc;
}

int cutc_ao_to_mo_device(int n_grid1, int n_ao, int n_mo, const double *mo_l_coef, const double *mo_r_coef,
                         const double *int2_grad1_u12_ao, double *int2_grad1_u12, size_t mem_limit, int64_t mo_chunk,
                         void *stream, cutc_mo_stats *stats)
{
    cutc_internal_lock();
    int rc = mo_path(n_grid1, n_ao, n_mo, mo_l_coef, mo_r_coef, int2_grad1_u12_ao, int2_grad1_u12, mem_limit, mo_chunk,
                     (hipStream_t)stream, stats);
    cutc_internal_unlock();
    cutc_set_last_error(rc);
    return rc;
}

/* Host arrays: upload, transform, download, free every device buffer.  The
 * two big arrays move through the library's transfer pipeline (pinned
 * staging, host copy threads), which is several times faster than copies
 * from and to pageable memory; the transform itself is short next to them. */
int cutc_ao_to_mo(int n_grid1, int n_ao, int n_mo, double *h_mo_l_coef, double *h_mo_r_coef,
                  double *h_int2_grad1_u12_ao, double *h_int2_grad1_u12)
{
    int rc = CUTC_OK;
    double *d[4] = {NULL, NULL, NULL, NULL};
    cutc_xfer *xf = NULL;
    if (n_grid1 < 0 || n_ao < 0 || n_mo < 0) {
        cutc_set_last_error(CUTC_ERR_ARG);
        return CUTC_ERR_ARG;
    }
    const size_t n1 = (size_t)n_grid1, na = (size_t)n_ao, nm = (size_t)n_mo;
    const size_t sz[4] = {na * nm, na * nm, na * na * 3 * n1, 3 * n1 * nm * nm};
    const double *h_in[3] = {h_mo_l_coef, h_mo_r_coef, h_int2_grad1_u12_ao};
    for (int i = 0; i < 4; i++)
        if (sz[i] > 0 && (i < 3 ? !h_in[i] : !h_int2_grad1_u12)) {
            cutc_set_last_error(CUTC_ERR_ARG);
            return CUTC_ERR_ARG;
        }
    /* the library lock from the first allocation to the last free */
    cutc_internal_lock();
    for (int i = 0; i < 4; i++)
        if (sz[i] > 0 && (rc = hrc(hipMalloc((void **)&d[i], 8 * sz[i]))) != CUTC_OK)
            goto out;
    for (int i = 0; i < 2; i++)
        if (sz[i] > 0 && (rc = hrc(hipMemcpy(d[i], h_in[i], 8 * sz[i], hipMemcpyHostToDevice))) != CUTC_OK)
            goto out;
    if (sz[2] > 0 || sz[3] > 0) {
        int64_t t = 0;
        if ((rc = cutc_xfer_open(&xf, (size_t)64 << 20, 4, cutc_internal_copy_threads())) != CUTC_OK)
            goto out;
        if (sz[2] > 0 && ((rc = cutc_xfer_h2d(xf, d[2], h_in[2], 8 * sz[2], &t)) != CUTC_OK ||
                          (rc = cutc_xfer_wait(xf, t)) != CUTC_OK))
            goto out;
    }
    rc = mo_path(n_grid1, n_ao, n_mo, d[0], d[1], d[2], d[3], 0, 0, NULL, NULL);
    if (rc == CUTC_OK && sz[3] > 0) {
        rc = cutc_xfer_d2h(xf, h_int2_grad1_u12, d[3], 8 * sz[3], NULL);
        if (rc == CUTC_OK)
            rc = cutc_xfer_finish(xf, NULL);
    }
out:
    cutc_xfer_close(xf);
    for (int i = 0; i < 4; i++)
        if (d[i])
            hipFree(d[i]);
    cutc_internal_release_workspace();
    cutc_internal_unlock();
    cutc_set_last_error(rc);
    return rc;
}
