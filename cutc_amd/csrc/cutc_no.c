/*
 * Normal ordering (cutc_no): host driver.
 *
 * Reference: src/cutc_no.c:14-778.  The reference launches one kernel per
 * intermediate (closed- or open-shell set), two DGEMMs for the 2e part, two
 * DGEMVs and a DGEMM for the 1e part, a DDOT and a block reduction finished on
 * the host for the 0e part, then N + N^T and a 4-index transpose.  Here the
 * grid runs in chunks sized to the device memory; per chunk three fused
 * kernels (kernels/no.hip) build every intermediate, the 2e contractions run
 * on the FP64 MFMA GEMM with K-contiguous zero-padded operands and
 * accumulate over chunks, the 1e products are a GEMV and a split-K GEMM, and
 * the 0e sum is reduced on the device in a fixed order.  The symmetrisation
 * and the (q, s) swap run once at the end as one out-of-place kernel.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <hip/hip_runtime_api.h>

#include "cutc_no.h"
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

static size_t rnd(size_t b) { return (b + 255) & ~(size_t)255; }
static int64_t rup(int64_t a, int64_t m) { return (a + m - 1) / m * m; }

/* device bytes per padded grid point of a chunk, and the fixed part */
static size_t bytes_per_point(int64_t nm, int ne, int open)
{
    const int64_t nm2 = nm * nm;
    return (size_t)(8 * (8 * nm2 + 10 * nm + (open ? 6 * nm : 0) + 6 * nm + 6 * (int64_t)ne + 9));
}

/* split of the small 1e GEMMs (n_mo x n_mo outputs, K of several hundred
 * thousand): enough K slices to give every CU work */
static int split_1e(int64_t nm, int64_t K)
{
    const int64_t tiles = ((nm + 127) / 128) * ((nm + 127) / 128);
    int64_t s = 512 / (tiles > 0 ? tiles : 1);
    if (s > 64)
        s = 64;
    if (s > K / 1024)
        s = K / 1024;
    return s < 1 ? 1 : (int)s;
}

typedef struct {
    hipEvent_t e[2];
    int on;
} evpair;

static double ev_ms(evpair *p)
{
    float ms = 0.0f;
    if (p->on && hipEventSynchronize(p->e[1]) == hipSuccess && hipEventElapsedTime(&ms, p->e[0], p->e[1]) == hipSuccess)
        return ms;
    return 0.0;
}

/* phase timers: events recorded around each phase of each chunk */
#define NPH 6
typedef struct {
    hipEvent_t *ev;
    int *ph;
    int n, cap;
    hipStream_t st;
} phtimer;

static int ph_mark(phtimer *t, int phase)
{
    if (!t->ev)
        return 0;
    if (t->n == t->cap)
        return 0;
    if (hipEventCreate(&t->ev[t->n]) != hipSuccess)
        return 0;
    hipEventRecord(t->ev[t->n], t->st);
    t->ph[t->n++] = phase;
    return 1;
}

/* called before the kernels of each grid chunk are queued, with the end of
 * the chunk: the host entry point waits there for that part of its upload */
typedef struct {
    int (*before_chunk)(void *ctx, int64_t g_end);
    void *ctx;
} no_hook;

static int no_path(int64_t ng, int nm, int na, int nb, const double *w, const double *mol, const double *mor,
                   const double *G, double *o2, double *o1, double *o0, size_t mem_limit, int64_t req_chunk,
                   hipStream_t st, cutc_no_stats *stats, const no_hook *hook)
{
    if (ng < 0 || nm < 0 || na < 0 || nb < 0 || na > nm || nb > nm)
        return CUTC_ERR_ARG;
    if (!o0 || (nm > 0 && (!o1 || !o2)) || (ng > 0 && (!w || (nm > 0 && (!mol || !mor || !G)))))
        return CUTC_ERR_ARG;
    if (stats)
        memset(stats, 0, sizeof *stats);
    const int ne = na > nb ? na : nb, open = na != nb;
    const int64_t nm2 = (int64_t)nm * nm;
    int rc = CUTC_OK;
    phtimer pt = {0};
    evpair tot = {{0, 0}, 0};
    double *N = NULL, *work = NULL, *part = NULL;
    char *base = NULL;

    if (nm == 0 || ng == 0) { /* nothing to sum: zero outputs */
        if (nm > 0) {
            TRY(cutc_launch_fill(o2, nm2 * nm2, 0.0, st));
            TRY(cutc_launch_fill(o1, nm2, 0.0, st));
        }
        TRY(cutc_launch_fill(o0, 1, 0.0, st));
        TRY(hrc(hipStreamSynchronize(st)));
        return CUTC_OK;
    }

    /* chunk of the grid */
    const size_t per = bytes_per_point(nm, ne, open);
    const int ns_max = 64;
    size_t fixed = rnd(8 * (size_t)nm2 * (size_t)nm2) + rnd(8 * (size_t)ns_max * (size_t)nm2) + (1 << 20);
    const size_t budget = cutc_internal_budget(mem_limit);
    int64_t chunk = budget > fixed ? (int64_t)((budget - fixed) / per) : 0;
    if (req_chunk > 0 && req_chunk < chunk)
        chunk = req_chunk;
    if (chunk > ng)
        chunk = ng;
    if (chunk < 16 && chunk < ng)
        return CUTC_ERR_NOMEM;
    const int64_t nch = (ng + chunk - 1) / chunk;
    chunk = (ng + nch - 1) / nch;
    const int64_t ldk = rup(chunk, 16);
    const int64_t bpc = (ldk + 255) / 256; /* point-kernel blocks per chunk */
    const int64_t nblk = bpc * nch;
    fixed = rnd(8 * (size_t)nm2 * (size_t)nm2) + rnd(8 * (size_t)ns_max * (size_t)nm2) + rnd(16 * (size_t)nblk);
    const size_t need = fixed + per * (size_t)ldk + 64 * 256;
    TRY(cutc_internal_workspace(need, &base));
    if (stats) {
        stats->chunk = chunk;
        stats->n_chunks = nch;
        stats->workspace_bytes = need;
        stats->flops_gemm2e = 2.0 * (double)nm2 * (double)nm2 * 4.0 * (double)ng;
    }
    {
        char *q = base;
#define TAKE(p, n)                                                                                                   \
    do {                                                                                                             \
        p = (double *)q;                                                                                             \
        q += rnd(8 * (size_t)(n));                                                                                   \
    } while (0)
        TAKE(N, nm2 * nm2);
        TAKE(work, ns_max * nm2);
        TAKE(part, 2 * nblk);
        cutc_no_args a;
        memset(&a, 0, sizeof a);
        a.ng = ng;
        a.ldk = ldk;
        a.nblk = nblk;
        a.nm = nm;
        a.na = na;
        a.nb = nb;
        a.ne = ne;
        a.w = w;
        a.mol = mol;
        a.mor = mor;
        a.G = G;
        TAKE(a.J, 3 * ldk);
        TAKE(a.O, ldk);
        TAKE(a.S, ldk);
        TAKE(a.TE, 3 * (int64_t)ne * ldk + 1);
        TAKE(a.TF, 3 * (int64_t)ne * ldk + 1);
        TAKE(a.Dv, 4 * ldk);
        TAKE(a.A, 3 * (int64_t)nm * ldk);
        TAKE(a.Rt, 3 * (int64_t)nm * ldk);
        TAKE(a.E, 5 * (int64_t)nm * ldk);
        TAKE(a.F, 5 * (int64_t)nm * ldk);
        if (open) {
            TAKE(a.Gt, 3 * (int64_t)nm * ldk);
            TAKE(a.Ht, 3 * (int64_t)nm * ldk);
        }
        TAKE(a.C1x, 3 * nm2 * ldk);
        TAKE(a.Gx, 3 * nm2 * ldk);
        TAKE(a.C2, nm2 * ldk);
        TAKE(a.D2, nm2 * ldk);
#undef TAKE
        if ((size_t)(q - base) > need) {
            rc = CUTC_ERR_ARG; /* planner and carve disagree */
            goto done;
        }

        if (stats) {
            pt.cap = 8 * (int)nch + 8;
            pt.ev = calloc((size_t)pt.cap, sizeof(hipEvent_t));
            pt.ph = calloc((size_t)pt.cap, sizeof(int));
            pt.st = st;
            if (hipEventCreate(&tot.e[0]) == hipSuccess && hipEventCreate(&tot.e[1]) == hipSuccess) {
                tot.on = 1;
                hipEventRecord(tot.e[0], st);
            }
        }
        TRY(cutc_launch_fill(part, 2 * nblk, 0.0, st));
        TRY(cutc_launch_fill(o1, nm2, 0.0, st));
        const int64_t K1 = 5 * ldk, K3 = 3 * ldk;
        const int ns1 = cutc_gemm_tn_split_count(K1, split_1e(nm, K1));
        const int ns3 = cutc_gemm_tn_split_count(K3, split_1e(nm, K3));
        for (int64_t c = 0; c < nch; c++) {
            a.g0 = c * chunk;
            a.ngc = (ng - a.g0 < chunk) ? ng - a.g0 : chunk;
            if (hook)
                TRY(hook->before_chunk(hook->ctx, a.g0 + a.ngc));
            ph_mark(&pt, 0);
            TRY(cutc_no_point(&a, part + c * bpc, st));
            TRY(cutc_no_orb(&a, st));
            ph_mark(&pt, 1);
            TRY(cutc_no_pair(&a, st));
            ph_mark(&pt, 2);
            /* 2e: N (+)= 1/2 C1^T G over (x, k) and 1/2 C2^T D2 (src/cutc_no.c:419-446) */
            for (int x = 0; x < 3; x++)
                TRY(cutc_gemm_tn_f64(a.C1x + x * nm2 * ldk, ldk, a.Gx + x * nm2 * ldk, ldk, nm2, nm2, ldk, 0.5,
                                     (c == 0 && x == 0) ? 0.0 : 1.0, N, nm2, 1, NULL, st));
            TRY(cutc_gemm_tn_f64(a.C2, ldk, a.D2, ldk, nm2, nm2, ldk, 0.5, 1.0, N, nm2, 1, NULL, st));
            ph_mark(&pt, 3);
            /* 1e: 2 G^T Dv + 2 C2^T Dv3 + E^T F (+ Gt^T Ht) (src/cutc_no.c:481-525, 614-621) */
            TRY(cutc_no_gemv(a.Gx, a.C2, a.Dv, nm, ldk, a.ngc, o1, st));
            /* one slice writes its sum to work directly, several their partial sums */
            TRY(cutc_gemm_tn_f64(a.E, K1, a.F, K1, nm, nm, K1, 1.0, 0.0, work, nm, split_1e(nm, K1), work, st));
            TRY(cutc_no_add_slices(work, ns1, nm2, o1, st));
            if (open) {
                TRY(cutc_gemm_tn_f64(a.Gt, K3, a.Ht, K3, nm, nm, K3, 1.0, 0.0, work, nm, split_1e(nm, K3), work, st));
                TRY(cutc_no_add_slices(work, ns3, nm2, o1, st));
            }
            ph_mark(&pt, 4);
        }
        ph_mark(&pt, 5);
        TRY(cutc_no_sym_swap(N, nm, o2, st));
        TRY(cutc_no_finish_0e(part, nblk, o0, st));
        ph_mark(&pt, 6);
        if (tot.on)
            hipEventRecord(tot.e[1], st);
        TRY(hrc(hipStreamSynchronize(st)));
    }
    if (stats) {
        stats->ms_total = ev_ms(&tot);
        double *acc[NPH] = {&stats->ms_point, &stats->ms_pair, &stats->ms_gemm2e, &stats->ms_1e, NULL, &stats->ms_symm};
        for (int i = 0; i + 1 < pt.n; i++) {
            const int ph = pt.ph[i];
            if (ph < NPH && acc[ph]) {
                float ms = 0.0f;
                if (hipEventElapsedTime(&ms, pt.ev[i], pt.ev[i + 1]) == hipSuccess)
                    *acc[ph] += ms;
            }
        }
    }
done:
    if (rc != CUTC_OK)
        hipStreamSynchronize(st);
    for (int i = 0; i < pt.n; i++)
        hipEventDestroy(pt.ev[i]);
    free(pt.ev);
    free(pt.ph);
    if (tot.e[0])
        hipEventDestroy(tot.e[0]);
    if (tot.e[1])
        hipEventDestroy(tot.e[1]);
    return rc;
}

int cutc_no_device(int n_grid1, int n_mo, int ne_a, int ne_b, const double *wr1, const double *mos_l_in_r,
                   const double *mos_r_in_r, const double *int2_grad1_u12, double *no_2e, double *no_1e,
                   double *no_0e, size_t mem_limit, int64_t grid_chunk, void *stream, cutc_no_stats *stats)
{
    cutc_internal_lock();
    int rc = no_path(n_grid1, n_mo, ne_a, ne_b, wr1, mos_l_in_r, mos_r_in_r, int2_grad1_u12, no_2e, no_1e, no_0e,
                     mem_limit, grid_chunk, (hipStream_t)stream, stats, NULL);
    cutc_internal_unlock();
    cutc_set_last_error(rc);
    return rc;
}

/* Host entry point: the uploads of int2_grad1_u12 and the build overlap.
 * The grid is cut into NO_UP_CHUNKS parts; each part of int2_grad1_u12 (a
 * strided block of the (n_grid1, 3, n_mo, n_mo) array) goes through the
 * transfer pipeline, and the build starts on a part as soon as it is in
 * device memory.  no_2e comes back through the same pipeline. */
#define NO_UP_CHUNKS 8

typedef struct {
    cutc_xfer *xf;
    int64_t part; /* grid points per upload part */
    int64_t tickets[NO_UP_CHUNKS];
    int nparts;
} no_upload_t;

static int wait_upload(void *ctx, int64_t g_end)
{
    no_upload_t *u = (no_upload_t *)ctx;
    int64_t k = (g_end + u->part - 1) / u->part - 1;
    if (k >= u->nparts)
        k = u->nparts - 1;
    return k >= 0 ? cutc_xfer_wait(u->xf, u->tickets[k]) : CUTC_OK;
}

/* Host arrays: upload, build, download (src/cutc_no.c:154-157, 703-704),
 * then free every device buffer, as the reference does. */
int cutc_no(int n_grid1, int n_mo, int ne_a, int ne_b, double *h_wr1, double *h_mos_l_in_r, double *h_mos_r_in_r,
            double *h_int2_grad1_u12, double *h_no_2e, double *h_no_1e, double *h_no_0e)
{
    int rc = CUTC_OK;
    double *d[7] = {NULL, NULL, NULL, NULL, NULL, NULL, NULL};
    cutc_no_stats stt;
    no_upload_t up;
    memset(&up, 0, sizeof up);
    if (n_grid1 < 0 || n_mo < 0 || ne_a < 0 || ne_b < 0 || ne_a > n_mo || ne_b > n_mo) {
        cutc_set_last_error(CUTC_ERR_ARG);
        return CUTC_ERR_ARG;
    }
    printf(" Computing 0e, 1e, and 2e-Elements for Normal-Ordering With CuTC\n");
    const size_t ng = (size_t)n_grid1, nm = (size_t)n_mo, nm2 = nm * nm;
    const size_t sz[7] = {ng, ng * nm, ng * nm, 3 * ng * nm2, nm2 * nm2, nm2, 1};
    const double *h_in[4] = {h_wr1, h_mos_l_in_r, h_mos_r_in_r, h_int2_grad1_u12};
    for (int i = 0; i < 4; i++)
        if (sz[i] > 0 && !h_in[i]) {
            cutc_set_last_error(CUTC_ERR_ARG);
            return CUTC_ERR_ARG;
        }
    /* the library lock from the first allocation to the last free: other
     * callers plan their workspace against the same free memory */
    cutc_internal_lock();
    for (int i = 0; i < 7; i++)
        if (sz[i] > 0 && (rc = hrc(hipMalloc((void **)&d[i], 8 * sz[i]))) != CUTC_OK)
            goto out;
    for (int i = 0; i < 3; i++)
        if (sz[i] > 0 && (rc = hrc(hipMemcpy(d[i], h_in[i], 8 * sz[i], hipMemcpyHostToDevice))) != CUTC_OK)
            goto out;
    /* the pipeline needs one grid row of a part (8 part bytes) within a staging buffer */
    const size_t stage = (size_t)64 << 20;
    int64_t part = (int64_t)((ng + NO_UP_CHUNKS - 1) / NO_UP_CHUNKS);
    part = (part + 15) / 16 * 16;
    if (sz[3] > 0 && ng >= 2 * 16 && (size_t)part * 8 <= stage) {
        if ((rc = cutc_xfer_open(&up.xf, stage, 4, cutc_internal_copy_threads())) != CUTC_OK)
            goto out;
        up.part = part;
        for (int64_t g0 = 0; g0 < (int64_t)ng && up.nparts < NO_UP_CHUNKS; g0 += part) {
            const int64_t cnt = ((int64_t)ng - g0 < part) ? (int64_t)ng - g0 : part;
            rc = cutc_xfer_h2d_2d(up.xf, d[3] + g0, 8 * ng, h_int2_grad1_u12 + g0, 8 * ng, 8 * (size_t)cnt, 3 * nm2,
                                  &up.tickets[up.nparts]);
            if (rc != CUTC_OK)
                goto out;
            up.nparts++;
        }
    } else if (sz[3] > 0 && (rc = hrc(hipMemcpy(d[3], h_in[3], 8 * sz[3], hipMemcpyHostToDevice))) != CUTC_OK) {
        goto out;
    }
    {
        no_hook hook = {wait_upload, &up};
        rc = no_path(n_grid1, n_mo, ne_a, ne_b, d[0], d[1], d[2], d[3], d[4], d[5], d[6], 0, up.xf ? up.part : 0,
                     NULL, &stt, up.xf ? &hook : NULL);
    }
    if (rc == CUTC_OK) {
        double *h_out[3] = {h_no_2e, h_no_1e, h_no_0e};
        if (up.xf) { /* no_2e through the pipeline, the small outputs directly */
            rc = cutc_xfer_d2h(up.xf, h_out[0], d[4], 8 * sz[4], NULL);
            if (rc == CUTC_OK)
                rc = cutc_xfer_finish(up.xf, NULL);
        } else if (sz[4] > 0) {
            rc = hrc(hipMemcpy(h_out[0], d[4], 8 * sz[4], hipMemcpyDeviceToHost));
        }
        for (int i = 1; i < 3 && rc == CUTC_OK; i++)
            if (sz[4 + i] > 0)
                rc = hrc(hipMemcpy(h_out[i], d[4 + i], 8 * sz[4 + i], hipMemcpyDeviceToHost));
        printf("Ellapsed (effective) time on GPU for cutc_no = %.3f sec\n", stt.ms_total / 1000.0);
    }
out:
    cutc_xfer_close(up.xf); /* drains queued copies before the buffers go */
    for (int i = 0; i < 7; i++)
        if (d[i])
            hipFree(d[i]);
    cutc_internal_release_workspace();
    cutc_internal_unlock();
    cutc_set_last_error(rc);
    return rc;
}
