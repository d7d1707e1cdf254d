/*
 * Host <-> device transfer pipeline of the host entry points.
 *
 * The reference copies int2_grad1_u12_ao and int_2e_ao back with blocking
 * cudaMemcpy calls after the whole build (src/cutc_int.c:210-211).  Here each
 * block of device memory is queued as soon as the stream that writes it has
 * been given the work (an event marks the point); one driver thread moves it
 * through a ring of pinned staging buffers: the DMA of a piece on a
 * non-blocking copy stream overlaps the host copy of the previous piece into
 * the caller's (pageable) array, which a small pool of threads performs in
 * parallel.  The build keeps running on its own stream meanwhile, so the
 * transfers of finished r1 chunks hide behind the contractions of later
 * ones.  Uploads go the other way through the same ring (host threads
 * gather rows into a staging buffer while the DMA of the previous one runs)
 * and are waited for by ticket, so a build can start on the first part of
 * an input while the rest is still on its way.
 */
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#include <hip/hip_runtime_api.h>

#include "cutc_int.h"
#include "xfer.h"

#define XF_MAX_STAGE 8
#define XF_MAX_THREADS 16

typedef struct job {
    char *dst;
    const char *src;
    size_t bytes;
    hipEvent_t ready;
    /* host -> device jobs: rows x width bytes, row pitches dpitch (device)
     * and spitch (host); ticket = their sequence number */
    int h2d;
    size_t dpitch, spitch, width, rows;
    int64_t ticket;
    struct job *next;
} job_t;

struct cutc_xfer {
    hipStream_t cs;
    char *stage[XF_MAX_STAGE];
    hipEvent_t done[XF_MAX_STAGE];
    int nstage;
    size_t stage_bytes;

    pthread_mutex_t mu;
    pthread_cond_t cv, cv_h2d;
    job_t *head, *tail;
    int closing;
    int err;
    size_t bytes_done;
    int64_t h2d_issued, h2d_done;
    int driver_exited;
    pthread_t driver;
    int driver_on;

    /* host copy pool: nthr workers plus the driver, one generation per copy */
    int nthr;
    pthread_t thr[XF_MAX_THREADS];
    pthread_mutex_t pm;
    pthread_cond_t pgo, pdone;
    unsigned gen;
    int remaining;
    int cp_quit;
    char *cp_dst;
    const char *cp_src;
    size_t cp_bytes;  /* one row: bytes; several rows: row width */
    size_t cp_rows, cp_dpitch, cp_spitch;
};

static int hrc(hipError_t e)
{
    if (e == hipSuccess)
        return CUTC_OK;
    return e == hipErrorOutOfMemory ? CUTC_ERR_NOMEM : CUTC_ERR_HIP;
}

/* share w of nthr + 1 of the current host copy: one row is cut into
 * 64-byte aligned pieces, several rows are dealt out by row */
static void copy_share(cutc_xfer *x, int w)
{
    const int parts = x->nthr + 1;
    if (x->cp_rows <= 1) {
        size_t per = (x->cp_bytes / parts + 63) & ~(size_t)63;
        size_t a = per * (size_t)w, b = a + per;
        if (a > x->cp_bytes)
            a = x->cp_bytes;
        if (b > x->cp_bytes)
            b = x->cp_bytes;
        if (b > a)
            memcpy(x->cp_dst + a, x->cp_src + a, b - a);
        return;
    }
    const size_t per = (x->cp_rows + parts - 1) / parts;
    size_t a = per * (size_t)w, b = a + per;
    if (b > x->cp_rows)
        b = x->cp_rows;
    for (size_t r = a; r < b; r++)
        memcpy(x->cp_dst + r * x->cp_dpitch, x->cp_src + r * x->cp_spitch, x->cp_bytes);
}

typedef struct {
    cutc_xfer *x;
    int w;
} worker_arg;

static void *pool_main(void *arg)
{
    cutc_xfer *x = ((worker_arg *)arg)->x;
    const int w = ((worker_arg *)arg)->w;
    free(arg);
    unsigned seen = 0;
    for (;;) {
        pthread_mutex_lock(&x->pm);
        while (x->gen == seen && !x->cp_quit)
            pthread_cond_wait(&x->pgo, &x->pm);
        const int quit = x->cp_quit;
        seen = x->gen;
        pthread_mutex_unlock(&x->pm);
        if (quit)
            break;
        copy_share(x, w);
        pthread_mutex_lock(&x->pm);
        if (--x->remaining == 0)
            pthread_cond_signal(&x->pdone);
        pthread_mutex_unlock(&x->pm);
    }
    return NULL;
}

/* rows x width bytes from src (row pitch spitch) to dst (row pitch dpitch),
 * shared with the pool when large */
static void host_copy2d(cutc_xfer *x, char *dst, size_t dpitch, const char *src, size_t spitch, size_t width,
                        size_t rows)
{
    if (x->nthr == 0 || width * rows < ((size_t)4 << 20)) {
        for (size_t r = 0; r < rows; r++)
            memcpy(dst + r * dpitch, src + r * spitch, width);
        return;
    }
    pthread_mutex_lock(&x->pm);
    x->cp_dst = dst;
    x->cp_src = src;
    x->cp_bytes = width;
    x->cp_rows = rows;
    x->cp_dpitch = dpitch;
    x->cp_spitch = spitch;
    x->remaining = x->nthr;
    x->gen++;
    pthread_cond_broadcast(&x->pgo);
    pthread_mutex_unlock(&x->pm);
    copy_share(x, x->nthr);
    pthread_mutex_lock(&x->pm);
    while (x->remaining > 0)
        pthread_cond_wait(&x->pdone, &x->pm);
    pthread_mutex_unlock(&x->pm);
}

/* one piece in flight: staging buffer b holds bytes for dst */
typedef struct {
    char *dst;
    size_t bytes;
} piece_t;

static void set_err(cutc_xfer *x, int rc)
{
    pthread_mutex_lock(&x->mu);
    if (!x->err)
        x->err = rc;
    pthread_cond_broadcast(&x->cv_h2d);
    pthread_mutex_unlock(&x->mu);
}

/* One host -> device job, run when no download piece is in flight: the rows
 * are gathered into the staging ring (host threads) while the DMAs of the
 * previous pieces run; returns once every piece is in device memory. */
static int run_h2d(cutc_xfer *x, const job_t *j)
{
    const size_t per = x->stage_bytes / j->width;
    size_t r0 = 0;
    int64_t issued = 0;
    int rc = CUTC_OK;
    while (r0 < j->rows && rc == CUTC_OK) {
        const int b = (int)(issued % x->nstage);
        const size_t nr = (j->rows - r0 < per) ? j->rows - r0 : per;
        if (issued >= x->nstage && (rc = hrc(hipEventSynchronize(x->done[b]))) != CUTC_OK)
            break;
        host_copy2d(x, x->stage[b], j->width, j->src + r0 * j->spitch, j->spitch, j->width, nr);
        rc = hrc(hipMemcpy2DAsync(j->dst + r0 * j->dpitch, j->dpitch, x->stage[b], j->width, j->width, nr,
                                  hipMemcpyHostToDevice, x->cs));
        if (rc == CUTC_OK)
            rc = hrc(hipEventRecord(x->done[b], x->cs));
        issued++;
        r0 += nr;
    }
    const int rs = hrc(hipStreamSynchronize(x->cs));
    return rc != CUTC_OK ? rc : rs;
}

static void *driver_main(void *arg)
{
    cutc_xfer *x = (cutc_xfer *)arg;
    piece_t fl[XF_MAX_STAGE];
    int first = 0, inflight = 0, err = 0;
    job_t *cur = NULL, *retired = NULL;
    size_t off = 0;
    for (;;) {
        /* issue DMAs while a staging buffer is free and work is queued */
        while (inflight < x->nstage && !err) {
            if (!cur) {
                pthread_mutex_lock(&x->mu);
                cur = x->head;
                if (cur && cur->h2d && inflight > 0)
                    cur = NULL; /* an upload waits for the downloads in flight */
                if (cur) {
                    x->head = cur->next;
                    if (!x->head)
                        x->tail = NULL;
                }
                pthread_mutex_unlock(&x->mu);
                if (!cur)
                    break;
                off = 0;
                if (cur->h2d) {
                    err = run_h2d(x, cur);
                    pthread_mutex_lock(&x->mu);
                    if (err == CUTC_OK)
                        x->h2d_done = cur->ticket;
                    pthread_cond_broadcast(&x->cv_h2d);
                    pthread_mutex_unlock(&x->mu);
                    cur->next = retired;
                    retired = cur;
                    cur = NULL;
                    if (err)
                        break;
                    continue;
                }
                if ((err = hrc(hipStreamWaitEvent(x->cs, cur->ready, 0))) != CUTC_OK)
                    break;
            }
            const int b = (first + inflight) % x->nstage;
            const size_t n = (cur->bytes - off < x->stage_bytes) ? cur->bytes - off : x->stage_bytes;
            if ((err = hrc(hipMemcpyAsync(x->stage[b], cur->src + off, n, hipMemcpyDeviceToHost, x->cs))) != CUTC_OK)
                break;
            if ((err = hrc(hipEventRecord(x->done[b], x->cs))) != CUTC_OK)
                break;
            fl[b].dst = cur->dst + off;
            fl[b].bytes = n;
            inflight++;
            off += n;
            if (off >= cur->bytes) { /* its event may still be waited on: free after the stream drains */
                cur->next = retired;
                retired = cur;
                cur = NULL;
            }
        }
        if (err) {
            set_err(x, err);
            break;
        }
        if (inflight == 0) {
            pthread_mutex_lock(&x->mu);
            while (!x->head && !x->closing)
                pthread_cond_wait(&x->cv, &x->mu);
            const int stop = !x->head && x->closing;
            pthread_mutex_unlock(&x->mu);
            if (stop)
                break;
            continue;
        }
        /* finish the oldest piece */
        if ((err = hrc(hipEventSynchronize(x->done[first]))) != CUTC_OK) {
            set_err(x, err);
            break;
        }
        host_copy2d(x, fl[first].dst, 0, x->stage[first], 0, fl[first].bytes, 1);
        pthread_mutex_lock(&x->mu);
        x->bytes_done += fl[first].bytes;
        pthread_mutex_unlock(&x->mu);
        first = (first + 1) % x->nstage;
        inflight--;
    }
    hipStreamSynchronize(x->cs);
    if (cur) {
        cur->next = retired;
        retired = cur;
    }
    while (retired) {
        job_t *j = retired;
        retired = j->next;
        if (j->ready)
            hipEventDestroy(j->ready);
        free(j);
    }
    pthread_mutex_lock(&x->mu);
    x->driver_exited = 1;
    pthread_cond_broadcast(&x->cv_h2d);
    pthread_mutex_unlock(&x->mu);
    return NULL;
}

int cutc_xfer_open(cutc_xfer **out, size_t stage_bytes, int nstage, int nthreads)
{
    *out = NULL;
    if (nstage < 2)
        nstage = 2;
    if (nstage > XF_MAX_STAGE)
        nstage = XF_MAX_STAGE;
    if (nthreads < 0)
        nthreads = 0;
    if (nthreads > XF_MAX_THREADS)
        nthreads = XF_MAX_THREADS;
    cutc_xfer *x = (cutc_xfer *)calloc(1, sizeof *x);
    if (!x)
        return CUTC_ERR_NOMEM;
    pthread_mutex_init(&x->mu, NULL);
    pthread_cond_init(&x->cv, NULL);
    pthread_cond_init(&x->cv_h2d, NULL);
    x->nstage = nstage;
    x->stage_bytes = stage_bytes;
    int rc = hrc(hipStreamCreateWithFlags(&x->cs, hipStreamNonBlocking));
    for (int b = 0; b < nstage && rc == CUTC_OK; b++) {
        rc = hrc(hipHostMalloc((void **)&x->stage[b], stage_bytes, hipHostMallocDefault));
        if (rc == CUTC_OK)
            rc = hrc(hipEventCreateWithFlags(&x->done[b], hipEventDisableTiming));
    }
    pthread_mutex_init(&x->pm, NULL);
    pthread_cond_init(&x->pgo, NULL);
    pthread_cond_init(&x->pdone, NULL);
    for (int i = 0; i < nthreads && rc == CUTC_OK; i++) { /* fewer workers if a thread cannot start */
        worker_arg *a = (worker_arg *)malloc(sizeof *a);
        if (!a)
            break;
        a->x = x;
        a->w = i;
        if (pthread_create(&x->thr[i], NULL, pool_main, a) != 0) {
            free(a);
            break;
        }
        x->nthr++;
    }
    if (rc == CUTC_OK) {
        if (pthread_create(&x->driver, NULL, driver_main, x) == 0)
            x->driver_on = 1;
        else
            rc = CUTC_ERR_HIP;
    }
    if (rc != CUTC_OK) {
        cutc_xfer_close(x);
        return rc;
    }
    *out = x;
    return CUTC_OK;
}

int cutc_xfer_d2h(cutc_xfer *x, void *h_dst, const void *d_src, size_t bytes, hipStream_t producer)
{
    if (bytes == 0)
        return CUTC_OK;
    job_t *j = (job_t *)calloc(1, sizeof *j);
    if (!j)
        return CUTC_ERR_NOMEM;
    int rc = hrc(hipEventCreateWithFlags(&j->ready, hipEventDisableTiming));
    if (rc == CUTC_OK)
        rc = hrc(hipEventRecord(j->ready, producer));
    if (rc != CUTC_OK) {
        if (j->ready)
            hipEventDestroy(j->ready);
        free(j);
        set_err(x, rc);
        return rc;
    }
    j->dst = (char *)h_dst;
    j->src = (const char *)d_src;
    j->bytes = bytes;
    pthread_mutex_lock(&x->mu);
    if (x->tail)
        x->tail->next = j;
    else
        x->head = j;
    x->tail = j;
    pthread_cond_signal(&x->cv);
    pthread_mutex_unlock(&x->mu);
    return CUTC_OK;
}

int cutc_xfer_h2d_2d(cutc_xfer *x, void *d_dst, size_t dpitch, const void *h_src, size_t spitch, size_t width,
                     size_t rows, int64_t *ticket)
{
    *ticket = 0;
    if (width == 0 || rows == 0)
        return CUTC_OK;
    if (width > x->stage_bytes || (rows > 1 && (dpitch < width || spitch < width)))
        return CUTC_ERR_ARG;
    job_t *j = (job_t *)calloc(1, sizeof *j);
    if (!j)
        return CUTC_ERR_NOMEM;
    j->h2d = 1;
    j->dst = (char *)d_dst;
    j->src = (const char *)h_src;
    j->dpitch = dpitch;
    j->spitch = spitch;
    j->width = width;
    j->rows = rows;
    pthread_mutex_lock(&x->mu);
    j->ticket = ++x->h2d_issued;
    *ticket = j->ticket;
    if (x->tail)
        x->tail->next = j;
    else
        x->head = j;
    x->tail = j;
    pthread_cond_signal(&x->cv);
    pthread_mutex_unlock(&x->mu);
    return CUTC_OK;
}

int cutc_xfer_h2d(cutc_xfer *x, void *d_dst, const void *h_src, size_t bytes, int64_t *ticket)
{
    /* whole rows of 1 MiB, then the rest as one short row */
    const size_t w = (size_t)1 << 20;
    const size_t rows = bytes / w, rest = bytes - rows * w;
    int64_t t = 0;
    int rc = CUTC_OK;
    *ticket = 0;
    if (rows > 0 && (rc = cutc_xfer_h2d_2d(x, d_dst, w, h_src, w, w, rows, &t)) != CUTC_OK)
        return rc;
    *ticket = t;
    if (rest > 0 && (rc = cutc_xfer_h2d_2d(x, (char *)d_dst + rows * w, rest, (const char *)h_src + rows * w, rest,
                                           rest, 1, &t)) != CUTC_OK)
        return rc;
    *ticket = t;
    return CUTC_OK;
}

int cutc_xfer_wait(cutc_xfer *x, int64_t ticket)
{
    pthread_mutex_lock(&x->mu);
    while (x->h2d_done < ticket && !x->err && !x->driver_exited)
        pthread_cond_wait(&x->cv_h2d, &x->mu);
    const int rc = (x->h2d_done >= ticket) ? CUTC_OK : (x->err ? x->err : CUTC_ERR_HIP);
    pthread_mutex_unlock(&x->mu);
    return rc;
}

int cutc_xfer_finish(cutc_xfer *x, size_t *bytes_done)
{
    pthread_mutex_lock(&x->mu);
    x->closing = 1;
    pthread_cond_signal(&x->cv);
    pthread_mutex_unlock(&x->mu);
    if (x->driver_on) {
        pthread_join(x->driver, NULL);
        x->driver_on = 0;
    }
    if (bytes_done)
        *bytes_done = x->bytes_done;
    return x->err;
}

void cutc_xfer_close(cutc_xfer *x)
{
    if (!x)
        return;
    cutc_xfer_finish(x, NULL);
    while (x->head) { /* jobs left after an error */
        job_t *j = x->head;
        x->head = j->next;
        if (j->ready)
            hipEventDestroy(j->ready);
        free(j);
    }
    pthread_mutex_lock(&x->pm);
    x->cp_quit = 1;
    pthread_cond_broadcast(&x->pgo);
    pthread_mutex_unlock(&x->pm);
    for (int i = 0; i < x->nthr; i++)
        pthread_join(x->thr[i], NULL);
    pthread_mutex_destroy(&x->pm);
    pthread_cond_destroy(&x->pgo);
    pthread_cond_destroy(&x->pdone);
    for (int b = 0; b < x->nstage; b++) {
        if (x->stage[b])
            hipHostFree(x->stage[b]);
        if (x->done[b])
            hipEventDestroy(x->done[b]);
    }
    if (x->cs)
        hipStreamDestroy(x->cs);
    pthread_mutex_destroy(&x->mu);
    pthread_cond_destroy(&x->cv);
    pthread_cond_destroy(&x->cv_h2d);
    free(x);
}
