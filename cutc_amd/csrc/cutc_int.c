/*
 * C host layer of libcutcint (gfx950): the reference's cutc_int path behind
 * its C ABI, driving the HIP kernels (kernels/ directory) through launch.h.
 *
 * Reference host code replaced here:
 *   cutc_int               src/cutc_int.c:15-243
 *   get_int2_grad1_u12_ao  src/build_int2_grad1_u12_ao.cu:9-347
 *   get_int_2e_ao          src/build_int_2e_ao.cu:11-145
 *   exported launch wrappers src/jast_bh.cu:268-303, src/long_range_integ.cu:52-66,
 *     src/short_range_integ_herm.cu:51-63, src/short_range_integ_nonherm.cu:69-81,
 *     src/add_trans_inplace.cu:24-34
 *
 * Differences that are deliberate: 64-bit sizes; a memory plan instead of
 * the reference's chunk heuristic; every r1 chunk starts its contraction
 * fresh (the reference keeps beta = 1, src/build_int2_grad1_u12_ao.cu:166);
 * errors are returned instead of exit(0); kernels run back to back on one
 * stream with timing events read once at the end.
 */
#define _POSIX_C_SOURCE 200809L /* clock_gettime */
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include <unistd.h>

#include <hip/hip_runtime_api.h>

#include "cutc_int.h"
#include "internal.h"
#include "kernels/launch.h"
#include "xfer.h"

static double now_ms(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec * 1e3 + ts.tv_nsec * 1e-6;
}

/* CAT_LONGR: stage-1 pair weights (the long-range integrand); CAT_SHORTR:
 * stage-2 r1 integrands (herm and non-herm in one kernel); CAT_STAGE1/2:
 * each stage as a whole */
#define NCAT 10
enum { CAT_JAST = 0, CAT_LONGR, CAT_GEMM1, CAT_GEMM2, CAT_SYMM, CAT_SPLIT, CAT_TOTAL, CAT_SHORTR, CAT_STAGE1,
       CAT_STAGE2 };

typedef struct {
    hipStream_t st;
    int timing;
    int capture; /* recording into a graph: no stream syncs, no timing events */
    int nev;
    int cap;
    hipEvent_t *ev; /* pairs (start, stop) */
    int *cat;
    int launches[NCAT];
    /* called after the work writing rows [i0, i0 + n1e) of every X
     * component has been queued on st (host entry points: starts their
     * download) */
    void (*x_ready)(void *ctx, int64_t i0, int64_t n1e, hipStream_t st);
    void *x_ctx;
} tracer_t;

static void tr_init(tracer_t *t, hipStream_t st, int timing)
{
    memset(t, 0, sizeof *t);
    t->st = st;
    t->timing = timing;
}

/* events of category cat around work queued on stream s (tr_begin: on t->st) */
static int tr_begin_on(tracer_t *t, int cat, hipStream_t s)
{
    t->launches[cat]++;
    if (!t->timing)
        return -1;
    if (t->nev == t->cap) {
        int nc = t->cap ? 2 * t->cap : 64;
        hipEvent_t *ne = realloc(t->ev, sizeof(hipEvent_t) * 2 * nc);
        int *nk = realloc(t->cat, sizeof(int) * nc);
        if (!ne || !nk) {
            free(ne ? ne : t->ev);
            free(nk ? nk : t->cat);
            t->ev = NULL;
            t->cat = NULL;
            t->timing = 0;
            return -1;
        }
        t->ev = ne;
        t->cat = nk;
        t->cap = nc;
    }
    int i = t->nev;
    if (hipEventCreate(&t->ev[2 * i]) != hipSuccess)
        return -1;
    if (hipEventCreate(&t->ev[2 * i + 1]) != hipSuccess) {
        hipEventDestroy(t->ev[2 * i]);
        return -1;
    }
    t->cat[i] = cat;
    hipEventRecord(t->ev[2 * i], s);
    t->nev++;
    return i;
}

static void tr_end_on(tracer_t *t, int i, hipStream_t s)
{
    if (i >= 0)
        hipEventRecord(t->ev[2 * i + 1], s);
}

static int tr_begin(tracer_t *t, int cat) { return tr_begin_on(t, cat, t->st); }
static void tr_end(tracer_t *t, int i) { tr_end_on(t, i, t->st); }

/* After the stream is synchronised: per-category milliseconds. */
static void tr_collect(tracer_t *t, double *ms)
{
    for (int c = 0; c < NCAT; c++)
        ms[c] = 0.0;
    for (int i = 0; i < t->nev; i++) {
        float f = 0.0f;
        if (hipEventElapsedTime(&f, t->ev[2 * i], t->ev[2 * i + 1]) == hipSuccess)
            ms[t->cat[i]] += f;
    }
}

static void tr_free(tracer_t *t)
{
    for (int i = 0; i < t->nev; i++) {
        hipEventDestroy(t->ev[2 * i]);
        hipEventDestroy(t->ev[2 * i + 1]);
    }
    free(t->ev);
    free(t->cat);
    t->ev = NULL;
    t->cat = NULL;
    t->nev = t->cap = 0;
}

static int hip_rc(hipError_t e)
{
    if (e == hipSuccess)
        return CUTC_OK;
    if (e == hipErrorOutOfMemory)
        return CUTC_ERR_NOMEM;
    return CUTC_ERR_HIP;
}

#define TRY(x)                                                                                                       \
    do {                                                                                                             \
        rc = (x);                                                                                                    \
        if (rc != CUTC_OK)                                                                                           \
            goto done;                                                                                               \
    } while (0)
#define TRYH(x) TRY(hip_rc(x))

/* slices = 0: every precision takes its own default (cutc_default_slices) */
void cutc_default_options(cutc_options *opt)
{
    memset(opt, 0, sizeof *opt);
    opt->precision = CUTC_PREC_FP64;
    opt->slices = 0;
}

static int is_crt(int precision)
{
    return precision == CUTC_PREC_INT8 || precision == CUTC_PREC_BF16_CRT || precision == CUTC_PREC_FP16_CRT;
}

/* residue-plane kind of a modular precision */
static int crt_kind(int precision)
{
    return precision == CUTC_PREC_BF16_CRT ? CUTC_CRT_BF16
                                           : (precision == CUTC_PREC_FP16_CRT ? CUTC_CRT_FP16 : CUTC_CRT_I8);
}

static int valid_precision(const cutc_options *o)
{
    if (o->precision == CUTC_PREC_FP64)
        return 1;
    if (is_crt(o->precision) && o->slices >= 6 && o->slices <= CUTC_CRT_MAX)
        return 1;
    if ((o->precision == CUTC_PREC_BF16 || o->precision == CUTC_PREC_FP16) && o->slices >= 2 && o->slices <= 4)
        return 1;
    return 0;
}

/* a copy of *in with defaults filled in (slices = 0 -> the precision's default) */
static cutc_options resolved_options(const cutc_options *in)
{
    cutc_options o = *in;
    if (o.slices == 0)
        o.slices = cutc_default_slices(o.precision);
    return o;
}

/* Boys-Handy exponents must be non-negative; a^b is evaluated by repeated
 * products, so large exponents are refused as meaningless (> 64). */
static int check_exponents(const int *m, const int *n, const int *o, int64_t cnt)
{
    for (int64_t i = 0; i < cnt; i++)
        if (m[i] < 0 || n[i] < 0 || o[i] < 0 || m[i] > 64 || n[i] > 64 || o[i] > 64)
            return CUTC_ERR_ARG;
    return CUTC_OK;
}

/* Copy the (small) Boys-Handy tables to the host, check the exponents and
 * expand them into the Jastrow monomial program. */
static int program_from_device(int n_nuc, int size_bh, const double *dc, const int *dm, const int *dn,
                               const int *dno, cutc_jprog *prog)
{
    const int64_t cnt = (int64_t)n_nuc * size_bh;
    memset(prog, 0, sizeof *prog);
    double *c = malloc(sizeof(double) * (cnt + 1));
    int *h = malloc(sizeof(int) * (3 * cnt + 1));
    int rc = (c && h) ? CUTC_OK : CUTC_ERR_NOMEM;
    if (rc == CUTC_OK && cnt > 0 &&
        (hipMemcpy(c, dc, sizeof(double) * cnt, hipMemcpyDeviceToHost) != hipSuccess ||
         hipMemcpy(h, dm, sizeof(int) * cnt, hipMemcpyDeviceToHost) != hipSuccess ||
         hipMemcpy(h + cnt, dn, sizeof(int) * cnt, hipMemcpyDeviceToHost) != hipSuccess ||
         hipMemcpy(h + 2 * cnt, dno, sizeof(int) * cnt, hipMemcpyDeviceToHost) != hipSuccess))
        rc = CUTC_ERR_HIP;
    if (rc == CUTC_OK)
        rc = check_exponents(h, h + cnt, h + 2 * cnt, cnt);
    if (rc == CUTC_OK)
        rc = cutc_jast_program(n_nuc, size_bh, c, h, h + cnt, h + 2 * cnt, prog);
    free(c);
    free(h);
    return rc;
}

/* A Jastrow program with its device copies (monomials and classes).
 * borrowed: host and device arrays live in static host memory and the
 * device scratch (the allocation-free entry point) and are not freed. */
typedef struct {
    cutc_jprog h;
    cutc_jmono *dmono;
    int *dcls;
    int borrowed;
} jprog_t;

/* Device copies of a program's monomials and classes. */
static int program_upload(const cutc_jprog *prog, cutc_jmono **dmono, int **dcls)
{
    *dmono = NULL;
    *dcls = NULL;
    int rc = hip_rc(hipMalloc((void **)dmono, sizeof(cutc_jmono) * (prog->n_mono + 1)));
    if (rc == CUTC_OK)
        rc = hip_rc(hipMalloc((void **)dcls, sizeof(int) * (2 * prog->n_cls + 2)));
    if (rc == CUTC_OK && prog->n_mono > 0)
        rc = hip_rc(hipMemcpy(*dmono, prog->mono, sizeof(cutc_jmono) * prog->n_mono, hipMemcpyHostToDevice));
    if (rc == CUTC_OK && prog->n_cls > 0)
        rc = hip_rc(hipMemcpy(*dcls, prog->cls, sizeof(int) * 2 * prog->n_cls, hipMemcpyHostToDevice));
    return rc;
}

/* Program of a system: decoded from the device tables and uploaded (need = 0:
 * nothing, for calls that run no Jastrow). */
static int jprog_load(int n_nuc, int size_bh, const double *c_bh, const int *m_bh, const int *n_bh, const int *o_bh,
                      int need, jprog_t *jp)
{
    memset(jp, 0, sizeof *jp);
    if (!need)
        return CUTC_OK;
    int rc = program_from_device(n_nuc, size_bh, c_bh, m_bh, n_bh, o_bh, &jp->h);
    if (rc == CUTC_OK)
        rc = program_upload(&jp->h, &jp->dmono, &jp->dcls);
    return rc;
}

static void dfree(void *p);

static void jprog_free(jprog_t *jp)
{
    if (!jp->borrowed) {
        dfree(jp->dmono);
        dfree(jp->dcls);
        cutc_jast_program_free(&jp->h);
    }
    memset(jp, 0, sizeof *jp);
}

/* jprog_load without allocating: the Boys-Handy tables are copied into
 * static host arrays, expanded by cutc_jast_program_static, and the
 * program's monomials and classes copied to the front of the device
 * region base (*used: bytes taken there, 256-byte multiples). */
static double g_nat_c[CUTC_JPROG_STATIC_TERMS];
static int g_nat_h[3 * CUTC_JPROG_STATIC_TERMS];

static int jprog_load_static(int n_nuc, int size_bh, const double *dc, const int *dm, const int *dn,
                             const int *dno, int need, char *base, size_t cap, size_t *used, jprog_t *jp)
{
    memset(jp, 0, sizeof *jp);
    jp->borrowed = 1;
    *used = 0;
    if (!need)
        return CUTC_OK;
    const int64_t cnt = (int64_t)n_nuc * size_bh;
    if (n_nuc < 0 || size_bh < 0 || cnt > CUTC_JPROG_STATIC_TERMS)
        return CUTC_ERR_ARG;
    if (cnt > 0 &&
        (hipMemcpy(g_nat_c, dc, sizeof(double) * cnt, hipMemcpyDeviceToHost) != hipSuccess ||
         hipMemcpy(g_nat_h, dm, sizeof(int) * cnt, hipMemcpyDeviceToHost) != hipSuccess ||
         hipMemcpy(g_nat_h + cnt, dn, sizeof(int) * cnt, hipMemcpyDeviceToHost) != hipSuccess ||
         hipMemcpy(g_nat_h + 2 * cnt, dno, sizeof(int) * cnt, hipMemcpyDeviceToHost) != hipSuccess))
        return CUTC_ERR_HIP;
    int rc = check_exponents(g_nat_h, g_nat_h + cnt, g_nat_h + 2 * cnt, cnt);
    if (rc == CUTC_OK)
        rc = cutc_jast_program_static(n_nuc, size_bh, g_nat_c, g_nat_h, g_nat_h + cnt, g_nat_h + 2 * cnt, &jp->h);
    if (rc != CUTC_OK)
        return rc;
    const size_t bm = (sizeof(cutc_jmono) * (jp->h.n_mono + 1) + 255) & ~(size_t)255;
    const size_t bc = (sizeof(int) * (2 * jp->h.n_cls + 2) + 255) & ~(size_t)255;
    if (bm + bc > cap)
        return CUTC_ERR_NOMEM;
    jp->dmono = (cutc_jmono *)base;
    jp->dcls = (int *)(base + bm);
    *used = bm + bc;
    if (jp->h.n_mono > 0)
        rc = hip_rc(hipMemcpy(jp->dmono, jp->h.mono, sizeof(cutc_jmono) * jp->h.n_mono, hipMemcpyHostToDevice));
    if (rc == CUTC_OK && jp->h.n_cls > 0)
        rc = hip_rc(hipMemcpy(jp->dcls, jp->h.cls, sizeof(int) * 2 * jp->h.n_cls, hipMemcpyHostToDevice));
    return rc;
}

static size_t arena_size_current(void);

/* Set by the allocation-free entry point while it runs (under the library
 * lock): the stages carve their workspace from this device region instead
 * of the arena, and the memory plan is made against its size. */
static struct {
    char *base;
    size_t size;
} g_ws_fixed;

/* free device memory plus this device's cached workspace, minus a margin */
/* device bytes a call may plan with: free memory plus the workspace arena
 * (and `held`, memory of the library's the call will reuse), less a 1 GiB
 * margin, at most mem_limit */
static size_t device_budget_plus(const cutc_options *o, size_t held)
{
    if (g_ws_fixed.base)
        return (o && o->mem_limit && o->mem_limit < g_ws_fixed.size) ? o->mem_limit : g_ws_fixed.size;
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) != hipSuccess)
        return 0;
    fr += arena_size_current() + held;
    const size_t margin = (size_t)1 << 30;
    size_t b = fr > margin ? fr - margin : 0;
    if (o && o->mem_limit && o->mem_limit < b)
        b = o->mem_limit;
    return b;
}

static size_t device_budget(const cutc_options *o) { return device_budget_plus(o, 0); }

static void dfree(void *p)
{
    if (p)
        hipFree(p);
}

/* Device workspace: one grow-only arena per device, kept between calls of
 * the additive device entry point so that repeated builds allocate nothing
 * (the reference allocates and frees every buffer in every call).  A stage
 * sums its buffer sizes, makes sure the arena is at least that large, and
 * carves the buffers with a 256-byte-aligned bump pointer.  All calls are
 * serialised by g_lock because they share the arenas. */
#define WS_MAX_DEV 64
static struct {
    char *base;
    size_t size;
} g_arena[WS_MAX_DEV];
static pthread_mutex_t g_lock = PTHREAD_MUTEX_INITIALIZER;

typedef struct {
    char *base;
    size_t off, cap;
} ws_t;

static size_t ws_round(size_t b) { return (b + 255) & ~(size_t)255; }

static size_t arena_size_current(void)
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= WS_MAX_DEV)
        return 0;
    return g_arena[dev].size;
}

/* set while a graph is being recorded: the arena may not be reallocated */
static int g_ws_frozen = 0;

static int ws_begin(ws_t *w, size_t need)
{
    int dev = 0;
    memset(w, 0, sizeof *w);
    if (g_ws_fixed.base) {
        if (need > g_ws_fixed.size)
            return CUTC_ERR_NOMEM;
        w->base = g_ws_fixed.base;
        w->cap = g_ws_fixed.size;
        return CUTC_OK;
    }
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= WS_MAX_DEV)
        return CUTC_ERR_HIP;
    if (g_arena[dev].size < need && g_ws_frozen)
        return CUTC_ERR_NOMEM;
    if (g_arena[dev].size < need) {
        if (g_arena[dev].base) {
            hipDeviceSynchronize();
            hipFree(g_arena[dev].base);
        }
        g_arena[dev].base = NULL;
        g_arena[dev].size = 0;
        void *p = NULL;
        int rc = hip_rc(hipMalloc(&p, need));
        if (rc)
            return rc;
        g_arena[dev].base = p;
        g_arena[dev].size = need;
    }
    w->base = g_arena[dev].base;
    w->cap = g_arena[dev].size;
    return CUTC_OK;
}

static void *ws_take(ws_t *w, size_t bytes)
{
    bytes = ws_round(bytes);
    if (w->off + bytes > w->cap)
        return NULL;
    void *p = w->base + w->off;
    w->off += bytes;
    return p;
}

/* Internal interface for the other host files of the library (internal.h):
 * the library lock, the device workspace arena and the memory budget. */
void cutc_internal_lock(void) { pthread_mutex_lock(&g_lock); }
void cutc_internal_unlock(void) { pthread_mutex_unlock(&g_lock); }
size_t cutc_internal_budget(size_t mem_limit)
{
    cutc_options o;
    cutc_default_options(&o);
    o.mem_limit = mem_limit;
    return device_budget(&o);
}
int cutc_internal_workspace(size_t need, char **base)
{
    ws_t w;
    int rc = ws_begin(&w, need);
    *base = (rc == CUTC_OK) ? w.base : NULL;
    return rc;
}

/* Device copies of the host entry points' arrays (inputs, X, int_2e_ao): a
 * second grow-only allocation per device, kept and released with the
 * workspace, so that repeated cutc_int_ex calls do not allocate (a fresh
 * 90 GB hipMalloc beside a full workspace took 2.5 s).  The streamed build's
 * X chunk and kept stage-1 orderings: a third (42 GB on the 40-nucleus
 * workload, ~1 s to allocate and free per call). */
typedef struct {
    char *base;
    size_t size;
} devbuf_t;
static devbuf_t g_hbuf[WS_MAX_DEV], g_sbuf[WS_MAX_DEV];

static void free_on_device(int d, char **base, size_t *size)
{
    if (*base) {
        int cur = 0;
        hipGetDevice(&cur);
        hipSetDevice(d);
        hipDeviceSynchronize();
        hipFree(*base);
        hipSetDevice(cur);
    }
    *base = NULL;
    *size = 0;
}

static void release_arenas(void)
{
    for (int d = 0; d < WS_MAX_DEV; d++) {
        free_on_device(d, &g_arena[d].base, &g_arena[d].size);
        free_on_device(d, &g_hbuf[d].base, &g_hbuf[d].size);
        free_on_device(d, &g_sbuf[d].base, &g_sbuf[d].size);
    }
}

/* buffer bufs[current device], at least need bytes (shrink: and at most
 * 2 need, reallocated otherwise); when it has to be allocated and the
 * device is full, the workspace is given back first */
static int buf_get(devbuf_t *bufs, size_t need, int shrink, char **base)
{
    int dev = 0;
    *base = NULL;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= WS_MAX_DEV)
        return CUTC_ERR_HIP;
    if (bufs[dev].size < need || (shrink && bufs[dev].size / 2 > need)) {
        free_on_device(dev, &bufs[dev].base, &bufs[dev].size);
        void *p = NULL;
        hipError_t e = hipMalloc(&p, need);
        if (e == hipErrorOutOfMemory) {
            (void)hipGetLastError();
            free_on_device(dev, &g_arena[dev].base, &g_arena[dev].size);
            e = hipMalloc(&p, need);
        }
        if (e != hipSuccess)
            return hip_rc(e);
        bufs[dev].base = (char *)p;
        bufs[dev].size = need;
    }
    *base = bufs[dev].base;
    return CUTC_OK;
}

static int hbuf_get(size_t need, char **base) { return buf_get(g_hbuf, need, 0, base); }

/* bytes of bufs[current device] */
static size_t buf_size_current(const devbuf_t *bufs)
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= WS_MAX_DEV)
        return 0;
    return bufs[dev].size;
}

void cutc_internal_release_workspace(void) { release_arenas(); }

void cutc_release_workspace(void)
{
    pthread_mutex_lock(&g_lock);
    release_arenas();
    pthread_mutex_unlock(&g_lock);
}

#define WS_TAKE(ptr, bytes)                                                                                         \
    do {                                                                                                             \
        *(void **)&(ptr) = ws_take(&ws, (bytes));                                                                    \
        if (!(ptr)) {                                                                                                \
            rc = CUTC_ERR_NOMEM;                                                                                     \
            goto done;                                                                                               \
        }                                                                                                            \
    } while (0)

/* A second stream per device (non-blocking, default priority) for the
 * Jastrow of the next r1 chunk of a split stage 1, and the two events that
 * order it against the build stream; made on first use, kept until the
 * process ends. */
static struct {
    hipStream_t st;
    hipEvent_t eq, ej;
} g_side[WS_MAX_DEV];

static int side_stream(hipStream_t *s, hipEvent_t *eq, hipEvent_t *ej)
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= WS_MAX_DEV)
        return CUTC_ERR_HIP;
    if (!g_side[dev].st) {
        hipStream_t x = NULL;
        hipEvent_t a = NULL, b = NULL;
        int rc = hip_rc(hipStreamCreateWithFlags(&x, hipStreamNonBlocking));
        if (rc == CUTC_OK)
            rc = hip_rc(hipEventCreateWithFlags(&a, hipEventDisableTiming));
        if (rc == CUTC_OK)
            rc = hip_rc(hipEventCreateWithFlags(&b, hipEventDisableTiming));
        if (rc != CUTC_OK) {
            if (a)
                hipEventDestroy(a);
            if (b)
                hipEventDestroy(b);
            if (x)
                hipStreamDestroy(x);
            hipGetLastError(); /* the build goes on in order on its own stream */
            return rc;
        }
        g_side[dev].st = x;
        g_side[dev].eq = a;
        g_side[dev].ej = b;
    }
    *s = g_side[dev].st;
    *eq = g_side[dev].eq;
    *ej = g_side[dev].ej;
    return CUTC_OK;
}

/* ------------------------------------------------------------ stage 1 */

static int stage1(const cutc_options *opt, const cutc_plan *pl, int64_t n1, int64_t n2, int n_ao, int n_nuc,
                  const double *r1, const double *r2, const double *wr2, const double *rn, const double *aos2,
                  const jprog_t *jp, double *X, tracer_t *tr, cutc_stats *stats)
{
    const hipStream_t st = tr->st;
    const int64_t P = pl->n_pairs, ldk = pl->ldk, n1c = pl->n1_chunk;
    const cutc_jprog *prog = &jp->h;
    const cutc_jmono *dmono = jp->dmono;
    const int *dcls = jp->dcls;
    int *pairs = NULL;
    double *Apk = NULL, *G = NULL, *f2tab = NULL, *tab = NULL;
    int rc = CUTC_OK, ev;
    (void)opt;

    ws_t ws;
    {
        const size_t need = ws_round(sizeof(int) * 2 * P) + ws_round(sizeof(double) * P * ldk) +
                            ws_round(sizeof(double) * 4 * n1c * ldk) +
                            (n_nuc > 32 ? ws_round(sizeof(double) * n_nuc * ldk) : 0) +
                            ws_round(sizeof(double) * (prog->ts * n1c + 1));
        TRY(ws_begin(&ws, need));
        if (stats)
            stats->workspace_bytes = need;
    }
    WS_TAKE(pairs, sizeof(int) * 2 * P);
    WS_TAKE(Apk, sizeof(double) * P * ldk);
    WS_TAKE(G, sizeof(double) * 4 * n1c * ldk);
    if (n_nuc > 32)
        WS_TAKE(f2tab, sizeof(double) * n_nuc * ldk);
    WS_TAKE(tab, sizeof(double) * (prog->ts * n1c + 1));

    TRY(cutc_launch_pair_table(n_ao, pairs, st));
    ev = tr_begin(tr, CAT_LONGR);
    rc = cutc_launch_pair_weights(wr2, aos2, n2, 0, n2, ldk, n_ao, Apk, ldk, st);
    tr_end(tr, ev);
    TRY(rc);

    int nchunks = 0;
    for (int64_t i0 = 0; i0 < n1; i0 += n1c) {
        const int64_t n1e = (n1 - i0 < n1c) ? n1 - i0 : n1c;
        nchunks++;
        ev = tr_begin(tr, CAT_JAST);
        rc = cutc_launch_jast_prep(r1, i0, n1e, rn, n_nuc, dmono, prog->n_mono, tab, prog->ts, st);
        if (rc == CUTC_OK)
            rc = cutc_launch_jast_pairs(r1, r2, rn, n_nuc, tab, prog->ts, dcls, prog->n_cls, prog->kmax, i0, n1e, 0,
                                        n2, ldk, G, ldk, n1e * ldk, f2tab, st);
        tr_end(tr, ev);
        TRY(rc);
        ev = tr_begin(tr, CAT_GEMM1);
        rc = cutc_gemm1_f64(Apk, ldk, G, ldk, P, n1e, ldk, pairs, n_ao, X, n1, i0, 0, st);
        tr_end(tr, ev);
        TRY(rc);
        if (tr->x_ready)
            tr->x_ready(tr->x_ctx, i0, n1e, st);
        if (stats) {
            /* executed contraction work, FP64-equivalent: 2 x pairs x (4 n1e) x n2 */
            stats->flops_exec += 2.0 * (double)P * 4.0 * (double)n1e * (double)n2;
        }
    }
    if (stats) {
        stats->n1_chunk = n1c;
        stats->n1_chunks = nchunks;
    }
    if (!tr->capture)
        TRYH(hipStreamSynchronize(st));
done:
    if (rc != CUTC_OK && !tr->capture)
        hipStreamSynchronize(st);
    return rc;
}

/* sum of the n step counts of the lists (device), after the work queued on st */
static int kcount_sum(const int *kcount, int64_t n, hipStream_t st, double *sum)
{
    int *h = (int *)malloc(sizeof(int) * (n > 0 ? n : 1));
    if (!h)
        return CUTC_ERR_NOMEM;
    int rc = hipMemcpyAsync(h, kcount, sizeof(int) * n, hipMemcpyDeviceToHost, st) == hipSuccess &&
                     hipStreamSynchronize(st) == hipSuccess
                 ? CUTC_OK
                 : CUTC_ERR_HIP;
    double s = 0.0;
    for (int64_t i = 0; i < n && rc == CUTC_OK; i++)
        s += h[i];
    free(h);
    *sum = s;
    return rc;
}

/* The r1-independent part of the block-sparse stage 1 (the permuted copies
 * of r2, wr2 and the AO values, the clustered pair table and its exponents),
 * kept across the r1 chunks of one streamed call: the first chunk computes
 * it into these buffers, the others reuse it. */
typedef struct s1pre {
    double *r2p, *wr2p, *aos2p;
    int *pairs, *rexpA;
    int valid;
} s1pre_t;

static size_t s1pre_bytes(int64_t n2, int n_ao)
{
    const int64_t P = (int64_t)n_ao * (n_ao + 1) / 2;
    return ws_round(sizeof(double) * 3 * n2) + ws_round(sizeof(double) * n2) +
           ws_round(sizeof(double) * n_ao * n2) + ws_round(sizeof(int) * 2 * P) + ws_round(sizeof(int) * P);
}

/* carve the buffers of s1pre_bytes(n2, n_ao) bytes at base */
static void s1pre_carve(s1pre_t *pre, char *base, int64_t n2, int n_ao)
{
    const int64_t P = (int64_t)n_ao * (n_ao + 1) / 2;
    pre->r2p = (double *)base;
    base += ws_round(sizeof(double) * 3 * n2);
    pre->wr2p = (double *)base;
    base += ws_round(sizeof(double) * n2);
    pre->aos2p = (double *)base;
    base += ws_round(sizeof(double) * n_ao * n2);
    pre->pairs = (int *)base;
    base += ws_round(sizeof(int) * 2 * P);
    pre->rexpA = (int *)base;
    pre->valid = 0;
}

/* What the Jastrow of one r1 chunk of stage 1 reads and writes. */
typedef struct {
    const jprog_t *jp;
    const double *r1, *r2, *rn;
    int n_nuc;
    int64_t n2, ldk;
    double *tab, *G, *gpart, *f2tab;
} jchunk_t;

/* The Jastrow of r1 points [i0, i0 + n1e) on stream s: the per-point table,
 * then the 4 G rows of every point (row stride ldk, component stride
 * n1e ldk).  With gpart (modular) the staged pair kernel also writes the
 * partial row sums of squares and *norms is set; otherwise the caller takes
 * the row exponents from G. */
static int jast_chunk(const jchunk_t *a, int64_t i0, int64_t n1e, hipStream_t s, int *norms)
{
    const cutc_jprog *prog = &a->jp->h;
    int rc = cutc_launch_jast_prep(a->r1, i0, n1e, a->rn, a->n_nuc, a->jp->dmono, prog->n_mono, a->tab, prog->ts, s);
    *norms = 0;
    if (rc == CUTC_OK && a->gpart) {
        rc = cutc_launch_jast_pairs_norms(a->r1, a->r2, a->rn, a->n_nuc, a->tab, prog->ts, a->jp->dcls, prog->n_cls,
                                          prog->kmax, i0, n1e, a->n2, a->ldk, a->G, a->ldk, n1e * a->ldk, a->gpart, s);
        *norms = rc == CUTC_OK;
        if (rc == CUTC_ERR_UNSUPPORTED)
            rc = CUTC_OK;
    }
    if (rc == CUTC_OK && !*norms)
        rc = cutc_launch_jast_pairs(a->r1, a->r2, a->rn, a->n_nuc, a->tab, prog->ts, a->jp->dcls, prog->n_cls,
                                    prog->kmax, i0, n1e, 0, a->n2, a->ldk, a->G, a->ldk, n1e * a->ldk, a->f2tab, s);
    return rc;
}

/* jast_chunk on the side stream js, after the work queued so far on the
 * build stream; ej marks its end */
static int jast_queue(tracer_t *tr, const jchunk_t *a, int64_t i0, int64_t n1e, hipStream_t js, hipEvent_t eq,
                      hipEvent_t ej, int *norms)
{
    int rc = hip_rc(hipEventRecord(eq, tr->st));
    if (rc == CUTC_OK)
        rc = hip_rc(hipStreamWaitEvent(js, eq, 0));
    if (rc == CUTC_OK) {
        const int ev = tr_begin_on(tr, CAT_JAST, js);
        rc = jast_chunk(a, i0, n1e, js, norms);
        tr_end_on(tr, ev, js);
    }
    if (rc == CUTC_OK)
        rc = hip_rc(hipEventRecord(ej, js));
    return rc;
}

/* Stage 1 with a split-precision contraction.  INT8 / BF16_CRT / FP16_CRT:
 * the modular (CRT) scheme, pair residues quantised on the fly from w2 and
 * chi, each G chunk quantised after the Jastrow block, one int8 (or bf16 /
 * fp16) GEMM per modulus, CRT rebuild scattered into X.  BF16 / FP16: the slice scheme, digit planes of the
 * pairs and of each G chunk, one slice GEMM into FP64, scattered into X.
 * Modular: block-sparse.  The r2 points are reordered along a Morton
 * curve (the Jastrow, pair weights and AO values all use the permuted
 * copies, so the contraction over r2 is unchanged), the pair rows clustered
 * into 256-pair tiles by the K steps where their quantised values are
 * non-zero, and every GEMM runs, per tile, only the K steps where that tile
 * has a non-zero residue (opt->dense: all of them); pre (optional) keeps the
 * orderings across the r1 chunks of a streamed call.
 * With one K pass the Jastrow runs on a second stream, one chunk ahead: that
 * of the first chunk beside the pair orderings (it reads only the permuted
 * r2), that of chunk c + 1 once chunk c's G has been quantised, beside chunk
 * c's contraction (the FP VALU work of the one against the memory-bound
 * GEMMs of the other); chunk c + 1's quantiser waits for it.  Not under
 * graph capture or in the allocation-free entry point. */
static int stage1_split(const cutc_options *opt, const cutc_plan *pl, int64_t n1, int64_t n2, int n_ao, int n_nuc,
                        const double *r1, const double *r2, const double *wr2, const double *rn, const double *aos2,
                        const jprog_t *jp, double *X, tracer_t *tr, cutc_stats *stats, s1pre_t *pre)
{
    const cutc_jprog *prog = &jp->h;
    const hipStream_t st = tr->st;
    hipStream_t js = st; /* the stream of the Jastrow */
    hipEvent_t eq = NULL, ej = NULL;
    const int64_t P = pl->n_pairs, ldk = pl->ldk, n1c = pl->n1_chunk;
    const int nmod = opt->slices;
    int *pairs = NULL;
    double *G = NULL, *f2tab = NULL, *tab = NULL;
    int *rexpA = NULL, *rexpG = NULL, *rexpN = NULL, *kcount = NULL;
    int8_t *Apl = NULL, *Gpl = NULL, *Cpl = NULL;
    double *gpart = NULL; /* the Jastrow kernel's partial row sums of G (modular) */
    uint8_t *occ = NULL;
    uint32_t *klist = NULL, *svals = NULL;
    uint64_t *skeys = NULL;
    double *r2p = NULL, *wr2p = NULL, *aos2p = NULL, *box = NULL;
    void *cws = NULL;
    int rc = CUTC_OK, ev;
    const int slice = !is_crt(opt->precision), f16 = opt->precision == CUTC_PREC_FP16;
    const int emode = slice ? CUTC_ROWEXP_MAX : CUTC_ROWEXP_NORM;
    const size_t rb = CUTC_CRT_RBYTES(crt_kind(opt->precision));
    cutc_crt crt;
    if (!slice)
        TRY(cutc_crt_setup(n2, nmod, crt_kind(opt->precision), &crt));
    /* modular: residue planes (int8, or 16-bit for bf16 / fp16) and int8
     * results; slices: 16-bit digit planes and FP64 results.  The modular
     * planes hold one K pass of kc r2 points (kc = ldk: a single pass); the
     * GEMM of each further pass adds its residues into the results. */
    const int64_t kc = (slice || pl->k_chunk <= 0 || pl->k_chunk > ldk) ? ldk : pl->k_chunk;
    const int64_t kpass = (n2 + kc - 1) / kc;
    const size_t bA = slice ? (size_t)nmod * 2 * CUTC_OZ_ROWS(P) * ldk : rb * nmod * CUTC_PK_ROWS_B(P) * kc;
    const size_t bG = slice ? (size_t)nmod * 2 * CUTC_OZ_ROWS(4 * n1c) * ldk
                            : rb * nmod * CUTC_PK_ROWS_A(4 * n1c) * kc;
    const size_t bC = slice ? sizeof(double) * 4 * n1c * P : (size_t)nmod * 4 * n1c * P;
    /* block-sparse state (the planner's sparse_bytes) */
    const int sparse = !slice && n2 > 0 && P > 0;
    const int kind = slice ? 0 : crt_kind(opt->precision);
    /* KT: the steps of all r2 points (pair supports); KTp: of one K pass (flags and lists) */
    const int64_t tiles = CUTC_PK_ROWS_B(P) / CUTC_SP_ROWS, KT = sparse ? ldk / CUTC_CRT_KSTEP(kind) : 0;
    const int64_t KTp = sparse ? kc / CUTC_CRT_KSTEP(kind) : 0;
    const int64_t nsort = sparse ? cutc_sort_len(n2) : 0;

    ws_t ws;
    {
        size_t need = ws_round(sizeof(int) * 2 * P) + ws_round(sizeof(int) * P) +
                      ws_round(sizeof(int) * 4 * n1c) + ws_round(bA) + ws_round(bG) + ws_round(bC) +
                      ws_round(sizeof(double) * 4 * n1c * ldk) +
                      (n_nuc > 32 ? ws_round(sizeof(double) * n_nuc * ldk) : 0) +
                      ws_round(sizeof(double) * (prog->ts * n1c + 1));
        if (!slice)
            need += ws_round(sizeof(double) * 4 * n1c * cutc_jast_norm_blocks(ldk));
        if (sparse)
            need += ws_round(sizeof(double) * 3 * n2) + ws_round(sizeof(double) * n2) +
                    ws_round(sizeof(double) * n_ao * n2) + ws_round(sizeof(uint64_t) * nsort) +
                    ws_round(sizeof(uint32_t) * nsort) + ws_round(sizeof(int) * P) + ws_round(tiles * KTp) +
                    ws_round(sizeof(uint32_t) * tiles * KTp) + ws_round(sizeof(int) * tiles) +
                    ws_round(sizeof(double) * 6) + ws_round(cutc_cluster_bytes(P, KT));
        TRY(ws_begin(&ws, need));
        if (stats)
            stats->workspace_bytes = need;
    }
    WS_TAKE(pairs, sizeof(int) * 2 * P);
    WS_TAKE(rexpA, sizeof(int) * P);
    WS_TAKE(rexpG, sizeof(int) * 4 * n1c);
    WS_TAKE(Apl, bA);
    WS_TAKE(Gpl, bG);
    WS_TAKE(Cpl, bC);
    WS_TAKE(G, sizeof(double) * 4 * n1c * ldk);
    if (n_nuc > 32)
        WS_TAKE(f2tab, sizeof(double) * n_nuc * ldk);
    WS_TAKE(tab, sizeof(double) * (prog->ts * n1c + 1));
    if (!slice)
        WS_TAKE(gpart, sizeof(double) * 4 * n1c * cutc_jast_norm_blocks(ldk));
    if (sparse) {
        WS_TAKE(r2p, sizeof(double) * 3 * n2);
        WS_TAKE(wr2p, sizeof(double) * n2);
        WS_TAKE(aos2p, sizeof(double) * n_ao * n2);
        WS_TAKE(skeys, sizeof(uint64_t) * nsort);
        WS_TAKE(svals, sizeof(uint32_t) * nsort);
        WS_TAKE(rexpN, sizeof(int) * P);
        WS_TAKE(occ, tiles * KTp);
        WS_TAKE(klist, sizeof(uint32_t) * tiles * KTp);
        WS_TAKE(kcount, sizeof(int) * tiles);
        WS_TAKE(box, sizeof(double) * 6);
        WS_TAKE(cws, cutc_cluster_bytes(P, KT));
    }
    /* the Jastrow one chunk ahead on the side stream js (js == st: in order) */
    if (kpass == 1 && n1 > 0 && (sparse || n1 > n1c) && !tr->capture && !g_ws_fixed.base &&
        side_stream(&js, &eq, &ej) != CUTC_OK)
        js = st;
    jchunk_t jc = {jp, r1, r2, rn, n_nuc, n2, ldk, tab, G, gpart, f2tab}; /* r2: the permuted copy once made */
    int queued = 0, norms = 0; /* the next chunk's Jastrow is on js; its G row sums are written */

    ev = tr_begin(tr, CAT_SPLIT);
    if (sparse && pre) {
        /* the orderings of an earlier chunk of the call, or computed here for the later ones */
        r2p = pre->r2p;
        wr2p = pre->wr2p;
        aos2p = pre->aos2p;
        pairs = pre->pairs;
        rexpA = pre->rexpA;
    }
    if (sparse) {
        /* r2 order, then everything of stage 1 reads the permuted copies */
        if (!(pre && pre->valid))
            rc = cutc_order_r2(r2, wr2, aos2, n2, n_ao, box, skeys, svals, nsort, r2p, wr2p, aos2p, st);
        r2 = r2p;
        wr2 = wr2p;
        aos2 = aos2p;
        jc.r2 = r2;
        if (rc == CUTC_OK && js != st) {
            rc = jast_queue(tr, &jc, 0, n1 < n1c ? n1 : n1c, js, eq, ej, &norms);
            queued = rc == CUTC_OK;
        }
        if (rc == CUTC_OK && !(pre && pre->valid))
            rc = cutc_rowexp_pairw(wr2, aos2, n2, n_ao, emode, rexpN, st);
        if (rc == CUTC_OK && !(pre && pre->valid))
            rc = cutc_order_pairs(wr2, aos2, n2, n_ao, crt.bits, rexpN, CUTC_CRT_KSTEP(kind), KT, cws, pairs, rexpA,
                                  st);
        if (rc == CUTC_OK && pre)
            pre->valid = 1;
        if (rc == CUTC_OK && kpass == 1)
            rc = cutc_zero_u8(occ, tiles * KT, st);
        if (rc == CUTC_OK && kpass == 1)
            rc = cutc_crt_quant_pairw_pk_occ(&crt, wr2, aos2, n2, n2, n_ao, ldk, rexpA, pairs, Apl, occ, st);
        if (rc == CUTC_OK && kpass == 1)
            rc = cutc_klists(occ, tiles, KT, CUTC_CRT_LIST_ALIGN(kind), klist, kcount, st);
    } else {
        rc = cutc_launch_pair_table(n_ao, pairs, st);
        if (rc == CUTC_OK)
            rc = cutc_rowexp_pairw(wr2, aos2, n2, n_ao, emode, rexpA, st);
        if (rc == CUTC_OK && kpass == 1)
            rc = slice ? cutc_oz_slice_pairw(nmod, f16, wr2, aos2, n2, n_ao, ldk, CUTC_OZ_ROWS(P), rexpA,
                                             (uint16_t *)Apl, st)
                       : cutc_crt_quant_pairw_pk(&crt, wr2, aos2, n2, n_ao, ldk, rexpA, Apl, st);
    }
    tr_end(tr, ev);
    TRY(rc);
    const uint32_t *kl = (sparse && !opt->dense) ? klist : NULL;
    const int *kc_ = (sparse && !opt->dense) ? kcount : NULL;
    double dens_sum = 0.0, dens_den = 0.0; /* executed tiles of the passes of the first chunk */

    int nchunks = 0;
    for (int64_t i0 = 0; i0 < n1; i0 += n1c) {
        const int64_t n1e = (n1 - i0 < n1c) ? n1 - i0 : n1c;
        nchunks++;
        /* modular: the pair kernel also sums the squares of the G rows it
         * writes (their 2-norm exponents), where its staged form runs */
        if (queued) {
            rc = hip_rc(hipStreamWaitEvent(st, ej, 0));
            queued = 0;
        } else {
            ev = tr_begin(tr, CAT_JAST);
            rc = jast_chunk(&jc, i0, n1e, st, &norms);
            tr_end(tr, ev);
        }
        TRY(rc);
        ev = tr_begin(tr, CAT_SPLIT);
        rc = norms ? cutc_rowexp_finish(gpart, cutc_jast_norm_blocks(ldk), 4 * n1e, CUTC_ROWEXP_NORM, rexpG, st)
                   : cutc_rowexp_k(G, ldk, 4 * n1e, n2, emode, rexpG, st);
        if (rc == CUTC_OK && kpass == 1)
            rc = slice ? cutc_oz_slice_k(nmod, f16, G, ldk, 4 * n1e, n2, ldk, CUTC_OZ_ROWS(4 * n1e), rexpG,
                                         (uint16_t *)Gpl, st)
                       : cutc_crt_quant_k_pk(&crt, G, ldk, 4 * n1e, n2, ldk, rexpG, Gpl, st);
        tr_end(tr, ev);
        TRY(rc);
        /* G, its row sums and the point table are free: the next chunk's
         * Jastrow goes on js, beside this chunk's contraction */
        if (js != st && i0 + n1c < n1) {
            const int64_t i1 = i0 + n1c;
            TRY(jast_queue(tr, &jc, i1, n1 - i1 < n1c ? n1 - i1 : n1c, js, eq, ej, &norms));
            queued = 1;
        }
        if (kpass == 1) {
            ev = tr_begin(tr, CAT_GEMM1);
            rc = slice ? cutc_oz_gemm(nmod, f16, (const uint16_t *)Gpl, 4 * n1e, (const uint16_t *)Apl, P, ldk, rexpG,
                                      rexpA, (double *)Cpl, st)
                       : cutc_crt_gemm_pk_list(&crt, Gpl, 4 * n1e, Apl, P, ldk, kl, kc_, 0, 1, Cpl, 0, st);
            tr_end(tr, ev);
            TRY(rc);
        }
        /* several K passes (modular only): the residue planes of both
         * operands for r2 points [k0, k0 + kn) (block-sparse: with the flags
         * and step lists of the pass), products added into Cpl */
        for (int64_t kp = 0; kp < kpass && kpass > 1; kp++) {
            const int64_t k0 = kp * kc, kn = (n2 - k0 < kc) ? n2 - k0 : kc;
            ev = tr_begin(tr, CAT_SPLIT);
            if (sparse) {
                rc = cutc_zero_u8(occ, tiles * KTp, st);
                if (rc == CUTC_OK)
                    rc = cutc_crt_quant_pairw_pk_occ(&crt, wr2 + k0, aos2 + k0, n2, kn, n_ao, kc, rexpA, pairs, Apl,
                                                     occ, st);
                if (rc == CUTC_OK)
                    rc = cutc_klists(occ, tiles, KTp, CUTC_CRT_LIST_ALIGN(kind), klist, kcount, st);
            } else {
                rc = cutc_crt_quant_pairw_pk_range(&crt, wr2 + k0, aos2 + k0, n2, kn, n_ao, kc, rexpA, Apl, st);
            }
            if (rc == CUTC_OK)
                rc = cutc_crt_quant_k_pk(&crt, G + k0, ldk, 4 * n1e, kn, kc, rexpG, Gpl, st);
            tr_end(tr, ev);
            TRY(rc);
            ev = tr_begin(tr, CAT_GEMM1);
            rc = cutc_crt_gemm_pk_list(&crt, Gpl, 4 * n1e, Apl, P, kc, kl, kc_, 0, 1, Cpl,
                                       (kp > 0) | CUTC_GEMM_KPASS, st);
            tr_end(tr, ev);
            TRY(rc);
            if (kl && i0 == 0 && stats && !tr->capture) {
                double sum = 0.0;
                TRY(kcount_sum(kcount, tiles, st, &sum));
                dens_sum += sum;
                dens_den += (double)tiles * (double)KTp;
            }
        }
        ev = tr_begin(tr, CAT_SPLIT);
        rc = slice ? cutc_oz_scatter_x((const double *)Cpl, P, 4 * n1e, pairs, n_ao, X, n1, n1e, i0, st)
                   : cutc_crt_rebuild_x(&crt, Cpl, P, 4 * n1e, rexpA, rexpG, pairs, n_ao, X, n1, n1e, i0, st);
        tr_end(tr, ev);
        TRY(rc);
        if (tr->x_ready)
            tr->x_ready(tr->x_ctx, i0, n1e, st);
        if (stats)
            stats->flops_exec += 2.0 * (double)P * 4.0 * (double)n1e * (double)n2;
    }
    if (stats) {
        stats->n1_chunk = n1c;
        stats->n1_chunks = nchunks;
    }
    if (!tr->capture)
        TRYH(hipStreamSynchronize(st));
    if (stats && !slice && !tr->capture) {
        /* executed fraction of the (256-pair tile, K step) tiles */
        double d = 1.0;
        if (kl && tiles > 0 && kpass == 1) {
            double sum = 0.0;
            TRY(kcount_sum(kcount, tiles, st, &sum));
            d = sum / ((double)tiles * (double)KT);
        } else if (kl && dens_den > 0) {
            d = dens_sum / dens_den;
        }
        stats->k_density = d;
    }
done:
    if (rc != CUTC_OK && !tr->capture) {
        hipStreamSynchronize(st);
        if (js != st)
            hipStreamSynchronize(js);
    }
    return rc;
}

/* ------------------------------------------------------------ stage 2 */

/* K splits of the FP64 step-2 contraction for n_pairs P and K = n1 chunk:
 * split only when the P x P grid of 128 x 128 tiles is small (< 256 blocks),
 * into at most 16 slices of at least 1024 K rows, keeping the partial sums
 * within `room` bytes (split2_room). */
static int split2_heuristic(int64_t P, int64_t K, size_t room)
{
    const int64_t tiles = ((P + 127) / 128) * ((P + 127) / 128);
    if (tiles >= 256 || K < 2048)
        return 1;
    int s = 1;
    while (s < 16 && tiles * s < 512 && K / (2 * s) >= 1024 && (size_t)(2 * s) * P * P * 8 <= room)
        s *= 2;
    return s;
}

/* K splits of the block-sparse modular step-2 contraction (256 x 256 blocks,
 * one per pair of 256-pair tiles, lists per H tile): split when the tiles
 * give fewer than 512 blocks, into at most 16 parts whose int8 residues
 * beyond the first part's (nmod P^2 bytes each) fit in `room` bytes; stage 2
 * runs fewer parts where a contraction has too few steps for them. */
static int split2_crt_heuristic(int64_t P, int nmod, size_t room)
{
    const int64_t t = (P + 255) / 256, tiles = t * t;
    int s = 1;
    while (s < 16 && tiles * s < 512 && (size_t)(2 * s - 1) * nmod * P * P <= room)
        s *= 2;
    return s;
}

/* the step-2 K split for the precision of opt (resolved options): FP64
 * and the block-sparse modular forms split small contractions, the slice
 * forms and the dense modular ones do not */
static int split2_for(const cutc_options *opt, int64_t P, int64_t K, size_t room)
{
    if (opt->precision == CUTC_PREC_FP64)
        return split2_heuristic(P, K, room);
    if (is_crt(opt->precision) && !opt->dense)
        return split2_crt_heuristic(P, opt->slices, room);
    return 1;
}

/* device bytes a step-2 K split may take beside the plan's stage 2: a
 * sixteenth of the budget, within what the plan leaves free */
static size_t split2_room(const cutc_plan *pl, size_t budget)
{
    const size_t left = budget > pl->bytes_stage2 ? budget - pl->bytes_stage2 : 0;
    return left < budget / 16 ? left : budget / 16;
}

/* Autotuner of the split: times the FP64 step-2 contraction and its reduce
 * for the candidate splits 1, 2, 4, 8, 16 (up to the heuristic's memory and
 * K limits) on scratch operands of the real shape (P pairs of n_ao orbitals,
 * K rows) and keeps the fastest; results are cached per (P, K). */
#define TUNE_CACHE 64
static struct {
    int64_t P, K;
    int split;
} g_tune[TUNE_CACHE];
static int g_ntune = 0;

static int tune_split2(int64_t P, int n_ao, int64_t K, size_t budget, hipStream_t st, double *ms_tune)
{
    for (int i = 0; i < g_ntune; i++)
        if (g_tune[i].P == P && g_tune[i].K == K)
            return g_tune[i].split;
    int smax = split2_heuristic(P, K, budget / 16);
    if (smax <= 1)
        return 1;
    while (smax < 16 && K / (2 * smax) >= 1024 && (size_t)(2 * smax) * P * P * 8 <= budget / 16)
        smax *= 2;
    const int64_t nao2 = (int64_t)n_ao * n_ao;
    double *X = NULL, *B = NULL, *M = NULL, *work = NULL;
    int *pairs = NULL;
    hipEvent_t e0 = NULL, e1 = NULL;
    int best = 1;
    float best_ms = 0.0f;
    const double t0 = now_ms();
    int rc = hip_rc(hipMalloc((void **)&X, sizeof(double) * K * nao2));
    if (rc == CUTC_OK)
        rc = hip_rc(hipMalloc((void **)&B, sizeof(double) * ((K + 15) / 16 * 16) * P));
    if (rc == CUTC_OK)
        rc = hip_rc(hipMalloc((void **)&M, sizeof(double) * nao2 * nao2));
    if (rc == CUTC_OK)
        rc = hip_rc(hipMalloc((void **)&work, sizeof(double) * smax * P * P));
    if (rc == CUTC_OK)
        rc = hip_rc(hipMalloc((void **)&pairs, sizeof(int) * 2 * P));
    if (rc == CUTC_OK)
        rc = cutc_launch_fill(X, K * nao2, 1.0, st);
    if (rc == CUTC_OK)
        rc = cutc_launch_fill(B, (K + 15) / 16 * 16 * P, 0.5, st);
    if (rc == CUTC_OK)
        rc = cutc_launch_pair_table(n_ao, pairs, st);
    if (rc == CUTC_OK)
        rc = hip_rc(hipEventCreate(&e0));
    if (rc == CUTC_OK)
        rc = hip_rc(hipEventCreate(&e1));
    for (int s = 1; s <= smax && rc == CUTC_OK; s *= 2) {
        float ms = 0.0f;
        for (int rep = 0; rep < 2 && rc == CUTC_OK; rep++) { /* the first run warms up */
            rc = hip_rc(hipEventRecord(e0, st));
            if (rc == CUTC_OK && s == 1)
                rc = cutc_gemm2_f64(X, nao2, B, P, P, K, 1.0, pairs, n_ao, M, 1, st);
            if (rc == CUTC_OK && s > 1)
                rc = cutc_gemm2_f64_split(X, nao2, B, P, P, K, pairs, n_ao, s, work, st);
            if (rc == CUTC_OK && s > 1)
                rc = cutc_scatter_m(work, cutc_gemm2_split_count(K, s), P, 1.0, pairs, n_ao, M, 1, st);
            if (rc == CUTC_OK)
                rc = hip_rc(hipEventRecord(e1, st));
            if (rc == CUTC_OK)
                rc = hip_rc(hipEventSynchronize(e1));
            if (rc == CUTC_OK)
                rc = hip_rc(hipEventElapsedTime(&ms, e0, e1));
        }
        if (rc == CUTC_OK && (s == 1 || ms < best_ms)) {
            best = s;
            best_ms = ms;
        }
    }
    if (e0)
        hipEventDestroy(e0);
    if (e1)
        hipEventDestroy(e1);
    dfree(X);
    dfree(B);
    dfree(M);
    dfree(work);
    dfree(pairs);
    if (ms_tune)
        *ms_tune += now_ms() - t0;
    if (rc != CUTC_OK)
        return split2_heuristic(P, K, budget / 16);
    if (g_ntune < TUNE_CACHE) {
        g_tune[g_ntune].P = P;
        g_tune[g_ntune].K = K;
        g_tune[g_ntune].split = best;
        g_ntune++;
    }
    return best;
}

/* Stage 2 over the r1 points r0 .. r0 + cnt - 1: r1 / wr1 / aos1 are the
 * caller's full arrays (leading dimension n1), X holds the cnt points' four
 * components with component stride cnt.  accumulate = 0: the first
 * contraction stores into out, else everything adds; symm = 1: out = M + M^T
 * at the end.  nsplit2 > 1: the FP64 contractions run split-K (for small
 * n_pairs, whose P x P tile grid leaves most of the GPU idle).
 * Modular precisions with r1 given: block-sparse, as stage 1 with the roles
 * of the operands swapped.  Each r1 chunk's points run in Morton order (the
 * integrands are formed at the permuted points and X is read through the
 * same permutation), the H-side pairs are clustered by where their quantised
 * H values are non-zero, and the contraction runs, per 256-pair H tile (the
 * GEMM's A side), only the K steps where that tile has a non-zero residue
 * (opt->dense: all of them).  Its results go into halves of M (see
 * cutc_crt_rebuild_m), completed with the symmetrisation. */
static int stage2(const cutc_options *opt, const cutc_plan *pl, int64_t n1, int64_t r0, int64_t cnt, int n_ao,
                  const double *r1, const double *wr1, const double *aos1,
                  const double *X, double *out, int accumulate, int symm, int nsplit2, tracer_t *tr,
                  cutc_stats *stats)
{
    const hipStream_t st = tr->st;
    const int64_t P = pl->n_pairs, n1c = pl->n1_chunk_2e;
    const int64_t nao2 = (int64_t)n_ao * n_ao;
    int *pairs = NULL, *pairsH = NULL, *kcount = NULL;
    double *Hp = NULL, *NHp = NULL;
    int *rexpA = NULL, *rexpB = NULL;
    double *npart = NULL, *work = NULL;
    int8_t *Apl = NULL, *Bpl = NULL, *Cpl = NULL;
    uint8_t *occ = NULL;
    uint32_t *perm1 = NULL, *klist = NULL;
    uint64_t *skeys = NULL;
    double *box = NULL;
    void *cws = NULL;
    int rc = CUTC_OK, ev;
    const int split = opt->precision != CUTC_PREC_FP64;
    const int slice = split && !is_crt(opt->precision), f16 = opt->precision == CUTC_PREC_FP16;
    const int emode = slice ? CUTC_ROWEXP_MAX : CUTC_ROWEXP_NORM;
    const int nmod = opt->slices;
    const size_t rb = CUTC_CRT_RBYTES(crt_kind(opt->precision));

    const int64_t rows = (n1c + 15) / 16 * 16;
    const int64_t plane = rows * P;
    const int64_t kp_max = (n1c + CUTC_PK_KALIGN - 1) / CUTC_PK_KALIGN * CUTC_PK_KALIGN;
    ws_t ws;
    /* operand planes of both sides and the contraction's results */
    const size_t bpl = slice ? (size_t)nmod * 2 * CUTC_OZ_ROWS(P) * 3 * kp_max
                             : rb * nmod * CUTC_PK_ROWS_A(P) * 3 * kp_max;
    /* block-sparse state: r1 order of a chunk, sort keys, the H-side pair
     * table, occupancy flags, step lists and counts per 256-pair H tile */
    const int sparse = split && !slice && r1 != NULL && P > 0;
    /* modular, list-driven: up to nsc parts of a K split, each its own residues */
    const int nsc = (sparse && !opt->dense && nsplit2 > 1) ? nsplit2 : 1;
    const size_t bcp = slice ? sizeof(double) * P * P : (size_t)nsc * nmod * P * P;
    const int kind = slice ? 0 : crt_kind(opt->precision);
    const int64_t tilesH = CUTC_PK_ROWS_A(P) / CUTC_SP_ROWS;
    const int64_t KT2 = sparse ? 3 * kp_max / CUTC_CRT_KSTEP(kind) : 0;
    const int64_t nsort = sparse ? cutc_sort_len(n1c) : 0;
    const int64_t KTH = sparse ? kp_max / CUTC_CRT_KSTEP(kind) : 0; /* steps of the H segment */
    {
        size_t need = ws_round(sizeof(int) * 2 * P) + ws_round(sizeof(double) * plane) +
                      ws_round(sizeof(double) * 3 * plane);
        if (!split && nsplit2 > 1)
            need += ws_round(sizeof(double) * nsplit2 * P * P);
        if (split)
            need += 2 * ws_round(sizeof(int) * P) + ws_round(sizeof(double) * CUTC_CRT_KSPLIT * P) + 2 * ws_round(bpl) +
                    ws_round(bcp);
        if (sparse)
            need += ws_round(sizeof(uint64_t) * nsort) + ws_round(sizeof(uint32_t) * nsort) +
                    ws_round(sizeof(double) * 6) + ws_round(cutc_cluster_bytes(P, KTH)) +
                    ws_round(sizeof(int) * 2 * P) + ws_round(tilesH * KT2) + ws_round(sizeof(uint32_t) * tilesH * KT2) +
                    ws_round(sizeof(int) * tilesH);
        TRY(ws_begin(&ws, need));
        if (stats && need > stats->workspace_bytes)
            stats->workspace_bytes = need;
    }
    WS_TAKE(pairs, sizeof(int) * 2 * P);
    WS_TAKE(Hp, sizeof(double) * plane);
    WS_TAKE(NHp, sizeof(double) * 3 * plane);
    if (!split && nsplit2 > 1)
        WS_TAKE(work, sizeof(double) * nsplit2 * P * P);
    if (split) {
        WS_TAKE(rexpA, sizeof(int) * P);
        WS_TAKE(rexpB, sizeof(int) * P);
        WS_TAKE(npart, sizeof(double) * CUTC_CRT_KSPLIT * P);
        WS_TAKE(Apl, bpl);
        WS_TAKE(Bpl, bpl);
        WS_TAKE(Cpl, bcp);
    }
    if (sparse) {
        WS_TAKE(skeys, sizeof(uint64_t) * nsort);
        WS_TAKE(perm1, sizeof(uint32_t) * nsort); /* the point sort's values: the permutation */
        WS_TAKE(box, sizeof(double) * 6);
        WS_TAKE(cws, cutc_cluster_bytes(P, KTH));
        WS_TAKE(pairsH, sizeof(int) * 2 * P);
        WS_TAKE(occ, tilesH * KT2);
        WS_TAKE(klist, sizeof(uint32_t) * tilesH * KT2);
        WS_TAKE(kcount, sizeof(int) * tilesH);
    }
    TRY(cutc_launch_pair_table(n_ao, pairs, st));
    double dens_sum = 0.0, dens_den = 0.0;

    for (int64_t i0 = 0; i0 < cnt; i0 += n1c) {
        const int64_t n1e = (cnt - i0 < n1c) ? cnt - i0 : n1c;
        const int first = i0 == 0 && !accumulate;
        const int64_t pad0 = n1e * P, padn = ((n1e + 15) / 16 * 16 - n1e) * P;
        ev = tr_begin(tr, CAT_SHORTR);
        if (sparse) {
            /* the chunk's r1 order, then the integrands at the permuted points */
            rc = cutc_order_points(r1 + 3 * (r0 + i0), n1e, box, skeys, perm1, cutc_sort_len(n1e), st);
        }
        /* (block-sparse: H alone, which orders the H pairs; H and NH follow
         * in that order below) */
        if (rc == CUTC_OK)
            rc = cutc_launch_r1_integrands(wr1, aos1, n1, r0 + i0, n1e, n_ao, pairs, Hp, sparse ? NULL : NHp, P,
                                           plane, sparse ? perm1 : NULL, st);
        for (int m = 0; m < 4 && rc == CUTC_OK && padn > 0; m++)
            rc = cutc_launch_fill((m == 0 ? Hp : NHp + (m - 1) * plane) + pad0, padn, 0.0, st);
        tr_end(tr, ev);
        TRY(rc);
        if (sparse) {
            /* H-side pair order by support (keys from H = w1 chi_a chi_b at the
             * bits of the first part), then the integrands again in that order */
            cutc_crt crt0;
            cutc_colsrc sh;
            memset(&sh, 0, sizeof sh);
            sh.nsrc = 1;
            sh.p[0] = Hp;
            sh.ld = P;
            sh.K = n1e;
            ev = tr_begin(tr, CAT_SPLIT);
            rc = cutc_crt_setup(n1e, nmod, kind, &crt0);
            if (rc == CUTC_OK)
                rc = cutc_rowexp_colk(&sh, P, emode, npart, rexpB, st);
            if (rc == CUTC_OK)
                rc = cutc_order_pairs_colk(&sh, n_ao, crt0.bits, rexpB, CUTC_CRT_KSTEP(kind),
                                           (n1e + CUTC_CRT_KSTEP(kind) - 1) / CUTC_CRT_KSTEP(kind), cws, pairsH, st);
            tr_end(tr, ev);
            TRY(rc);
            ev = tr_begin(tr, CAT_SHORTR);
            rc = cutc_launch_r1_integrands(wr1, aos1, n1, r0 + i0, n1e, n_ao, pairsH, Hp, NHp, P, plane, perm1, st);
            tr_end(tr, ev);
            TRY(rc);
        }
        if (!split && nsplit2 > 1) {
            const int ns = cutc_gemm2_split_count(n1e, nsplit2);
            ev = tr_begin(tr, CAT_GEMM2);
            for (int m = 3; m >= 0 && rc == CUTC_OK; m--) {
                rc = cutc_gemm2_f64_split(X + (m * cnt + i0) * nao2, nao2, m == 3 ? Hp : NHp + m * plane, P, P, n1e,
                                          pairs, n_ao, nsplit2, work, st);
                if (rc == CUTC_OK)
                    rc = cutc_scatter_m(work, ns, P, m == 3 ? 1.0 : -0.5, pairs, n_ao, out,
                                        m == 3 ? (first ? 0 : 1) : 2, st);
            }
            tr_end(tr, ev);
            TRY(rc);
        } else if (!split) {
            ev = tr_begin(tr, CAT_GEMM2);
            rc = cutc_gemm2_f64(X + (3 * cnt + i0) * nao2, nao2, Hp, P, P, n1e, 1.0, pairs, n_ao, out, first ? 0 : 1,
                                st);
            for (int m = 0; m < 3 && rc == CUTC_OK; m++)
                rc = cutc_gemm2_f64(X + (m * cnt + i0) * nao2, nao2, NHp + m * plane, P, P, n1e, -0.5, pairs, n_ao,
                                    out, 2, st);
            tr_end(tr, ev);
            TRY(rc);
        } else {
            /* X^3 H, then X^{0..2} NH^{0..2} as one contraction over (m, r1);
             * the H side is the GEMM's A operand, so C[q][p] has the X-side
             * pair p contiguous */
            const int64_t kp = (n1e + CUTC_PK_KALIGN - 1) / CUTC_PK_KALIGN * CUTC_PK_KALIGN;
            for (int part = 0; part < 2; part++) {
                const int nseg = part == 0 ? 1 : 3;
                cutc_crt crt;
                if (!slice)
                    TRY(cutc_crt_setup(nseg * n1e, nmod, crt_kind(opt->precision), &crt));
                cutc_colsrc sa, sb;
                memset(&sa, 0, sizeof sa);
                memset(&sb, 0, sizeof sb);
                sa.nsrc = sb.nsrc = nseg;
                sa.ld = nao2;
                sb.ld = P;
                sa.K = sb.K = n1e;
                sa.pairs = pairs;
                sa.n_ao = n_ao;
                sa.kperm = sparse ? perm1 : NULL; /* X rows of the permuted points */
                for (int g = 0; g < nseg; g++) {
                    const int m = part == 0 ? 3 : g;
                    sa.p[g] = X + (m * cnt + i0) * nao2;
                    sb.p[g] = part == 0 ? Hp : NHp + g * plane;
                }
                ev = tr_begin(tr, CAT_SPLIT);
                rc = cutc_rowexp_colk(&sa, P, emode, npart, rexpA, st);
                if (rc == CUTC_OK)
                    rc = slice ? cutc_oz_slice_colk(nmod, f16, &sa, P, kp, CUTC_OZ_ROWS(P), rexpA, (uint16_t *)Apl, st)
                               : cutc_crt_quant_colk_pk(&crt, &sa, P, kp, CUTC_PK_ROWS_B(P), rexpA, Apl, st);
                if (rc == CUTC_OK)
                    rc = cutc_rowexp_colk(&sb, P, emode, npart, rexpB, st);
                const int64_t kt = sparse ? nseg * kp / CUTC_CRT_KSTEP(kind) : 0;
                if (rc == CUTC_OK && sparse)
                    rc = cutc_zero_u8(occ, tilesH * kt, st);
                if (rc == CUTC_OK)
                    rc = slice ? cutc_oz_slice_colk(nmod, f16, &sb, P, kp, CUTC_OZ_ROWS(P), rexpB, (uint16_t *)Bpl, st)
                               : cutc_crt_quant_colk_pk_occ(&crt, &sb, P, kp, CUTC_PK_ROWS_A(P), rexpB, Bpl,
                                                            sparse ? occ : NULL, st);
                if (rc == CUTC_OK && sparse)
                    rc = cutc_klists(occ, tilesH, kt, CUTC_CRT_LIST_ALIGN(kind), klist, kcount, st);
                tr_end(tr, ev);
                TRY(rc);
                const int lists = sparse && !opt->dense;
                /* K split parts: at least 8 padding units of the lists each */
                int ns = 1;
                if (lists && nsc > 1) {
                    const int64_t np = nseg * kp / CUTC_CRT_KSTEP(kind) / (8 * CUTC_CRT_LIST_ALIGN(kind));
                    ns = np >= nsc ? nsc : (np > 1 ? (int)np : 1);
                }
                ev = tr_begin(tr, CAT_GEMM2);
                rc = slice ? cutc_oz_gemm(nmod, f16, (const uint16_t *)Bpl, P, (const uint16_t *)Apl, P, nseg * kp, rexpB,
                                          rexpA, (double *)Cpl, st)
                           : cutc_crt_gemm_pk_list(&crt, Bpl, P, Apl, P, nseg * kp, lists ? klist : NULL,
                                                   lists ? kcount : NULL, 1, ns, Cpl, 0, st);
                tr_end(tr, ev);
                TRY(rc);
                if (lists && stats && !tr->capture) {
                    /* executed fraction of the (256-pair H tile, K step) tiles */
                    int *h = (int *)malloc(sizeof(int) * tilesH);
                    if (!h)
                        TRY(CUTC_ERR_NOMEM);
                    rc = hipMemcpyAsync(h, kcount, sizeof(int) * tilesH, hipMemcpyDeviceToHost, st) == hipSuccess &&
                                 hipStreamSynchronize(st) == hipSuccess
                             ? CUTC_OK
                             : CUTC_ERR_HIP;
                    for (int64_t i = 0; i < tilesH && rc == CUTC_OK; i++)
                        dens_sum += (double)h[i] * (double)nmod;
                    free(h);
                    TRY(rc);
                    dens_den += (double)tilesH * (double)kt * (double)nmod;
                }
                ev = tr_begin(tr, CAT_SPLIT);
                const double alpha = part == 0 ? 1.0 : -0.5;
                /* slices: the full layout (store, add, add antisymmetric); modular:
                 * the halves of M (store or add), completed by cutc_crt_finalize_m */
                rc = slice ? cutc_scatter_m((const double *)Cpl, 1, P, alpha, pairs, n_ao, out,
                                            part == 0 ? (first ? 0 : 1) : 2, st)
                           : cutc_crt_rebuild_m(&crt, Cpl, P, rexpA, rexpB, alpha, pairs, sparse ? pairsH : NULL,
                                                n_ao, out, part, first ? 0 : 1, ns, st);
                tr_end(tr, ev);
                TRY(rc);
            }
        }
        if (stats)
            stats->flops_exec += 2.0 * (double)P * (double)P * 4.0 * (double)n1e;
    }
    if (symm) {
        ev = tr_begin(tr, CAT_SYMM);
        if (split && !slice) /* the modular halves of M into the full layout */
            rc = cutc_crt_finalize_m(out, n_ao, st);
        if (rc == CUTC_OK)
            rc = cutc_launch_trans_inplace(out, nao2, st);
        tr_end(tr, ev);
        TRY(rc);
    }
    if (stats && dens_den > 0)
        stats->k_density_2e = dens_sum / dens_den;
    if (!tr->capture)
        TRYH(hipStreamSynchronize(st));
done:
    if (rc != CUTC_OK && !tr->capture)
        hipStreamSynchronize(st);
    return rc;
}

/* --------------------------------------------------------- entry points */


static int check_sizes(int n1, int n2, int n_ao, int n_nuc, int size_bh)
{
    if (n1 < 0 || n2 < 0 || n_ao < 0 || n_nuc < 0 || size_bh < 0)
        return CUTC_ERR_ARG;
    return CUTC_OK;
}

static void fill_stats(cutc_stats *stats, tracer_t *tr)
{
    if (!stats)
        return;
    double ms[NCAT];
    tr_collect(tr, ms);
    stats->ms_jastrow = ms[CAT_JAST];
    stats->ms_integrands = ms[CAT_LONGR] + ms[CAT_SHORTR];
    stats->ms_long_range = ms[CAT_LONGR];
    stats->ms_short_range = ms[CAT_SHORTR];
    stats->ms_stage1 = ms[CAT_STAGE1];
    stats->ms_stage2 = ms[CAT_STAGE2];
    stats->ms_gemm1 = ms[CAT_GEMM1];
    stats->ms_gemm2 = ms[CAT_GEMM2];
    stats->ms_symm = ms[CAT_SYMM];
    stats->ms_split = ms[CAT_SPLIT];
    stats->ms_total = ms[CAT_TOTAL];
    stats->n_gemm1 = tr->launches[CAT_GEMM1];
    stats->n_gemm2 = tr->launches[CAT_GEMM2];
    stats->n_jastrow = tr->launches[CAT_JAST];
}

/* The reference's chunk-plan lines (src/build_int2_grad1_u12_ao.cu:60-61,
 * 81-95) for this build's plan: free and total device memory, r1 rows per
 * chunk with the count of full chunks and the rows of the last, r2 points
 * per K pass likewise. */
static void print_plan(const cutc_plan *pl, int64_t n1, int64_t n2)
{
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) != hipSuccess)
        fr = tot = 0;
    printf(" free  memory = %.3f GB\n", (double)fr / 1073741824.0);
    printf(" total memory = %.3f GB\n", (double)tot / 1073741824.0);
    const int64_t p1 = pl->n1_chunk > 0 ? pl->n1_chunk : 1;
    int64_t p2 = pl->k_chunk < n2 ? pl->k_chunk : n2;
    if (p2 < 1)
        p2 = 1;
    printf("n_grid1_pass = %lld\n", (long long)(pl->n1_chunk < n1 ? pl->n1_chunk : n1));
    printf("n_grid1_rest = %lld\n", (long long)(n1 % p1));
    printf("n1_pass = %lld\n", (long long)(n1 / p1));
    printf("n_grid2_pass = %lld\n", (long long)(n2 > 0 ? p2 : 0));
    printf("n_grid2_rest = %lld\n", (long long)(n2 % p2));
    printf("n2_pass = %lld\n", (long long)(n2 / p2));
    fflush(stdout);
}

/* The reference's stage-time lines, in its order: stage 1
 * (src/build_int2_grad1_u12_ao.cu:342-344), the stage-1 total as cutc_int
 * prints it (src/cutc_int.c:153, host entry points only), stage 2
 * (src/build_int_2e_ao.cu:139-142) and its total (src/cutc_int.c:198).  The
 * herm and non-herm integrands run as one kernel here: its time is on the
 * herm line.  Split precisions add a line for their quantisation and CRT
 * rebuild kernels, which the reference does not have. */
static void print_stage_times(tracer_t *tr, int stages, int split, int host)
{
    double ms[NCAT];
    tr_collect(tr, ms);
    if (stages & 1) {
        printf("Ellapsed time for DGEMM to build int2_grad1_u12_ao = %.3f sec\n", ms[CAT_GEMM1] / 1000.0);
        printf("Ellapsed time for int_long_range kernel = %.3f sec\n", ms[CAT_LONGR] / 1000.0);
        printf("Ellapsed time for cutc_int_bh kernel = %.3f sec\n", ms[CAT_JAST] / 1000.0);
        if (host)
            printf("Ellapsed time for get_int2_grad1_u12_ao = %.3f sec\n", ms[CAT_STAGE1] / 1000.0);
    }
    if (stages & 2) {
        printf("Ellapsed time for DGEMM to build int_2e_ao = %.3f sec\n", ms[CAT_GEMM2] / 1000.0);
        printf("Ellapsed time for int_short_range_herm kernel = %.3f sec\n", ms[CAT_SHORTR] / 1000.0);
        printf("Ellapsed time for int_short_range_nonherm kernel = %.3f sec (computed by the herm kernel)\n", 0.0);
        printf("Ellapsed time for addT = %.3f sec\n", ms[CAT_SYMM] / 1000.0);
        if (host)
            printf("Ellapsed time for get_int_2e_ao = %.3f sec\n", ms[CAT_STAGE2] / 1000.0);
    }
    if (split)
        printf("Ellapsed time for split-precision quantisation and rebuild kernels = %.3f sec\n",
               ms[CAT_SPLIT] / 1000.0);
    fflush(stdout);
}

static int device_path(const cutc_options *opt, int stages, int n1, int n2, int n_ao, int n_nuc, int size_bh,
                       const double *r1, const double *wr1, const double *r2, const double *wr2, const double *rn,
                       const double *aos1, const double *aos2, const double *c_bh, const int *m_bh, const int *n_bh,
                       const int *o_bh, double *X, double *out, hipStream_t st, cutc_stats *stats,
                       const tracer_t *hook, const jprog_t *pre)
{
    int rc = check_sizes(n1, n2, n_ao, n_nuc, size_bh);
    if (rc)
        return rc;
    const cutc_options ropt = resolved_options(opt);
    opt = &ropt;
    if (!valid_precision(opt))
        return CUTC_ERR_ARG;
    if (n_ao > 0 && n1 > 0 && !X)
        return CUTC_ERR_ARG;
    if ((stages & 2) && n_ao > 0 && !out)
        return CUTC_ERR_ARG;
    if ((stages & 1) && n1 > 0 && n_ao > 0 && n2 > 0 &&
        (!r1 || !r2 || !wr2 || !aos2 || (n_nuc > 0 && !rn) ||
         ((int64_t)n_nuc * size_bh > 0 && (!c_bh || !m_bh || !n_bh || !o_bh))))
        return CUTC_ERR_ARG;
    if ((stages & 2) && n1 > 0 && n_ao > 0 && (!wr1 || !aos1))
        return CUTC_ERR_ARG;
    jprog_t prog;
    if (pre)
        prog = *pre;
    else
        rc = jprog_load(n_nuc, size_bh, c_bh, m_bh, n_bh, o_bh, (stages & 1) && n1 > 0 && n2 > 0 && n_ao > 0, &prog);
    if (rc) {
        jprog_free(&prog);
        return rc;
    }
    cutc_plan pl;
    rc = cutc_make_plan(n1, n2, n_ao, n_nuc, opt->precision, opt->slices, device_budget(opt), opt->n1_chunk, &pl);
    if (rc) {
        jprog_free(&prog);
        return rc;
    }
    if (stats)
        memset(stats, 0, sizeof *stats);
    tracer_t tr;
    tr_init(&tr, st, stats != NULL || opt->verbose);
    if (hook) {
        tr.x_ready = hook->x_ready;
        tr.x_ctx = hook->x_ctx;
    }
    if (opt->verbose && (stages & 1))
        print_plan(&pl, n1, n2);
    int evt = tr_begin(&tr, CAT_TOTAL);
    int evs = tr_begin(&tr, CAT_STAGE1);
    if ((stages & 1) && n1 > 0 && n_ao > 0) {
        if (n2 == 0) {
            int ev = tr_begin(&tr, CAT_GEMM1);
            rc = cutc_launch_fill(X, 4 * (int64_t)n1 * n_ao * n_ao, 0.0, st);
            tr_end(&tr, ev);
        } else if (opt->precision != CUTC_PREC_FP64) {
            rc = stage1_split(opt, &pl, n1, n2, n_ao, n_nuc, r1, r2, wr2, rn, aos2, &prog, X, &tr, stats, NULL);
        } else {
            rc = stage1(opt, &pl, n1, n2, n_ao, n_nuc, r1, r2, wr2, rn, aos2, &prog, X, &tr, stats);
        }
    }
    tr_end(&tr, evs);
    evs = tr_begin(&tr, CAT_STAGE2);
    if (rc == CUTC_OK && (stages & 2) && n_ao > 0) {
        if (n1 == 0) {
            int ev = tr_begin(&tr, CAT_GEMM2);
            rc = cutc_launch_fill(out, (int64_t)n_ao * n_ao * n_ao * n_ao, 0.0, st);
            tr_end(&tr, ev);
        } else {
            rc = stage2(opt, &pl, n1, 0, n1, n_ao, r1, wr1, aos1, X, out, 0, 1,
                        split2_for(opt, pl.n_pairs, pl.n1_chunk_2e, split2_room(&pl, device_budget(opt))), &tr,
                        stats);
        }
    }
    tr_end(&tr, evs);
    tr_end(&tr, evt);
    if (rc == CUTC_OK)
        rc = hip_rc(hipStreamSynchronize(st));
    else
        hipStreamSynchronize(st);
    if (rc == CUTC_OK) {
        fill_stats(stats, &tr);
        if (opt->verbose)
            print_stage_times(&tr, stages, opt->precision != CUTC_PREC_FP64, hook != NULL);
    }
    tr_free(&tr);
    jprog_free(&prog);
    return rc;
}

int cutc_int_device(const cutc_options *opt, int n_grid1, int n_grid2, int n_ao, int n_nuc, int size_bh,
                    const double *r1, const double *wr1, const double *r2, const double *wr2, const double *rn,
                    const double *aos_data1, const double *aos_data2, const double *c_bh, const int *m_bh,
                    const int *n_bh, const int *o_bh, double *int2_grad1_u12_ao, double *int_2e_ao, void *stream,
                    cutc_stats *stats)
{
    cutc_options def;
    if (!opt) {
        cutc_default_options(&def);
        opt = &def;
    }
    pthread_mutex_lock(&g_lock);
    int rc = device_path(opt, 3, n_grid1, n_grid2, n_ao, n_nuc, size_bh, r1, wr1, r2, wr2, rn, aos_data1, aos_data2,
                         c_bh, m_bh, n_bh, o_bh, int2_grad1_u12_ao, int_2e_ao, (hipStream_t)stream, stats, NULL, NULL);
    pthread_mutex_unlock(&g_lock);
    cutc_set_last_error(rc);
    return rc;
}

/* cutc_int_device without allocating: the Jastrow program goes through
 * static host arrays into the front of the device scratch (scratch.hip), the
 * stages carve their workspace from the rest of it and plan their r1 chunks
 * against its size, and no timing events are made.  Calls on one GPU share
 * its scratch (serialised by the library lock; each returns after its work
 * has completed). */
int cutc_int_device_noalloc(const cutc_options *opt, int n_grid1, int n_grid2, int n_ao, int n_nuc, int size_bh,
                            const double *r1, const double *wr1, const double *r2, const double *wr2,
                            const double *rn, const double *aos_data1, const double *aos_data2, const double *c_bh,
                            const int *m_bh, const int *n_bh, const int *o_bh, double *int2_grad1_u12_ao,
                            double *int_2e_ao, void *stream)
{
    cutc_options o;
    if (opt)
        o = *opt;
    else
        cutc_default_options(&o);
    o.verbose = 0; /* printing times needs events */
    int rc = check_sizes(n_grid1, n_grid2, n_ao, n_nuc, size_bh);
    const int need = n_grid1 > 0 && n_grid2 > 0 && n_ao > 0;
    if (rc == CUTC_OK && need && (int64_t)n_nuc * size_bh > 0 && (!c_bh || !m_bh || !n_bh || !o_bh))
        rc = CUTC_ERR_ARG;
    if (rc != CUTC_OK) {
        cutc_set_last_error(rc);
        return rc;
    }
    pthread_mutex_lock(&g_lock);
    char *base = NULL;
    size_t bytes = 0, used = 0;
    jprog_t jp;
    memset(&jp, 0, sizeof jp);
    rc = cutc_scratch((void **)&base, &bytes);
    if (rc == CUTC_OK)
        rc = jprog_load_static(n_nuc, size_bh, c_bh, m_bh, n_bh, o_bh, need, base, bytes, &used, &jp);
    if (rc == CUTC_OK) {
        g_ws_fixed.base = base + used;
        g_ws_fixed.size = bytes - used;
        rc = device_path(&o, 3, n_grid1, n_grid2, n_ao, n_nuc, size_bh, r1, wr1, r2, wr2, rn, aos_data1, aos_data2,
                         c_bh, m_bh, n_bh, o_bh, int2_grad1_u12_ao, int_2e_ao, (hipStream_t)stream, NULL, NULL, &jp);
        g_ws_fixed.base = NULL;
        g_ws_fixed.size = 0;
    }
    pthread_mutex_unlock(&g_lock);
    cutc_set_last_error(rc);
    return rc;
}

/* The r1 chunks of a streamed build (n1 > 0, n_ao > 0): stage 1 of each chunk
 * into the chunk buffer Xc (nc points), then stage 2 adds it into out, which
 * is symmetrised after the last chunk. */
static int streamed_chunks(const cutc_options *opt, const cutc_plan *pl, int64_t n1, int64_t n2, int n_ao, int n_nuc,
                           const double *r1, const double *wr1, const double *r2, const double *wr2, const double *rn,
                           const double *aos1, const double *aos2, const jprog_t *jp, double *out, double *Xc,
                           int64_t nc, int nsplit2, tracer_t *tr, cutc_stats *stats, s1pre_t *pre)
{
    const int64_t nao2 = (int64_t)n_ao * n_ao;
    int rc = CUTC_OK;
    for (int64_t r0 = 0; r0 < n1 && rc == CUTC_OK; r0 += nc) {
        const int64_t cnt = (n1 - r0 < nc) ? n1 - r0 : nc;
        int evs = tr_begin(tr, CAT_STAGE1);
        if (n2 == 0) {
            int ev = tr_begin(tr, CAT_GEMM1);
            rc = cutc_launch_fill(Xc, 4 * cnt * nao2, 0.0, tr->st);
            tr_end(tr, ev);
        } else if (opt->precision != CUTC_PREC_FP64) {
            rc = stage1_split(opt, pl, cnt, n2, n_ao, n_nuc, r1 + 3 * r0, r2, wr2, rn, aos2, jp, Xc, tr, stats, pre);
        } else {
            rc = stage1(opt, pl, cnt, n2, n_ao, n_nuc, r1 + 3 * r0, r2, wr2, rn, aos2, jp, Xc, tr, stats);
        }
        tr_end(tr, evs);
        if (rc == CUTC_OK) {
            evs = tr_begin(tr, CAT_STAGE2);
            rc = stage2(opt, pl, n1, r0, cnt, n_ao, r1, wr1, aos1, Xc, out, r0 > 0, r0 + cnt >= n1, nsplit2, tr,
                        stats);
            tr_end(tr, evs);
        }
    }
    return rc;
}

/* Streamed build: X is never stored whole.  The r1 points run in chunks of
 * nc; each chunk's X (4 x nc x n_ao^2, one device allocation for the call)
 * goes straight from stage 1 into stage 2, which adds it into out; out is
 * symmetrised after the last chunk.  nc takes half the memory left over by
 * the stages' fixed buffers (or opt->n1_chunk when smaller): every X chunk
 * repeats the pair quantisation of stage 1 and the orderings, quantisation
 * and rebuild of stage 2, while smaller G chunks of stage 1 cost little. */
static int streamed_path(const cutc_options *opt, int n1, int n2, int n_ao, int n_nuc, int size_bh,
                         const double *r1, const double *wr1, const double *r2, const double *wr2, const double *rn,
                         const double *aos1, const double *aos2, const double *c_bh, const int *m_bh,
                         const int *n_bh, const int *o_bh, double *out, hipStream_t st, cutc_stats *stats)
{
    int rc = check_sizes(n1, n2, n_ao, n_nuc, size_bh);
    if (rc)
        return rc;
    const cutc_options ropt = resolved_options(opt);
    opt = &ropt;
    if (!valid_precision(opt))
        return CUTC_ERR_ARG;
    if (n_ao > 0 && !out)
        return CUTC_ERR_ARG;
    if (n1 > 0 && n_ao > 0 &&
        (!r1 || !wr1 || !aos1 ||
         (n2 > 0 && (!r2 || !wr2 || !aos2 || (n_nuc > 0 && !rn) ||
                     ((int64_t)n_nuc * size_bh > 0 && (!c_bh || !m_bh || !n_bh || !o_bh))))))
        return CUTC_ERR_ARG;
    if (stats)
        memset(stats, 0, sizeof *stats);
    if (n_ao == 0)
        return CUTC_OK;
    const int64_t nao2 = (int64_t)n_ao * n_ao;
    if (n1 == 0) {
        rc = cutc_launch_fill(out, nao2 * nao2, 0.0, st);
        if (rc == CUTC_OK)
            rc = hip_rc(hipStreamSynchronize(st));
        return rc;
    }
    jprog_t prog;
    rc = jprog_load(n_nuc, size_bh, c_bh, m_bh, n_bh, o_bh, n2 > 0, &prog);
    if (rc) {
        jprog_free(&prog);
        return rc;
    }
    double *Xc = NULL;
    char *preb = NULL;
    s1pre_t pre;
    memset(&pre, 0, sizeof pre);
    tracer_t tr;
    int have_tr = 0;
    /* chunk of the stored X; the modular precisions also keep the
     * r1-independent orderings of stage 1 across the chunks */
    const int keep = is_crt(opt->precision) && n2 > 0;
    const size_t preby = keep ? s1pre_bytes(n2, n_ao) : 0;
    /* the buffer of an earlier call (Xc and the orderings) counts as free */
    const size_t budget0 = device_budget_plus(opt, buf_size_current(g_sbuf)),
                 budget = budget0 > preby ? budget0 - preby : 0;
    cutc_plan p1;
    rc = cutc_make_plan(1, n2, n_ao, n_nuc, opt->precision, opt->slices, budget, 0, &p1);
    if (rc)
        goto done;
    {
        const size_t fixed = p1.bytes_stage1 > p1.bytes_stage2 ? p1.bytes_stage1 : p1.bytes_stage2;
        const size_t xper = 32 * (size_t)nao2;
        int64_t nc = budget > fixed ? (int64_t)((budget - fixed) / 2 / xper) : 0;
        if (opt->n1_chunk > 0 && opt->n1_chunk < nc)
            nc = opt->n1_chunk;
        if (nc < 1) {
            rc = CUTC_ERR_NOMEM;
            goto done;
        }
        if (nc < n1) {
            const int64_t nch = (n1 + nc - 1) / nc;
            nc = (n1 + nch - 1) / nch;
        } else {
            nc = n1;
        }
        /* X chunk and kept orderings in the grow-only buffer of the streamed
         * calls (at most twice this call's need, so that the stages keep
         * the memory the chunk size assumed) */
        const size_t xb = ws_round(sizeof(double) * 4 * (size_t)nc * nao2);
        char *sb = NULL;
        TRY(buf_get(g_sbuf, xb + (keep && nc < n1 ? preby : 0), 1, &sb));
        Xc = (double *)sb;
        if (keep && nc < n1) {
            preb = sb + xb;
            s1pre_carve(&pre, preb, n2, n_ao);
        }
        cutc_plan pl;
        TRY(cutc_make_plan(nc, n2, n_ao, n_nuc, opt->precision, opt->slices, device_budget(opt), opt->n1_chunk, &pl));
        tr_init(&tr, st, stats != NULL || opt->verbose);
        have_tr = 1;
        int evt = tr_begin(&tr, CAT_TOTAL);
        rc = streamed_chunks(opt, &pl, n1, n2, n_ao, n_nuc, r1, wr1, r2, wr2, rn, aos1, aos2, &prog, out, Xc, nc,
                             split2_for(opt, pl.n_pairs, pl.n1_chunk_2e, split2_room(&pl, device_budget(opt))),
                             &tr, stats, preb ? &pre : NULL);
        tr_end(&tr, evt);
        if (rc == CUTC_OK)
            rc = hip_rc(hipStreamSynchronize(st));
        if (rc == CUTC_OK) {
            fill_stats(stats, &tr);
            if (stats) {
                stats->n1_chunk = nc;
                stats->n1_chunks = (int)((n1 + nc - 1) / nc);
            }
            if (opt->verbose)
                print_stage_times(&tr, 3, opt->precision != CUTC_PREC_FP64, 0);
        }
    }
done:
    hipStreamSynchronize(st);
    if (have_tr)
        tr_free(&tr);
    jprog_free(&prog);
    return rc;
}

int cutc_int_streamed_device(const cutc_options *opt, int n_grid1, int n_grid2, int n_ao, int n_nuc, int size_bh,
                             const double *r1, const double *wr1, const double *r2, const double *wr2,
                             const double *rn, const double *aos_data1, const double *aos_data2, const double *c_bh,
                             const int *m_bh, const int *n_bh, const int *o_bh, double *int_2e_ao, void *stream,
                             cutc_stats *stats)
{
    cutc_options def;
    if (!opt) {
        cutc_default_options(&def);
        opt = &def;
    }
    pthread_mutex_lock(&g_lock);
    int rc = streamed_path(opt, n_grid1, n_grid2, n_ao, n_nuc, size_bh, r1, wr1, r2, wr2, rn, aos_data1, aos_data2,
                           c_bh, m_bh, n_bh, o_bh, int_2e_ao, (hipStream_t)stream, stats);
    pthread_mutex_unlock(&g_lock);
    cutc_set_last_error(rc);
    return rc;
}

/* Batched build: every system of the batch is built streamed (its X in r1
 * chunks of at most opt->n1_chunk points, by default as many as keep the
 * chunk buffer within a quarter of the device memory, at least 4096; one
 * chunk buffer shared by the batch) into its own int_2e_ao.  All programs, plans, the
 * workspace and the chunk buffer are set up first; then the launches of all
 * systems are recorded into one hipGraph, which is instantiated and launched
 * once on the stream. */
#define CUTC_BATCH_CHUNK 4096

static int check_system(const cutc_system *y)
{
    if (check_sizes(y->n_grid1, y->n_grid2, y->n_ao, y->n_nuc, y->size_bh))
        return CUTC_ERR_ARG;
    if (y->n_ao > 0 && !y->int_2e_ao)
        return CUTC_ERR_ARG;
    if (y->n_grid1 > 0 && y->n_ao > 0 &&
        (!y->r1 || !y->wr1 || !y->aos_data1 ||
         (y->n_grid2 > 0 &&
          (!y->r2 || !y->wr2 || !y->aos_data2 || (y->n_nuc > 0 && !y->rn) ||
           ((int64_t)y->n_nuc * y->size_bh > 0 && (!y->c_bh || !y->m_bh || !y->n_bh || !y->o_bh))))))
        return CUTC_ERR_ARG;
    return CUTC_OK;
}

/* executed contraction work of a system (FP64-equivalent flops), to balance
 * the branches of a batch */
static double system_work(const cutc_system *y)
{
    const double P = 0.5 * (double)y->n_ao * (y->n_ao + 1);
    return 2.0 * P * 4.0 * y->n_grid1 * (double)y->n_grid2 + 2.0 * P * P * 4.0 * y->n_grid1;
}

#define CUTC_BATCH_MAX_BRANCHES 4

static int batch_path(const cutc_options *opt, int nsys, const cutc_system *sys, hipStream_t st,
                      cutc_batch_stats *bst)
{
    const cutc_options ropt = resolved_options(opt);
    opt = &ropt;
    if (!valid_precision(opt) || nsys < 0 || (nsys > 0 && !sys) || opt->streams < 0 ||
        opt->streams > CUTC_BATCH_MAX_BRANCHES)
        return CUTC_ERR_ARG;
    for (int i = 0; i < nsys; i++)
        if (check_system(&sys[i]))
            return CUTC_ERR_ARG;
    if (bst)
        memset(bst, 0, sizeof *bst);
    if (nsys == 0)
        return CUTC_OK;
    int rc = CUTC_OK, nb = 1, capturing = 0;
    jprog_t *jp = (jprog_t *)calloc((size_t)nsys, sizeof(jprog_t));
    cutc_plan *pl = (cutc_plan *)calloc((size_t)nsys, sizeof(cutc_plan));
    int64_t *nc = (int64_t *)calloc((size_t)nsys, sizeof(int64_t));
    int *split2 = (int *)calloc((size_t)nsys, sizeof(int));
    int *branch = (int *)calloc((size_t)nsys, sizeof(int));
    double *Xc[CUTC_BATCH_MAX_BRANCHES] = {NULL};
    hipStream_t bs[CUTC_BATCH_MAX_BRANCHES] = {NULL};
    hipEvent_t ef = NULL, ej[CUTC_BATCH_MAX_BRANCHES] = {NULL};
    tracer_t tr[CUTC_BATCH_MAX_BRANCHES];
    hipGraph_t graph = NULL;
    hipGraphExec_t exec = NULL;
    hipEvent_t e0 = NULL, e1 = NULL;
    memset(tr, 0, sizeof tr);
    if (!jp || !pl || !nc || !split2 || !branch) {
        rc = CUTC_ERR_NOMEM;
        goto done;
    }
    {
        /* X chunk: opt->n1_chunk when set, else as many r1 points as keep a
         * chunk buffer within a quarter of the device budget (whole systems,
         * usually: fewer chunks mean fewer step-2 passes) */
        int64_t cmax = opt->n1_chunk;
        if (cmax <= 0) {
            const size_t quarter = device_budget_plus(opt, buf_size_current(g_sbuf)) / 4;
            int64_t hi = 1;
            for (int i = 0; i < nsys; i++)
                if (sys[i].n_grid1 > hi)
                    hi = sys[i].n_grid1;
            cmax = hi;
            for (int i = 0; i < nsys; i++) {
                const size_t xper = sizeof(double) * 4 * (size_t)sys[i].n_ao * sys[i].n_ao;
                if (xper > 0 && (size_t)cmax * xper > quarter)
                    cmax = (int64_t)(quarter / xper);
            }
            if (cmax < CUTC_BATCH_CHUNK)
                cmax = CUTC_BATCH_CHUNK;
        }
        size_t xbytes = 0;
        for (int i = 0; i < nsys; i++) {
            const int64_t n1 = sys[i].n_grid1, nch = (n1 + cmax - 1) / cmax;
            nc[i] = nch > 0 ? (n1 + nch - 1) / nch : 0; /* balanced, as the streamed call */
            const size_t b = sizeof(double) * 4 * (size_t)nc[i] * sys[i].n_ao * sys[i].n_ao;
            if (b > xbytes)
                xbytes = b;
        }
        const double t0 = now_ms();
        /* plans against the device budget; the workspace of a branch is the
         * largest stage of the batch */
        /* (the chunk buffers of an earlier batch or streamed call count as free) */
        const size_t budget = device_budget_plus(opt, buf_size_current(g_sbuf));
        size_t need = 0;
        for (int i = 0; i < nsys; i++) {
            const cutc_system *y = &sys[i];
            if (y->n_grid1 == 0 || y->n_ao == 0)
                continue;
            /* G chunks beyond a few thousand r1 points gain nothing and only
             * grow the workspace the batch allocates up front */
            TRY(cutc_make_plan(nc[i], y->n_grid2, y->n_ao, y->n_nuc, opt->precision, opt->slices, budget,
                               nc[i] < 8192 ? nc[i] : 8192, &pl[i]));
            TRY(jprog_load(y->n_nuc, y->size_bh, y->c_bh, y->m_bh, y->n_bh, y->o_bh, y->n_grid2 > 0, &jp[i]));
            /* the autotuner picks the K split of the FP64 step-2 contraction;
             * the modular ones split as in a single call */
            split2[i] = opt->precision == CUTC_PREC_FP64
                            ? tune_split2(pl[i].n_pairs, y->n_ao, pl[i].n1_chunk_2e, budget, st,
                                          bst ? &bst->ms_tune : NULL)
                            : split2_for(opt, pl[i].n_pairs, pl[i].n1_chunk_2e, budget / 16);
            /* stage 1 also holds the program's per-r1 table, stage 2 the split
             * partial sums (FP64) or the further parts' residues (modular) */
            const int64_t P2 = pl[i].n_pairs * pl[i].n_pairs;
            const size_t b1 = pl[i].bytes_stage1 + sizeof(double) * ((size_t)jp[i].h.ts * pl[i].n1_chunk + 1) + 4096;
            const size_t b2 = pl[i].bytes_stage2 +
                              (split2[i] <= 1 ? 0
                               : opt->precision == CUTC_PREC_FP64
                                   ? sizeof(double) * split2[i] * P2 + 256
                                   : (size_t)(split2[i] - 1) * opt->slices * P2 + 256);
            const size_t b = b1 > b2 ? b1 : b2;
            if (b > need)
                need = b;
        }
        need = ws_round(need + need / 64 + ((size_t)16 << 20));
        /* branches: opt->streams, by default 1 for FP64 and the slice forms
         * (on the 64-system benchmark batch two branches were slower, 40.2
         * against 39.7 s in FP64: every system's dense contractions fill the
         * device and two compete for L2 and power) and 4 for the modular
         * forms, whose list-driven contractions leave room (int8: 8.11 s on
         * one branch, 7.82 on two, 7.53 on three); as many as the device
         * holds a workspace region and a chunk buffer for */
        nb = opt->streams > 0 ? opt->streams : (is_crt(opt->precision) ? CUTC_BATCH_MAX_BRANCHES : 1);
        if (nb > nsys)
            nb = nsys;
        while (nb > 1 && (size_t)nb * (need + ws_round(xbytes)) > budget)
            nb--;
        /* each system goes to the branch with the least work so far */
        double load[CUTC_BATCH_MAX_BRANCHES] = {0.0};
        for (int i = 0; i < nsys; i++) {
            int b = 0;
            for (int r = 1; r < nb; r++)
                if (load[r] < load[b])
                    b = r;
            branch[i] = b;
            load[b] += system_work(&sys[i]);
        }
        /* the branches' chunk buffers in the grow-only buffer of the
         * streamed calls (at most twice what they need) */
        if (xbytes) {
            char *sb = NULL;
            TRY(buf_get(g_sbuf, (size_t)nb * ws_round(xbytes), 1, &sb));
            for (int r = 0; r < nb; r++)
                Xc[r] = (double *)(sb + (size_t)r * ws_round(xbytes));
        }
        ws_t ws;
        TRY(ws_begin(&ws, (size_t)nb * need));
        /* record on private streams (the legacy default stream cannot be
         * captured), the first ordered after the caller's stream, the others
         * forked from it inside the capture and joined back at the end */
        for (int r = 0; r < nb; r++)
            TRYH(hipStreamCreateWithFlags(&bs[r], hipStreamNonBlocking));
        TRYH(hipEventCreate(&e0));
        TRYH(hipEventCreate(&e1));
        TRYH(hipEventCreateWithFlags(&ef, hipEventDisableTiming));
        for (int r = 1; r < nb; r++)
            TRYH(hipEventCreateWithFlags(&ej[r], hipEventDisableTiming));
        TRYH(hipEventRecord(e0, st));
        TRYH(hipStreamWaitEvent(bs[0], e0, 0));
        for (int r = 0; r < nb; r++) {
            tr_init(&tr[r], bs[r], 0);
            tr[r].capture = 1;
        }
        g_ws_frozen = 1;
        TRYH(hipStreamBeginCapture(bs[0], hipStreamCaptureModeThreadLocal));
        capturing = 1;
        TRYH(hipEventRecord(ef, bs[0]));
        for (int r = 1; r < nb; r++)
            TRYH(hipStreamWaitEvent(bs[r], ef, 0));
        for (int i = 0; i < nsys && rc == CUTC_OK; i++) {
            const cutc_system *y = &sys[i];
            const int b = branch[i];
            const int64_t nao2 = (int64_t)y->n_ao * y->n_ao;
            if (y->n_ao == 0)
                continue;
            /* the branch's own workspace region, fixed while its launches are recorded */
            g_ws_fixed.base = ws.base + (size_t)b * need;
            g_ws_fixed.size = need;
            if (y->n_grid1 == 0)
                rc = cutc_launch_fill(y->int_2e_ao, nao2 * nao2, 0.0, bs[b]);
            else
                rc = streamed_chunks(opt, &pl[i], y->n_grid1, y->n_grid2, y->n_ao, y->n_nuc, y->r1, y->wr1, y->r2,
                                     y->wr2, y->rn, y->aos_data1, y->aos_data2, &jp[i], y->int_2e_ao, Xc[b], nc[i],
                                     split2[i], &tr[b], NULL, NULL);
            g_ws_fixed.base = NULL;
            g_ws_fixed.size = 0;
        }
        for (int r = 1; r < nb && rc == CUTC_OK; r++) {
            rc = hip_rc(hipEventRecord(ej[r], bs[r]));
            if (rc == CUTC_OK)
                rc = hip_rc(hipStreamWaitEvent(bs[0], ej[r], 0));
        }
        const int rc_end = hip_rc(hipStreamEndCapture(bs[0], &graph));
        capturing = 0;
        g_ws_frozen = 0;
        TRY(rc);
        TRY(rc_end);
        TRYH(hipGraphInstantiate(&exec, graph, NULL, NULL, 0));
        if (bst) {
            bst->ms_setup = now_ms() - t0;
            size_t nn = 0;
            if (hipGraphGetNodes(graph, NULL, &nn) == hipSuccess)
                bst->n_nodes = (int64_t)nn;
        }
        TRYH(hipEventRecord(e0, bs[0]));
        TRYH(hipGraphLaunch(exec, bs[0]));
        TRYH(hipEventRecord(e1, bs[0]));
        TRYH(hipStreamSynchronize(bs[0]));
        float ms = 0.0f;
        TRYH(hipEventElapsedTime(&ms, e0, e1));
        if (bst) {
            bst->ms_launch = ms;
            bst->n_systems = nsys;
            bst->n_branches = nb;
            for (int i = 0; i < nsys; i++)
                bst->n_split += split2[i] > 1;
        }
    }
done:
    g_ws_frozen = 0;
    g_ws_fixed.base = NULL;
    g_ws_fixed.size = 0;
    if (capturing) { /* an error while recording: end the capture, drop its graph */
        hipGraph_t g = NULL;
        if (hipStreamEndCapture(bs[0], &g) == hipSuccess && g)
            hipGraphDestroy(g);
        (void)hipGetLastError();
    }
    for (int r = 0; r < CUTC_BATCH_MAX_BRANCHES; r++) {
        if (bs[r]) {
            hipStreamSynchronize(bs[r]);
            hipStreamDestroy(bs[r]);
        }
        if (ej[r])
            hipEventDestroy(ej[r]);
        tr_free(&tr[r]);
    }
    if (exec)
        hipGraphExecDestroy(exec);
    if (graph)
        hipGraphDestroy(graph);
    if (e0)
        hipEventDestroy(e0);
    if (e1)
        hipEventDestroy(e1);
    if (ef)
        hipEventDestroy(ef);
    if (jp)
        for (int i = 0; i < nsys; i++)
            jprog_free(&jp[i]);
    free(jp);
    free(pl);
    free(nc);
    free(split2);
    free(branch);
    return rc;
}

int cutc_int_batch_device(const cutc_options *opt, int n_systems, const cutc_system *systems, void *stream,
                          cutc_batch_stats *stats)
{
    cutc_options def;
    if (!opt) {
        cutc_default_options(&def);
        opt = &def;
    }
    pthread_mutex_lock(&g_lock);
    int rc = batch_path(opt, n_systems, systems, (hipStream_t)stream, stats);
    pthread_mutex_unlock(&g_lock);
    cutc_set_last_error(rc);
    return rc;
}

void get_int2_grad1_u12_ao(cutc_dim3 dimGrid, cutc_dim3 dimBlock, int n_grid1, int n_grid2, int n_ao, int n_nuc,
                           int size_bh, double *r1, double *r2, double *wr2, double *rn, double *aos_data2,
                           double *c_bh, int *m_bh, int *n_bh, int *o_bh, double *int2_grad1_u12_ao)
{
    (void)dimGrid;
    (void)dimBlock;
    cutc_options opt;
    cutc_default_options(&opt);
    opt.verbose = 1; /* the reference prints its plan and stage times */
    pthread_mutex_lock(&g_lock);
    int rc = device_path(&opt, 1, n_grid1, n_grid2, n_ao, n_nuc, size_bh, r1, NULL, r2, wr2, rn, NULL, aos_data2,
                         c_bh, m_bh, n_bh, o_bh, int2_grad1_u12_ao, NULL, NULL, NULL, NULL, NULL);
    release_arenas(); /* the reference keeps no device memory between calls */
    pthread_mutex_unlock(&g_lock);
    cutc_set_last_error(rc);
    if (rc)
        fprintf(stderr, "get_int2_grad1_u12_ao: %s\n", cutc_error_string(rc));
}

void get_int_2e_ao(int n_grid1, int n_ao, double *wr1, double *aos_data1, double *int2_grad1_u12_ao,
                   double *tc_int_2e_ao)
{
    cutc_options opt;
    cutc_default_options(&opt);
    opt.verbose = 1; /* the reference prints its stage times */
    pthread_mutex_lock(&g_lock);
    int rc = device_path(&opt, 2, n_grid1, 0, n_ao, 0, 0, NULL, wr1, NULL, NULL, NULL, aos_data1, NULL, NULL, NULL,
                         NULL, NULL, int2_grad1_u12_ao, tc_int_2e_ao, NULL, NULL, NULL, NULL);
    release_arenas();
    pthread_mutex_unlock(&g_lock);
    cutc_set_last_error(rc);
    if (rc)
        fprintf(stderr, "get_int_2e_ao: %s\n", cutc_error_string(rc));
}

/* ---------------------------------------------------- host-array path */

static int geometry_ok(int nx, int ny, int nz, int bx, int by, int bz)
{
    int dev = 0;
    hipDeviceProp_t pr;
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&pr, dev) != hipSuccess)
        return CUTC_ERR_HIP;
    if (nx < 0 || ny < 0 || nz < 0 || bx < 0 || by < 0 || bz < 0)
        return CUTC_ERR_GEOMETRY;
    if ((long long)bx * by * bz > pr.maxThreadsPerBlock) {
        printf("Error: Too many threads per block!\n");
        return CUTC_ERR_GEOMETRY;
    }
    if (nx > pr.maxGridSize[0] || ny > pr.maxGridSize[1] || nz > pr.maxGridSize[2]) {
        printf("Error: Grid dimensions exceed device capabilities!\n");
        return CUTC_ERR_GEOMETRY;
    }
    return CUTC_OK;
}

/* carve the next 256-byte-aligned piece of bytes from the host-array buffer
 * and fill it from h (h == NULL: leave it for the build) */
static int carve_upload(char *base, size_t *off, void **d, const void *h, size_t bytes, int upload)
{
    *d = NULL;
    if (bytes == 0)
        return CUTC_OK;
    if (upload && !h)
        return CUTC_ERR_ARG;
    *d = base + *off;
    *off += ws_round(bytes);
    return upload ? hip_rc(hipMemcpy(*d, h, bytes, hipMemcpyHostToDevice)) : CUTC_OK;
}

/* Pipelined download of X by the host entry points: the rows of each r1
 * chunk (3 components; the 4th stays on the device, as in the reference)
 * are queued on the transfer pipeline as soon as the build has queued the
 * work that writes them. */
typedef struct {
    cutc_xfer *xf;
    double *h_X;
    const double *X;
    int64_t n1, nao2, queued;
    int rc;
} xdownload_t;

static void x_chunk_ready(void *ctx, int64_t i0, int64_t n1e, hipStream_t st)
{
    xdownload_t *d = (xdownload_t *)ctx;
    for (int m = 0; m < 3 && d->rc == CUTC_OK; m++) {
        const int64_t off = ((int64_t)m * d->n1 + i0) * d->nao2;
        d->rc = cutc_xfer_d2h(d->xf, d->h_X + off, d->X + off, sizeof(double) * (size_t)(n1e * d->nao2), st);
    }
    d->queued = i0 + n1e;
}

/* host copy workers beside the driver thread: CUTC_HOST_COPY_THREADS if set
 * (0..16), else one less than the online CPUs, at most 15 */
int cutc_internal_copy_threads(void)
{
    const char *e = getenv("CUTC_HOST_COPY_THREADS");
    if (e && *e) {
        const long v = strtol(e, NULL, 10);
        return (int)(v < 0 ? 0 : v > 16 ? 16 : v);
    }
    long n = sysconf(_SC_NPROCESSORS_ONLN);
    n = (n > 1) ? n - 1 : 0;
    return (int)(n > 15 ? 15 : n);
}

static int host_path(const cutc_options *opt, int n1, int n2, int n_ao, int n_nuc, int size_bh, const double *h_r1,
                     const double *h_wr1, const double *h_r2, const double *h_wr2, const double *h_rn,
                     const double *h_aos1, const double *h_aos2, const double *h_c, const int *h_m, const int *h_n,
                     const int *h_o, double *h_X, double *h_out, cutc_stats *stats)
{
    int rc = check_sizes(n1, n2, n_ao, n_nuc, size_bh);
    if (rc)
        return rc;
    const int64_t nao2 = (int64_t)n_ao * n_ao;
    const int64_t nt = (int64_t)n_nuc * size_bh;
    if ((n1 > 0 && nao2 > 0 && !h_X) || (nao2 > 0 && !h_out))
        return CUTC_ERR_ARG;
    if (nt > 0 && (!h_m || !h_n || !h_o))
        return CUTC_ERR_ARG;
    if (nt > 0 && (rc = check_exponents(h_m, h_n, h_o, nt)) != CUTC_OK)
        return rc;
    double *r1 = NULL, *wr1 = NULL, *r2 = NULL, *wr2 = NULL, *rn = NULL, *aos1 = NULL, *aos2 = NULL, *c = NULL;
    int *m = NULL, *n = NULL, *o = NULL;
    double *X = NULL, *out = NULL;
    hipEvent_t e0 = NULL, e1 = NULL, e2 = NULL, e3 = NULL;
    xdownload_t dl = {0};
    tracer_t hook;
    memset(&hook, 0, sizeof hook);
    const double t_start = now_ms();
    TRYH(hipEventCreate(&e0));
    TRYH(hipEventCreate(&e1));
    TRYH(hipEventCreate(&e2));
    TRYH(hipEventCreate(&e3));
    TRYH(hipEventRecord(e0, NULL));
    {
        /* every device array of the call in one cached buffer (hbuf_get) */
        void **dst[13] = {(void **)&r1, (void **)&wr1, (void **)&r2, (void **)&wr2, (void **)&rn, (void **)&aos1,
                          (void **)&aos2, (void **)&c, (void **)&m, (void **)&n, (void **)&o, (void **)&X,
                          (void **)&out};
        const void *src[13] = {h_r1, h_wr1, h_r2, h_wr2, h_rn, h_aos1, h_aos2, h_c, h_m, h_n, h_o, NULL, NULL};
        const size_t d8 = sizeof(double), i4 = sizeof(int);
        const size_t bytes[13] = {d8 * 3 * n1, d8 * n1, d8 * 3 * n2, d8 * n2, d8 * 3 * n_nuc,
                                  d8 * 4 * (size_t)n1 * n_ao, d8 * 4 * (size_t)n2 * n_ao, d8 * nt, i4 * nt, i4 * nt,
                                  i4 * nt, d8 * 4 * (size_t)(nao2 * n1), d8 * (size_t)(nao2 * nao2)};
        size_t need = 0, off = 0;
        for (int i = 0; i < 13; i++)
            need += ws_round(bytes[i]);
        char *base = NULL;
        if (need > 0)
            TRY(hbuf_get(need, &base));
        for (int i = 0; i < 13; i++)
            TRY(carve_upload(base, &off, dst[i], src[i], bytes[i], i < 11));
    }
    if (!opt->host_copy && (X || out)) {
        TRY(cutc_xfer_open(&dl.xf, (size_t)64 << 20, 4, cutc_internal_copy_threads()));
        dl.h_X = h_X;
        dl.X = X;
        dl.n1 = n1;
        dl.nao2 = nao2;
        if (X) {
            hook.x_ready = x_chunk_ready;
            hook.x_ctx = &dl;
        }
    }
    TRYH(hipEventRecord(e1, NULL));
    const double t_in = now_ms();
    TRY(device_path(opt, 3, n1, n2, n_ao, n_nuc, size_bh, r1, wr1, r2, wr2, rn, aos1, aos2, c, m, n, o, X, out, NULL,
                    stats, &hook, NULL));
    const double t_built = now_ms();
    TRYH(hipEventRecord(e2, NULL));
    if (dl.xf) {
        if (X && dl.queued < n1) /* rows written without r1 chunks (empty grid 2) */
            x_chunk_ready(&dl, dl.queued, n1 - dl.queued, NULL);
        TRY(dl.rc);
        if (out)
            TRY(cutc_xfer_d2h(dl.xf, h_out, out, sizeof(double) * (size_t)(nao2 * nao2), NULL));
        TRY(cutc_xfer_finish(dl.xf, NULL));
    } else {
        if (X)
            TRYH(hipMemcpy(h_X, X, sizeof(double) * 3 * nao2 * n1, hipMemcpyDeviceToHost));
        if (out)
            TRYH(hipMemcpy(h_out, out, sizeof(double) * nao2 * nao2, hipMemcpyDeviceToHost));
    }
    TRYH(hipEventRecord(e3, NULL));
    TRYH(hipEventSynchronize(e3));
    if (stats) {
        stats->ms_host_in = t_in - t_start;
        stats->ms_host_out = now_ms() - t_built;
    }
    if (opt->verbose) {
        float t_up = 0, t_comp = 0, t_down = 0;
        hipEventElapsedTime(&t_up, e0, e1);
        hipEventElapsedTime(&t_comp, e1, e2);
        hipEventElapsedTime(&t_down, e2, e3);
        printf("Ellapsed time for Device <-> Host transf = %.3f sec\n", (t_up + t_down) / 1000.0);
        printf("Ellapsed (total) time on GPU for cutc_int = %.3f sec\n", (t_up + t_comp + t_down) / 1000.0);
        fflush(stdout);
    }
done:
    cutc_xfer_close(dl.xf); /* drains queued copies before the buffers are reused */
    if (rc != CUTC_OK)
        hipDeviceSynchronize();
    if (e0)
        hipEventDestroy(e0);
    if (e1)
        hipEventDestroy(e1);
    if (e2)
        hipEventDestroy(e2);
    if (e3)
        hipEventDestroy(e3);
    return rc;
}

int cutc_int(int nxBlocks, int nyBlocks, int nzBlocks, int blockxSize, int blockySize, int blockzSize, int n_grid1,
             int n_grid2, int n_ao, int n_nuc, int size_bh, double *h_r1, double *h_wr1, double *h_r2, double *h_wr2,
             double *h_rn, double *h_aos_data1, double *h_aos_data2, double *h_c_bh, int *h_m_bh, int *h_n_bh,
             int *h_o_bh, double *h_int2_grad1_u12_ao, double *h_int_2e_ao)
{
    printf(" Computing TC-Integrals With CuTC\n");
    int rc = geometry_ok(nxBlocks, nyBlocks, nzBlocks, blockxSize, blockySize, blockzSize);
    if (rc == CUTC_OK) {
        cutc_options opt;
        cutc_default_options(&opt);
        opt.verbose = 1;
        pthread_mutex_lock(&g_lock);
        rc = host_path(&opt, n_grid1, n_grid2, n_ao, n_nuc, size_bh, h_r1, h_wr1, h_r2, h_wr2, h_rn, h_aos_data1,
                       h_aos_data2, h_c_bh, h_m_bh, h_n_bh, h_o_bh, h_int2_grad1_u12_ao, h_int_2e_ao, NULL);
        release_arenas(); /* the reference keeps no device memory between calls */
        pthread_mutex_unlock(&g_lock);
    }
    cutc_set_last_error(rc);
    return rc;
}

int cutc_int_ex(const cutc_options *opt, int n_grid1, int n_grid2, int n_ao, int n_nuc, int size_bh,
                const double *h_r1, const double *h_wr1, const double *h_r2, const double *h_wr2, const double *h_rn,
                const double *h_aos_data1, const double *h_aos_data2, const double *h_c_bh, const int *h_m_bh,
                const int *h_n_bh, const int *h_o_bh, double *h_int2_grad1_u12_ao, double *h_int_2e_ao,
                cutc_stats *stats)
{
    cutc_options def;
    if (!opt) {
        cutc_default_options(&def);
        opt = &def;
    }
    const cutc_options ropt = resolved_options(opt);
    opt = &ropt;
    int rc = valid_precision(opt) ? CUTC_OK : CUTC_ERR_ARG;
    if (rc == CUTC_OK) {
        pthread_mutex_lock(&g_lock);
        rc = host_path(opt, n_grid1, n_grid2, n_ao, n_nuc, size_bh, h_r1, h_wr1, h_r2, h_wr2, h_rn, h_aos_data1,
                       h_aos_data2, h_c_bh, h_m_bh, h_n_bh, h_o_bh, h_int2_grad1_u12_ao, h_int_2e_ao, stats);
        pthread_mutex_unlock(&g_lock);
    }
    cutc_set_last_error(rc);
    return rc;
}

/* ------------------------------------------ exported launch wrappers */

static void wrapper_status(int rc, const char *name)
{
    if (rc == CUTC_OK)
        rc = hip_rc(hipDeviceSynchronize());
    cutc_set_last_error(rc);
    if (rc)
        fprintf(stderr, "%s: %s\n", name, cutc_error_string(rc));
}

void cutc_int_bh(cutc_dim3 dimGrid, cutc_dim3 dimBlock, int ii0, int n_grid1_eff, int n_grid1_tot, int jj0,
                 int n_grid2_eff, int n_grid2_tot, int n_nuc, int size_bh, double *r1, double *r2, double *rn,
                 double *c_bh, int *m_bh, int *n_bh, int *o_bh, double *grad1_u12)
{
    (void)dimGrid;
    (void)dimBlock;
    double *f2tab = NULL, *tab = NULL;
    cutc_jmono *dmono = NULL;
    int *dcls = NULL;
    cutc_jprog prog;
    int rc = CUTC_OK;
    if (n_grid1_eff < 0 || n_grid2_eff < 0 || n_nuc < 0 || size_bh < 0)
        rc = CUTC_ERR_ARG;
    if (rc == CUTC_OK)
        rc = program_from_device(n_nuc, size_bh, c_bh, m_bh, n_bh, o_bh, &prog);
    else
        memset(&prog, 0, sizeof prog);
    if (rc == CUTC_OK && n_nuc > 32 && n_grid2_eff > 0)
        rc = hip_rc(hipMalloc((void **)&f2tab, sizeof(double) * n_nuc * (size_t)n_grid2_eff));
    if (rc == CUTC_OK)
        rc = hip_rc(hipMalloc((void **)&tab, sizeof(double) * (prog.ts * (int64_t)n_grid1_eff + 1)));
    if (rc == CUTC_OK)
        rc = program_upload(&prog, &dmono, &dcls);
    if (rc == CUTC_OK)
        rc = cutc_launch_jast_prep(r1, ii0, n_grid1_eff, rn, n_nuc, dmono, prog.n_mono, tab, prog.ts, NULL);
    if (rc == CUTC_OK)
        rc = cutc_launch_jast_pairs(r1, r2, rn, n_nuc, tab, prog.ts, dcls, prog.n_cls, prog.kmax, ii0, n_grid1_eff,
                                    jj0, n_grid2_eff, n_grid2_eff, grad1_u12, n_grid2_tot,
                                    (int64_t)n_grid1_tot * n_grid2_tot, f2tab, NULL);
    wrapper_status(rc, "cutc_int_bh");
    dfree(f2tab);
    dfree(tab);
    dfree(dmono);
    dfree(dcls);
    cutc_jast_program_free(&prog);
}

void int_long_range(int jj0, int n_grid2_eff, int n_grid2_tot, int n_grid2, int n_ao, double *wr2, double *aos_data2,
                    double *int_fct_long_range)
{
    wrapper_status(cutc_launch_long_range_ref(jj0, n_grid2_eff, n_grid2_tot, n_grid2, n_ao, wr2, aos_data2,
                                              int_fct_long_range, NULL),
                   "int_long_range");
}

void int_short_range_herm(int n_grid1, int n_ao, double *wr1, double *aos_data1, double *int_fct_short_range_herm)
{
    wrapper_status(cutc_launch_herm_ref(n_grid1, n_ao, wr1, aos_data1, int_fct_short_range_herm, NULL),
                   "int_short_range_herm");
}

void int_short_range_nonherm(int n_grid1, int n_ao, double *wr1, double *aos_data1,
                             double *int_fct_short_range_nonherm)
{
    wrapper_status(cutc_launch_nonherm_ref(n_grid1, n_ao, wr1, aos_data1, int_fct_short_range_nonherm, NULL),
                   "int_short_range_nonherm");
}

void trans_inplace(double *data, int size)
{
    wrapper_status(cutc_launch_trans_inplace(data, size, NULL), "trans_inplace");
}
