/* Host <-> device transfer pipeline (xfer.c): pinned staging ring, a
 * non-blocking copy stream and host copy threads; downloads are fed with
 * blocks of device memory as the build produces them, uploads are waited for
 * by ticket. */
#ifndef CUTC_XFER_H
#define CUTC_XFER_H

#include <stddef.h>
#include <stdint.h>

#include <hip/hip_runtime_api.h>

typedef struct cutc_xfer cutc_xfer;

/* nstage pinned buffers of stage_bytes (2..8), nthreads host copy workers
 * besides the driver thread (0..16) */
int cutc_xfer_open(cutc_xfer **x, size_t stage_bytes, int nstage, int nthreads);
/* copy bytes from d_src to h_dst once the work queued on producer so far
 * has run (an event is recorded on producer now) */
int cutc_xfer_d2h(cutc_xfer *x, void *h_dst, const void *d_src, size_t bytes, hipStream_t producer);
/* queue a copy of rows x width bytes from host memory (row pitch spitch)
 * to device memory (row pitch dpitch); width <= the staging size.  *ticket
 * numbers it for cutc_xfer_wait (0: nothing to copy). */
int cutc_xfer_h2d_2d(cutc_xfer *x, void *d_dst, size_t dpitch, const void *h_src, size_t spitch, size_t width,
                     size_t rows, int64_t *ticket);
/* queue a contiguous upload of bytes (as 1 MiB rows plus a short one) */
int cutc_xfer_h2d(cutc_xfer *x, void *d_dst, const void *h_src, size_t bytes, int64_t *ticket);
/* block until the upload with this ticket and every earlier one are in
 * device memory; returns the first error */
int cutc_xfer_wait(cutc_xfer *x, int64_t ticket);
/* wait until every queued block is in host memory; returns the first error */
int cutc_xfer_finish(cutc_xfer *x, size_t *bytes_done);
void cutc_xfer_close(cutc_xfer *x);

#endif
