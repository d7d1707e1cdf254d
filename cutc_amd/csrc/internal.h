/* Internal interface between the host files of the library (defined in
 * cutc_int.c): the library lock, the device workspace arena and the memory
 * budget, the thread-local last error and the copy-thread count. */
#ifndef CUTC_INTERNAL_H
#define CUTC_INTERNAL_H

#include <stddef.h>

void cutc_set_last_error(int code);
void cutc_internal_lock(void);
void cutc_internal_unlock(void);
/* free device memory (plus the arena already held) less a 1 GiB margin,
 * capped by mem_limit when that is nonzero */
size_t cutc_internal_budget(size_t mem_limit);
/* make the arena at least need bytes; *base: its start */
/* free the cached workspace and host-array buffers of every device (the
 * caller holds the library lock) */
void cutc_internal_release_workspace(void);
int cutc_internal_workspace(size_t need, char **base);
/* host copy threads of the transfer pipeline (CUTC_HOST_COPY_THREADS or the
 * CPU count) */
int cutc_internal_copy_threads(void);

#endif
