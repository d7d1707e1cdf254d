/* Error helpers and per-thread status of libcutcint. */
#include <stdio.h>
#include <stdlib.h>

#include "cutc_int.h"

static _Thread_local int g_last_error = 0;

void cutc_set_last_error(int code) { g_last_error = code; }
int cutc_get_last_error(void) { return g_last_error; }

const char *cutc_error_string(int code)
{
    switch (code) {
    case CUTC_OK: return "success";
    case CUTC_ERR_GEOMETRY: return "launch geometry exceeds the device limits";
    case CUTC_ERR_ARG: return "invalid argument";
    case CUTC_ERR_NOMEM: return "working set does not fit in device memory";
    case CUTC_ERR_HIP: return "HIP runtime error";
    case CUTC_ERR_UNSUPPORTED: return "unsupported shape or option";
    }
    return "unknown error";
}

/* The reference prints and calls exit(0) (src/utils.cu:8-14); a failure
 * status is returned to the shell here instead. */
void checkCudaErrors(int err, const char *msg, const char *file, int line)
{
    if (err != 0) {
        printf("CUDA Error in %s at line %d\n", file, line);
        printf("%s - error %d\n", msg, err);
        fflush(stdout);
        exit(1);
    }
}

/* src/utils.cu:43-51 (status 0 is success). */
void checkCublasErrors(int status, const char *msg, const char *file, int line)
{
    if (status != 0) {
        printf("CUBLAS Error in %s at line %d\n", file, line);
        printf("%s - status %d\n", msg, status);
        fflush(stdout);
        exit(1);
    }
}
