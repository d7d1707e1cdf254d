// Device scratch of the allocation-free entry point cutc_int_device_noalloc:
// one file-scope array of CUTC_SCRATCH_BYTES per GPU, placed by the HIP
// runtime when the library's code object is loaded on that GPU and held
// until the process ends.  Host code never passes the array's name to a
// kernel (in host code it is only a stand-in for the symbol): it takes the
// device address from cutc_scratch(), which asks the runtime for it once
// per device, with that device current, and keeps it.
#include <hip/hip_runtime.h>

#include "launch.h"

static __device__ __align__(256) char scratch[CUTC_SCRATCH_BYTES];

#define SCRATCH_MAX_DEV 64
static void *g_scratch_ptr[SCRATCH_MAX_DEV];

extern "C" int cutc_scratch(void **ptr, size_t *bytes)
{
    int dev = 0;
    *ptr = NULL;
    *bytes = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= SCRATCH_MAX_DEV)
        return CUTC_ERR_HIP;
    if (!g_scratch_ptr[dev]) {
        void *p = NULL;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p)
            return CUTC_ERR_HIP;
        g_scratch_ptr[dev] = p;
    }
    *ptr = g_scratch_ptr[dev];
    *bytes = CUTC_SCRATCH_BYTES;
    return CUTC_OK;
}
