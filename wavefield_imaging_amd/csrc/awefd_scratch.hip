/*
 * Device scratch of the allocation-free run (awefd_run_noalloc): a fixed
 * file-scope array, one instance per GPU, that stages the receiver traces of
 * the time loop.  It is created with the code object on each device the
 * library runs on and lives until the process ends; nothing here allocates.
 *
 * In host code `scratch` names a host-side stand-in, not a device address.
 * The device address of the current GPU's copy comes from
 * hipGetSymbolAddress, once per device (each GPU has its own copy at its own
 * address), and host code uses only that address: kernel arguments, copies.
 */
#include <hip/hip_runtime.h>
#include <pthread.h>

#include "awefd.h"
#include "awefd_internal.h"

static_assert(AWEFD_NOALLOC_TRACE_BYTES % 4 == 0, "float scratch");

static __device__ __align__(256) float scratch[AWEFD_NOALLOC_TRACE_BYTES / 4];

#define SCRATCH_MAX_DEV 64
static float *scratch_ptr[SCRATCH_MAX_DEV];
static pthread_mutex_t scratch_lock = PTHREAD_MUTEX_INITIALIZER;

/* Device address of device `dev`'s scratch; `dev` must be the current
 * device.  Returns a hipError_t. */
extern "C" int awefd__scratch(int dev, float **out)
{
    int cur = -1;
    hipError_t e = hipGetDevice(&cur);
    if (e != hipSuccess) return (int)e;
    if (dev < 0 || dev >= SCRATCH_MAX_DEV || cur != dev) return (int)hipErrorInvalidDevice;
    pthread_mutex_lock(&scratch_lock);
    if (!scratch_ptr[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e == hipSuccess && !p) e = hipErrorInvalidSymbol;
        if (e == hipSuccess) scratch_ptr[dev] = (float *)p;
    }
    *out = scratch_ptr[dev];
    pthread_mutex_unlock(&scratch_lock);
    return (int)e;
}
