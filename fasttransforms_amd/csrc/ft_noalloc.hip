// ft_noalloc.hip -- the device scratch of the allocation-free entry point
// (ft_gpu_execute_noalloc): a file-scope device array per GPU instead of
// memory the library allocates, and its device address.
//
// In host code `scratch` names a host-side stand-in, not device memory; its
// device address on the current device comes from hipGetSymbolAddress, once
// per device, and only that address is ever handed to kernels or copies.
#include <hip/hip_runtime.h>
#include <pthread.h>

#include "ft_hip_abi.h"

static __device__ __align__(256) unsigned char scratch[FTAMD_NOALLOC_SCRATCH_BYTES];

static pthread_mutex_t g_scratch_lock = PTHREAD_MUTEX_INITIALIZER;
static void *g_scratch_addr[FTAMD_MAX_DEVICES];

// Device address of the current device's scratch (looked up on the first call
// on that device, which also zeroes its error area with a plain copy).
extern "C" int ftamd_hip_scratch(void **ptr, size_t *bytes) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= FTAMD_MAX_DEVICES) return FTAMD_EDEVICE;
    pthread_mutex_lock(&g_scratch_lock);
    int rc = FTAMD_OK;
    if (!g_scratch_addr[dev]) {
        void *p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e == hipSuccess && p) {
            static const unsigned char zeros[256] = {0};
            e = hipMemcpy(p, zeros, sizeof zeros, hipMemcpyHostToDevice);
        }
        if (e == hipSuccess && p) g_scratch_addr[dev] = p;
        else rc = FTAMD_EDEVICE;
    }
    *ptr = g_scratch_addr[dev];
    *bytes = FTAMD_NOALLOC_SCRATCH_BYTES;
    pthread_mutex_unlock(&g_scratch_lock);
    return rc;
}
