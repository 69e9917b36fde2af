// ft_rotations.hip -- gfx950 kernels of the Givens-rotation stage and the thin
// C-ABI layer the C host code (ft_drivers.c) calls.
//
// What is computed (src/rotations/rotations_default.c:31-71): for every
// column, a chain of rotation sequences.  Sequence j rotates the row pairs
// (k, k+D) for k running over the sequence's rows, with s = s(k,j), c = c(k,j)
// from the plan; hi2lo applies sequences in decreasing j with k descending,
// lo2hi in increasing j with k ascending and the transposed rotation.
// sph: D = 2, j steps by 2; tri: D = 1, j steps by 1; disk: D = 1, j steps by 2.
//
// This file holds the pipelined engine's launcher and the HIP runtime calls
// of the C ABI; the engines themselves are templates in ft_pipeline.hpp
// (lanes are columns, waves are pipeline stages of L sequences) and
// ft_systolic.hpp (lanes are sequences, rows stream through a band of lanes).
// In the pipelined engine:
//
//  * Lanes are columns.  A "group" is up to 64 columns that need the same
//    sequences (equal order parity for sph/disk, consecutive orders for tri),
//    so every lane applies the same rotation (k, j) at a time.  A tile's
//    (s, c) pairs are fetched by ordinary vector loads, one pair per lane, and
//    broadcast to the wave where used with DPP row_newbcast.  Columns of lower
//    order skip the leading sequences by per-lane masks.
//  * Waves are pipeline stages.  Stage s applies sequences q = sL .. sL+L-1 of
//    the group over a window of rows held in registers, sweeping down the
//    column L rows per tile along an anti-diagonal (the reference's L x L
//    register tile, src/rotations/rotations.h:26-45), and hands L finished
//    rows per tile to the next stage through an LDS ring.
//  * lo2hi is run as hi2lo on reversed rows: the transposed rotation of the
//    reversed pair is the forward rotation, so one engine covers both.
//  * A task is one workgroup's pass over the rows (W-2 stage waves plus an
//    importer and an exporter wave).  Consecutive tasks of a group hand rows
//    over through row-major boundary buffers in global memory, accessed with
//    device-scope relaxed atomics, and publish progress counters; every task
//    of a group runs on one XCD (one L2).  The kernel is persistent:
//    workgroups take tasks from per-XCD ticket counters in a host-sorted
//    order (producers before consumers, long chains first).  Every wait is
//    bounded; on expiry the kernel records FT_GPU_ETIMEOUT and every wave
//    still reaches its exit.
//
// Arithmetic: x' = fma(c, x, s*y), y' = fma(c, y, -(s*x)) -- the fused form of
// the reference's dispatched SIMD kernels (src/rotations/rotations_AVX512F.c:4-15).
// This file is compiled with -ffp-contract=off so no other contraction occurs.

#include <hip/hip_runtime.h>
#include <string.h>

#include "ft_pipeline.hpp"

namespace {

thread_local int g_hip_err = 0;

// Records the XCC_ID of every workgroup (to count the device's XCDs).
__global__ void xcd_probe(int *out) {
    if (threadIdx.x == 0) out[blockIdx.x] = (int)__builtin_amdgcn_s_getreg((20) | (0 << 6) | ((4 - 1) << 11));
}

int map_err(hipError_t e) {
    if (e == hipSuccess) return FTAMD_OK;
    g_hip_err = (int)e;
    return e == hipErrorOutOfMemory ? FTAMD_ENOMEM : FTAMD_EDEVICE;
}

}  // namespace

// ================================================================ C ABI ====

extern "C" {

int ftamd_hip_device(int *device, int *num_cu) {
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return FTAMD_EDEVICE;
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return map_err(e);
    if (device) *device = dev;
    if (num_cu) {
        int cu = 0;
        e = hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, dev);
        if (e != hipSuccess) return map_err(e);
        *num_cu = cu;
    }
    return FTAMD_OK;
}

// Number of XCDs, from the XCC_IDs a spread of workgroups reports (cached
// per process; one device).
// Architecture name of the current device (e.g. "gfx950"), cached per device.
int ftamd_hip_arch(char *name, int cap) {
    static char cache[FTAMD_MAX_DEVICES][64];
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return map_err(e);
    if (dev < 0 || dev >= FTAMD_MAX_DEVICES || cap < 1) return FTAMD_EINVAL;
    if (!cache[dev][0]) {
        hipDeviceProp_t prop;
        e = hipGetDeviceProperties(&prop, dev);
        if (e != hipSuccess) return map_err(e);
        strncpy(cache[dev], prop.gcnArchName, 63);
        cache[dev][63] = 0;
    }
    strncpy(name, cache[dev], (size_t)cap - 1);
    name[cap - 1] = 0;
    return FTAMD_OK;
}

int ftamd_hip_xcd_count(int *nxcd) {
    static int cached = 0;
    if (cached > 0) {
        *nxcd = cached;
        return FTAMD_OK;
    }
    const int blocks = 4096;
    int *d = nullptr;
    hipError_t e = hipMalloc(&d, blocks * sizeof(int));
    if (e != hipSuccess) return map_err(e);
    hipLaunchKernelGGL(xcd_probe, dim3(blocks), dim3(64), 0, 0, d);
    int *h = (int *)malloc(blocks * sizeof(int));
    e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpy(h, d, blocks * sizeof(int), hipMemcpyDeviceToHost);
    hipFree(d);
    if (e != hipSuccess || !h) {
        free(h);
        return map_err(e != hipSuccess ? e : hipErrorOutOfMemory);
    }
    int mx = 0;
    for (int i = 0; i < blocks; ++i) mx = h[i] > mx ? h[i] : mx;
    free(h);
    cached = mx + 1 > FTAMD_MAX_XCD ? FTAMD_MAX_XCD : mx + 1;
    *nxcd = cached;
    return FTAMD_OK;
}

int ftamd_hip_pointer_kind(const void *ptr, int *kind) {
    hipPointerAttribute_t attr;
    memset(&attr, 0, sizeof(attr));
    hipError_t e = hipPointerGetAttributes(&attr, ptr);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        *kind = FTAMD_PTR_HOST;
        return FTAMD_OK;
    }
    *kind = (attr.type == hipMemoryTypeDevice || attr.type == hipMemoryTypeManaged) ? FTAMD_PTR_DEVICE
                                                                                     : FTAMD_PTR_HOST;
    return FTAMD_OK;
}

int ftamd_hip_malloc(void **ptr, size_t bytes) {
    *ptr = nullptr;
    if (bytes == 0) bytes = 1;
    return map_err(hipMalloc(ptr, bytes));
}

int ftamd_hip_free(void *ptr) { return ptr ? map_err(hipFree(ptr)) : FTAMD_OK; }

int ftamd_hip_memset(void *ptr, int value, size_t bytes) { return map_err(hipMemset(ptr, value, bytes)); }

int ftamd_hip_h2d(void *dst, const void *src, size_t bytes, void *stream) {
    return map_err(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, (hipStream_t)stream));
}

int ftamd_hip_d2h(void *dst, const void *src, size_t bytes, void *stream) {
    return map_err(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, (hipStream_t)stream));
}

int ftamd_hip_d2d(void *dst, const void *src, size_t bytes, void *stream) {
    return map_err(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, (hipStream_t)stream));
}

int ftamd_hip_device_sync(void) { return map_err(hipDeviceSynchronize()); }

int ftamd_hip_sync(void *stream) { return map_err(hipStreamSynchronize((hipStream_t)stream)); }

int ftamd_hip_last_runtime_error(void) { return g_hip_err; }

// Launch the pipelined kernel described by *a on `stream`.
int ftamd_hip_launch(const FtamdLaunch *a, int precision, int kind, int dir, int L, int W, int grid, void *stream) {
    ftamd::KernelFn fn = kind == FTAMD_SPH ? ftamd::pick_sph(precision, dir, L, W)
                       : kind == FTAMD_TRI ? ftamd::pick_tri(precision, dir, L, W)
                       : kind == FTAMD_DISK ? ftamd::pick_disk(precision, dir, L, W) : nullptr;
    if (!fn) return FTAMD_EINVAL;
    if (a->ntasks <= 0) return FTAMD_OK;
    if (grid < 1) grid = 1;
    if (grid > a->nxcd * a->ntasks) grid = a->nxcd * a->ntasks;
    hipLaunchKernelGGL(fn, dim3(grid), dim3(W * ftamd::kLanes), 0, (hipStream_t)stream, *a);
    return map_err(hipGetLastError());
}

}  // extern "C"
