// ft_k_sys.hip -- instantiates the band-systolic engine (ft_systolic.hpp) and
// its C-ABI launcher.
#include "ft_systolic.hpp"

// S = 4 in FP64 does not fit the 128 VGPRs a 16-wave workgroup allows
// without spilling; those plans run on the pipelined engine.
extern "C" int ftamd_hip_sys_supported(int S, int precision) { return S == 2 || (S == 4 && precision == 1); }

extern "C" int ftamd_hip_sys_launch(const FtamdSysLaunch *a, int precision, int S, void *stream) {
    using Fn = void (*)(FtamdSysLaunch);
    Fn fn = nullptr;
    if (!ftamd_hip_sys_supported(S, precision)) return FTAMD_EINVAL;
    // sph and disk chains pair the cosine and sine columns of each order
    // (one mask per lane row); tri chains have one order per column
    const bool rm = a->rowmask != 0;
    if (precision == 0) fn = rm ? ftsys::sys_kernel<double, 2, true> : ftsys::sys_kernel<double, 2, false>;
    else if (S == 2) fn = rm ? ftsys::sys_kernel<float, 2, true> : ftsys::sys_kernel<float, 2, false>;
    else if (S == 4) fn = rm ? ftsys::sys_kernel<float, 4, true> : ftsys::sys_kernel<float, 4, false>;
    if (!fn) return FTAMD_EINVAL;
    if (a->nchains <= 0) return FTAMD_OK;
    // the kernel indexes xcd_slot[0 .. nxcd] and one ticket per XCD
    if (a->nxcd < 1 || a->nxcd > FTAMD_MAX_XCD || a->grid < 1 || a->nslots < 1 || !a->ctl || (a->nprog > 0 && !a->prog))
        return FTAMD_EINVAL;
    hipLaunchKernelGGL(fn, dim3(a->grid), dim3(ftsys::kWaves * 64), 0, (hipStream_t)stream, *a);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) return FTAMD_OK;
    return e == hipErrorOutOfMemory ? FTAMD_ENOMEM : FTAMD_EDEVICE;
}
