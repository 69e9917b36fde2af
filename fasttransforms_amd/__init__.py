"""Python host mirror of the rotation path of FastTransforms on MI355X.

Thin ctypes bindings over ``libfasttransforms_amd.so`` (the C ABI declared in
``include/fasttransforms_amd.h``).  The names follow the reference's C API
(``ft_plan_rotsphere`` -> ``plan_rotsphere``, ``ft_execute_sph_hi2lo`` ->
``execute_sph_hi2lo`` ...), with the same argument meaning and the same
column-major layout: an n x M matrix is a float64 numpy array of shape (n, M)
in Fortran order, or a torch CUDA tensor whose memory is column-major (for
example ``torch.empty(M, n).T``).

Every computation runs in the HIP kernels.  If the shared library is not
built, importing this package raises; there is no CPU fallback.
"""
import ctypes
import os

import numpy as np

# executor kinds (FT_GPU_SPH .. FT_GPU_RECTDISK); a plan's kind is its
# constructor's: SPH, TRI, DISK or RECTDISK (sphv runs on a SPH plan)
SPH, TRI, DISK, SPHV, RECTDISK = 0, 1, 2, 3, 4
HI2LO, LO2HI = 0, 1
FP64, FP32 = 0, 1
# error codes of the C ABI (FT_GPU_* in include/fasttransforms_amd.h)
(FT_GPU_OK, FT_GPU_EINVAL, FT_GPU_ENOMEM, FT_GPU_EDEVICE, FT_GPU_ETIMEOUT, FT_GPU_EPTR,
 FT_GPU_EINCOMPLETE) = range(7)
_KIND = {"sph": SPH, "tri": TRI, "disk": DISK, "sphv": SPHV, "rectdisk": RECTDISK}
# table geometry of each kind: sphv and rectdisk use the sphere's
_GEOMETRY = {SPH: SPH, TRI: TRI, DISK: DISK, SPHV: SPH, RECTDISK: SPH}
_DIR = {"hi2lo": HI2LO, "lo2hi": LO2HI}

# One HIP runtime per process: PyTorch-ROCm bundles its own libamdhip64
# (SONAME libamdhip64.so.7) and requests it under another file name, so if this
# library loaded the system runtime first, torch would load a second copy that
# finds no GPU.  Importing torch first lets this library bind to torch's copy.
try:
    import torch as _torch  # noqa: F401
except ImportError:  # the C ABI works without torch
    _torch = None

_here = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_here, "libfasttransforms_amd.so")
if not os.path.exists(LIB_PATH):
    raise ImportError("fasttransforms_amd: %s is not built (run python fasttransforms_amd/build.py)" % LIB_PATH)
_lib = ctypes.CDLL(LIB_PATH)

_vp = ctypes.c_void_p
_i = ctypes.c_int
_d = ctypes.c_double

# Exported symbols of include/fasttransforms_amd.h, with their ctypes signatures.
SIGNATURES = {
    "ft_set_num_threads": (None, [_i]),
    "ft_destroy_rotation_plan": (None, [_vp]),
    "ft_plan_rotsphere": (_vp, [_i]),
    "ft_plan_rottriangle": (_vp, [_i, _d, _d, _d]),
    "ft_plan_rotdisk": (_vp, [_i, _d, _d]),
    "ft_plan_rotrectdisk": (_vp, [_i, _d]),
    "ft_plan_rotannulus": (_vp, [_i, _d, _d, _d, _d]),
    "ft_gpu_last_error": (_i, []),
    "ft_gpu_error_string": (ctypes.c_char_p, [_i]),
    "ft_gpu_execute": (_i, [_vp, _i, _i, _i, _vp, _i, _vp]),
    "ft_gpu_synchronize": (_i, [_vp]),
    "ft_gpu_execute_noalloc": (_i, [_vp, _i, _i, _i, _vp, _i, _vp]),
    "ft_gpu_execute_batch": (_i, [_vp, _i, _i, _i, _vp, _vp, _i, _vp]),
    "ft_gpu_plan_batch": (_vp, [_vp, _i, _i, _i, _vp, _vp, _i]),
    "ft_gpu_execute_batch_plan": (_i, [_vp, _vp]),
    "ft_gpu_destroy_batch": (None, [_vp]),
    "ft_gpu_set_tuning": (_i, [_i, _i, _i]),
    "ft_gpu_set_engine": (_i, [_i]),
    "ft_gpu_set_systolic": (_i, [_i]),
    "ft_gpu_last_engine": (_i, []),
    "ft_gpu_set_spin_limit": (_i, [_i]),
    "ft_gpu_set_tuning_table": (_i, [_vp, _i]),
    "ft_gpu_tuning_table": (_i, [_vp, _i]),
    "ft_gpu_tuned_choice": (_i, [_i, _i, _i, _i, _i, _vp]),
    "ft_plan_sph2fourier": (_vp, [_i]),
    "ft_plan_tri2cheb": (_vp, [_i, _d, _d, _d]),
    "ft_plan_disk2cxf": (_vp, [_i, _d, _d]),
    "ft_plan_rectdisk2cheb": (_vp, [_i, _d]),
    "ft_plan_ann2cxf": (_vp, [_i, _d, _d, _d, _d]),
    "ft_destroy_harmonic_plan": (None, [_vp]),
    "ft_gpu_execute_harmonic": (_i, [_i, ctypes.c_char, _vp, _vp, _i, _i, _vp]),
}
for _t in ("sph2fourier", "fourier2sph", "sphv2fourier", "fourier2sphv", "tri2cheb", "cheb2tri", "disk2cxf",
           "cxf2disk", "rectdisk2cheb", "cheb2rectdisk", "ann2cxf", "cxf2ann"):
    SIGNATURES["ft_execute_" + _t] = (None, [ctypes.c_char, _vp, _vp, _i, _i])
for _dn in ("hi2lo", "lo2hi"):
    for _k in ("sph", "tri", "disk", "rectdisk"):
        SIGNATURES["ft_kernel_%s_%s" % (_k, _dn)] = (None, [_vp, _i, _i, _vp, _i])
    for _k in ("sph", "tri", "disk", "sphv", "rectdisk"):
        SIGNATURES["ft_execute_%s_%s" % (_k, _dn)] = (None, [_vp, _vp, _vp, _i])

for _dn in ("hi2lo", "lo2hi"):
    SIGNATURES["ft_execute_tet_" + _dn] = (None, [_vp, _vp, _vp, _i, _i])
    SIGNATURES["ft_kernel_tet_" + _dn] = (None, [_vp, _i, _i, _vp])
SIGNATURES["ft_gpu_execute_tet"] = (_i, [_vp, _vp, _i, _vp, _i, _i, _vp])
SIGNATURES["ft_plan_rotspinsphere"] = (_vp, [_i, _i])
SIGNATURES["ft_destroy_spin_rotation_plan"] = (None, [_vp])
SIGNATURES["ft_gpu_execute_spinsph"] = (_i, [_vp, _i, _vp, _i, _vp])
for _dn in ("hi2lo", "lo2hi"):
    SIGNATURES["ft_execute_spinsph_" + _dn] = (None, [_vp, _vp, _vp, _i])
    SIGNATURES["ft_kernel_spinsph_" + _dn] = (None, [_vp, _i, _vp, _i])

for _t in ("sph_synthesis", "sph_analysis", "sphv_synthesis", "sphv_analysis"):
    SIGNATURES["ft_plan_" + _t] = (_vp, [_i, _i, ctypes.c_uint])
    SIGNATURES["ft_execute_" + _t] = (None, [ctypes.c_char, _vp, _vp, _i, _i])
SIGNATURES["ft_destroy_sphere_fftw_plan"] = (None, [_vp])
for _t in ("disk_synthesis", "disk_analysis"):
    SIGNATURES["ft_plan_" + _t] = (_vp, [_i, _i, ctypes.c_uint])
    SIGNATURES["ft_execute_" + _t] = (None, [ctypes.c_char, _vp, _vp, _i, _i])
SIGNATURES["ft_destroy_disk_fftw_plan"] = (None, [_vp])
for _t in ("annulus_synthesis", "annulus_analysis"):
    SIGNATURES["ft_plan_" + _t] = (_vp, [_i, _i, _d, ctypes.c_uint])
    SIGNATURES["ft_execute_" + _t] = (None, [ctypes.c_char, _vp, _vp, _i, _i])
SIGNATURES["ft_destroy_annulus_fftw_plan"] = (None, [_vp])
SIGNATURES["ft_get_rho_annulus_fftw_plan"] = (_d, [_vp])
for _t in ("tri_synthesis", "tri_analysis"):
    SIGNATURES["ft_plan_" + _t] = (_vp, [_i, _i, ctypes.c_uint])
    SIGNATURES["ft_execute_" + _t] = (None, [ctypes.c_char, _vp, _vp, _i, _i])
SIGNATURES["ft_destroy_triangle_fftw_plan"] = (None, [_vp])
for _t in ("spinsph_synthesis", "spinsph_analysis"):
    SIGNATURES["ft_plan_" + _t] = (_vp, [_i, _i, _i, ctypes.c_uint])
    SIGNATURES["ft_execute_" + _t] = (None, [ctypes.c_char, _vp, _vp, _i, _i])
SIGNATURES["ft_destroy_spinsphere_fftw_plan"] = (None, [_vp])
SIGNATURES["ft_get_spin_spinsphere_fftw_plan"] = (_i, [_vp])
SIGNATURES["ft_gpu_execute_spinsph_grid"] = (_i, [_i, ctypes.c_char, _vp, _vp, _i, _i, _vp])
for _t in ("rectdisk_synthesis", "rectdisk_analysis"):
    SIGNATURES["ft_plan_" + _t] = (_vp, [_i, _i, ctypes.c_uint])
    SIGNATURES["ft_execute_" + _t] = (None, [ctypes.c_char, _vp, _vp, _i, _i])
SIGNATURES["ft_destroy_rectdisk_fftw_plan"] = (None, [_vp])
SIGNATURES["ft_gpu_execute_rectdisk_grid"] = (_i, [_i, ctypes.c_char, _vp, _vp, _i, _i, _vp])
for _t in ("tet_synthesis", "tet_analysis"):
    SIGNATURES["ft_plan_" + _t] = (_vp, [_i, _i, _i, ctypes.c_uint])
    SIGNATURES["ft_execute_" + _t] = (None, [ctypes.c_char, _vp, _vp, _i, _i, _i])
SIGNATURES["ft_destroy_tetrahedron_fftw_plan"] = (None, [_vp])
SIGNATURES["ft_gpu_execute_tet_grid"] = (_i, [_i, ctypes.c_char, _vp, _vp, _i, _i, _i, _vp])
SIGNATURES["ft_gpu_execute_tri_grid"] = (_i, [_i, ctypes.c_char, _vp, _vp, _i, _i, _vp])
SIGNATURES["ft_gpu_execute_sph_grid"] = (_i, [_i, ctypes.c_char, _vp, _vp, _i, _i, _vp])

for _name, (_res, _args) in SIGNATURES.items():
    _fn = getattr(_lib, _name)
    _fn.restype = _res
    _fn.argtypes = _args


class _PlanHead(ctypes.Structure):
    # The reference's leading members {double *s; double *c; int n;}.
    _fields_ = [("s", ctypes.POINTER(ctypes.c_double)), ("c", ctypes.POINTER(ctypes.c_double)), ("n", ctypes.c_int)]


class GPUError(RuntimeError):
    pass


def last_error():
    return _lib.ft_gpu_last_error()


def _check():
    e = _lib.ft_gpu_last_error()
    if e:
        raise GPUError("fasttransforms_amd: %s (code %d)" % (_lib.ft_gpu_error_string(e).decode(), e))


class RotationPlan:
    """Owner of an ``ft_rotation_plan *`` (destroyed with the object)."""

    def __init__(self, ptr, kind, n, params):
        if not ptr:
            _check()
            raise GPUError("plan construction failed")
        self._ptr = ptr
        # kept on the object: module globals may be gone when it is collected at exit
        self._destroy = _lib.ft_destroy_rotation_plan
        self.kind = kind
        self.family = {SPH: "sph", TRI: "tri", DISK: "disk", RECTDISK: "rectdisk"}.get(kind)
        self.n = n
        self.params = params

    @property
    def ptr(self):
        return self._ptr

    def _head(self):
        return ctypes.cast(self._ptr, ctypes.POINTER(_PlanHead)).contents

    @property
    def size(self):
        n = self.n
        return n * n if self.kind == DISK else n * (n + 1) // 2

    @property
    def s(self):
        """Host FP64 sine table (a copy)."""
        return np.ctypeslib.as_array(self._head().s, shape=(self.size,)).copy()

    @property
    def c(self):
        """Host FP64 cosine table (a copy)."""
        return np.ctypeslib.as_array(self._head().c, shape=(self.size,)).copy()

    def __del__(self):
        ptr, self._ptr = getattr(self, "_ptr", None), None
        if ptr:
            self._destroy(ptr)


class _SpinHead(ctypes.Structure):
    # The reference's leading members {double *s1, *c1, *s2, *c2; int n, s;}.
    _dp = ctypes.POINTER(ctypes.c_double)
    _fields_ = [("s1", _dp), ("c1", _dp), ("s2", _dp), ("c2", _dp), ("n", ctypes.c_int), ("s", ctypes.c_int)]


class SpinRotationPlan:
    """Owner of an ``ft_spin_rotation_plan *`` (destroyed with the object)."""

    def __init__(self, ptr, n, s):
        if not ptr:
            _check()
            raise GPUError("plan construction failed")
        self._ptr = ptr
        self._destroy = _lib.ft_destroy_spin_rotation_plan
        self.n, self.s = n, s

    @property
    def ptr(self):
        return self._ptr

    def _table(self, name):
        head = ctypes.cast(self._ptr, ctypes.POINTER(_SpinHead)).contents
        size = self.n * self.n * (1 if name in ("s1", "c1") else abs(self.s))
        if size == 0:
            return np.zeros(0)
        return np.ctypeslib.as_array(getattr(head, name), shape=(size,)).copy()

    s1 = property(lambda self: self._table("s1"))
    c1 = property(lambda self: self._table("c1"))
    s2 = property(lambda self: self._table("s2"))
    c2 = property(lambda self: self._table("c2"))

    def __del__(self):
        ptr, self._ptr = getattr(self, "_ptr", None), None
        if ptr:
            self._destroy(ptr)


def plan_rotspinsphere(n, s):
    """ft_plan_rotspinsphere: spin-weighted spherical harmonics of spin s."""
    return SpinRotationPlan(_lib.ft_plan_rotspinsphere(int(n), int(s)), int(n), int(s))


def _complex_pointer(A, n, M):
    """Address of an n x M complex column-major buffer: a Fortran-ordered
    complex128 numpy array, or a contiguous torch complex128 tensor of shape
    (M, n) (row j = column j) or of n*M entries."""
    if _is_torch(A):
        import torch
        if A.dtype != torch.complex128 or not A.is_contiguous() or A.numel() < n * M:
            raise ValueError("expected a contiguous complex128 tensor of n*M entries")
        if A.dim() == 2 and A.shape[1] != n:
            raise ValueError("expected a (M, n) tensor")
        return _vp(A.data_ptr())
    if not isinstance(A, np.ndarray) or A.dtype != np.complex128:
        raise TypeError("expected a complex128 numpy array")
    if A.ndim == 2 and (A.shape[0] != n or not A.flags.f_contiguous):
        raise ValueError("expected a Fortran-ordered (n, M) array")
    if A.ndim != 2 and not A.flags.c_contiguous:
        raise ValueError("expected a contiguous buffer")
    if A.size < n * M:
        raise ValueError("array has fewer than n*M entries")
    return A.ctypes.data_as(_vp)


def _execute_spin(direction):
    fn = getattr(_lib, "ft_execute_spinsph_" + direction)

    def run(plan, A, M=None):
        M = int(A.shape[1]) if M is None and A.ndim == 2 else (A.size // plan.n if M is None else int(M))
        fn(plan.ptr, _complex_pointer(A, plan.n, M), None, M)
        _check()
        return A

    run.__name__ = "execute_spinsph_" + direction
    return run


def _kernel_spin(direction):
    fn = getattr(_lib, "ft_kernel_spinsph_" + direction)

    def run(plan, m, A, S=1):
        span = (plan.n - 1) * int(S) + 1
        if A.size < span:
            raise ValueError("buffer shorter than (n-1)*S+1")
        fn(plan.ptr, int(m), _complex_pointer(A.reshape(-1), span, 1), int(S))
        _check()
        return A

    run.__name__ = "kernel_spinsph_" + direction
    return run


execute_spinsph_hi2lo = _execute_spin("hi2lo")
execute_spinsph_lo2hi = _execute_spin("lo2hi")
kernel_spinsph_hi2lo = _kernel_spin("hi2lo")
kernel_spinsph_lo2hi = _kernel_spin("lo2hi")


def gpu_execute_spinsph(plan, direction, A, M=None, stream=None):
    """Asynchronous ft_gpu_execute_spinsph on a contiguous torch complex128
    tensor (shape (M, n): row j is column j)."""
    import torch
    direction = _DIR.get(direction, direction)
    M = (A.shape[0] if A.dim() == 2 else A.numel() // plan.n) if M is None else int(M)
    if stream is None:
        stream = torch.cuda.current_stream(A.device)
    rc = _lib.ft_gpu_execute_spinsph(plan.ptr, int(direction), _complex_pointer(A, plan.n, M), M,
                                     _vp(stream.cuda_stream))
    if rc:
        raise GPUError("ft_gpu_execute_spinsph: %s" % _lib.ft_gpu_error_string(rc).decode())
    return A


class SphereGridPlan:
    """Owner of an ``ft_sphere_fftw_plan *`` (the grid transforms' device plan)."""

    def __init__(self, ptr, N, M, synthesis, sphv):
        if not ptr:
            _check()
            raise GPUError("plan construction failed")
        self._ptr = ptr
        self._destroy = _lib.ft_destroy_sphere_fftw_plan
        self.N, self.M, self.synthesis, self.sphv = N, M, synthesis, sphv

    @property
    def ptr(self):
        return self._ptr

    def __del__(self):
        ptr, self._ptr = getattr(self, "_ptr", None), None
        if ptr:
            self._destroy(ptr)


def _plan_grid(name, synthesis, sphv):
    fn = getattr(_lib, "ft_plan_" + name)

    def make(N, M, flags=0):
        return SphereGridPlan(fn(int(N), int(M), int(flags)), int(N), int(M), synthesis, sphv)

    make.__name__ = "plan_" + name
    return make


plan_sph_synthesis = _plan_grid("sph_synthesis", True, False)
plan_sph_analysis = _plan_grid("sph_analysis", False, False)
plan_sphv_synthesis = _plan_grid("sphv_synthesis", True, True)
plan_sphv_analysis = _plan_grid("sphv_analysis", False, True)
plan_disk_synthesis = _plan_grid("disk_synthesis", True, False)
plan_disk_analysis = _plan_grid("disk_analysis", False, False)
_GRID_WHICH = {"sph_synthesis": 0, "sph_analysis": 1, "sphv_synthesis": 2, "sphv_analysis": 3,
               "disk_synthesis": 4, "disk_analysis": 5, "annulus_synthesis": 6, "annulus_analysis": 7}


def plan_annulus_synthesis(N, M, rho, flags=0):
    return SphereGridPlan(_lib.ft_plan_annulus_synthesis(int(N), int(M), float(rho), int(flags)), int(N), int(M),
                          True, False)


def plan_annulus_analysis(N, M, rho, flags=0):
    return SphereGridPlan(_lib.ft_plan_annulus_analysis(int(N), int(M), float(rho), int(flags)), int(N), int(M),
                          False, False)


def _execute_grid(name):
    fn = getattr(_lib, "ft_execute_" + name)

    def run(trans, plan, X):
        N, M = plan.N, plan.M
        fn(trans.encode() if isinstance(trans, str) else trans, plan.ptr, _pointer(X, N, M), N, M)
        _check()
        return X

    run.__name__ = "execute_" + name
    return run


execute_sph_synthesis = _execute_grid("sph_synthesis")
execute_sph_analysis = _execute_grid("sph_analysis")
execute_sphv_synthesis = _execute_grid("sphv_synthesis")
execute_sphv_analysis = _execute_grid("sphv_analysis")
execute_disk_synthesis = _execute_grid("disk_synthesis")
execute_disk_analysis = _execute_grid("disk_analysis")
execute_annulus_synthesis = _execute_grid("annulus_synthesis")
execute_annulus_analysis = _execute_grid("annulus_analysis")


def gpu_execute_sph_grid(name, trans, plan, A, stream=None):
    """Asynchronous ft_gpu_execute_sph_grid on a torch float64 tensor laid out
    column-major (shape (M, N): row j is column j)."""
    import torch
    if stream is None:
        stream = torch.cuda.current_stream(A.device)
    rc = _lib.ft_gpu_execute_sph_grid(_GRID_WHICH[name], trans.encode(), plan.ptr, _pointer(A, plan.N, plan.M),
                                      plan.N, plan.M, _vp(stream.cuda_stream))
    if rc:
        raise GPUError("ft_gpu_execute_sph_grid: %s" % _lib.ft_gpu_error_string(rc).decode())
    return A


class TriangleGridPlan(SphereGridPlan):
    """Owner of an ``ft_triangle_fftw_plan *``."""

    def __init__(self, ptr, N, M, synthesis):
        SphereGridPlan.__init__(self, ptr, N, M, synthesis, False)
        self._destroy = _lib.ft_destroy_triangle_fftw_plan


def plan_tri_synthesis(N, M, flags=0):
    return TriangleGridPlan(_lib.ft_plan_tri_synthesis(int(N), int(M), int(flags)), int(N), int(M), True)


def plan_tri_analysis(N, M, flags=0):
    return TriangleGridPlan(_lib.ft_plan_tri_analysis(int(N), int(M), int(flags)), int(N), int(M), False)


class RectdiskGridPlan(SphereGridPlan):
    """Owner of an ``ft_rectdisk_fftw_plan *``."""

    def __init__(self, ptr, N, M, synthesis):
        SphereGridPlan.__init__(self, ptr, N, M, synthesis, False)
        self._destroy = _lib.ft_destroy_rectdisk_fftw_plan


def plan_rectdisk_synthesis(N, M, flags=0):
    return RectdiskGridPlan(_lib.ft_plan_rectdisk_synthesis(int(N), int(M), int(flags)), int(N), int(M), True)


def plan_rectdisk_analysis(N, M, flags=0):
    return RectdiskGridPlan(_lib.ft_plan_rectdisk_analysis(int(N), int(M), int(flags)), int(N), int(M), False)


class SpinSphereGridPlan(SphereGridPlan):
    """Owner of an ``ft_spinsphere_fftw_plan *``."""

    def __init__(self, ptr, N, M, S, synthesis):
        SphereGridPlan.__init__(self, ptr, N, M, synthesis, S % 2 != 0)
        self._destroy = _lib.ft_destroy_spinsphere_fftw_plan
        self.S = S


def plan_spinsph_synthesis(N, M, S, flags=0):
    return SpinSphereGridPlan(_lib.ft_plan_spinsph_synthesis(int(N), int(M), int(S), int(flags)), int(N), int(M),
                              int(S), True)


def plan_spinsph_analysis(N, M, S, flags=0):
    return SpinSphereGridPlan(_lib.ft_plan_spinsph_analysis(int(N), int(M), int(S), int(flags)), int(N), int(M),
                              int(S), False)


def _execute_spin_grid(name):
    fn = getattr(_lib, "ft_execute_" + name)

    def run(trans, plan, X):
        N, M = plan.N, plan.M
        fn(trans.encode(), plan.ptr, _complex_pointer(X, N, M), N, M)
        _check()
        return X

    run.__name__ = "execute_" + name
    return run


execute_spinsph_synthesis = _execute_spin_grid("spinsph_synthesis")
execute_spinsph_analysis = _execute_spin_grid("spinsph_analysis")
execute_rectdisk_synthesis = _execute_grid("rectdisk_synthesis")
execute_rectdisk_analysis = _execute_grid("rectdisk_analysis")
execute_tri_synthesis = _execute_grid("tri_synthesis")
execute_tri_analysis = _execute_grid("tri_analysis")


def gpu_execute_tri_grid(name, trans, plan, A, stream=None):
    """Asynchronous ft_gpu_execute_tri_grid ("tri_synthesis" or "tri_analysis")
    on a torch float64 tensor of N*M entries."""
    import torch
    if stream is None:
        stream = torch.cuda.current_stream(A.device)
    rc = _lib.ft_gpu_execute_tri_grid(1 if name == "tri_analysis" else 0, trans.encode(), plan.ptr,
                                      _pointer(A, plan.N, plan.M), plan.N, plan.M, _vp(stream.cuda_stream))
    if rc:
        raise GPUError("ft_gpu_execute_tri_grid: %s" % _lib.ft_gpu_error_string(rc).decode())
    return A


class TetrahedronGridPlan:
    """Owner of an ``ft_tetrahedron_fftw_plan *``."""

    def __init__(self, ptr, N, L, M, synthesis):
        if not ptr:
            _check()
            raise GPUError("plan construction failed")
        self._ptr = ptr
        self._destroy = _lib.ft_destroy_tetrahedron_fftw_plan
        self.N, self.L, self.M, self.synthesis = N, L, M, synthesis

    @property
    def ptr(self):
        return self._ptr

    def __del__(self):
        ptr, self._ptr = getattr(self, "_ptr", None), None
        if ptr:
            self._destroy(ptr)


def plan_tet_synthesis(N, L, M, flags=0):
    return TetrahedronGridPlan(_lib.ft_plan_tet_synthesis(int(N), int(L), int(M), int(flags)), int(N), int(L), int(M),
                               True)


def plan_tet_analysis(N, L, M, flags=0):
    return TetrahedronGridPlan(_lib.ft_plan_tet_analysis(int(N), int(L), int(M), int(flags)), int(N), int(L), int(M),
                               False)


def _execute_tet_grid(name):
    fn = getattr(_lib, "ft_execute_" + name)

    def run(trans, plan, X):
        N, L, M = plan.N, plan.L, plan.M
        fn(trans.encode(), plan.ptr, _tet_pointer_raw(X, N * L * M), N, L, M)
        _check()
        return X

    run.__name__ = "execute_" + name
    return run


def _tet_pointer_raw(X, size):
    if not isinstance(X, np.ndarray) or X.dtype != np.float64 or not X.flags.f_contiguous or X.size < size:
        raise ValueError("expected a Fortran-ordered float64 array of N*L*M entries")
    return X.ctypes.data_as(_vp)


execute_tet_synthesis = _execute_tet_grid("tet_synthesis")
execute_tet_analysis = _execute_tet_grid("tet_analysis")


def plan_rotsphere(n):
    return RotationPlan(_lib.ft_plan_rotsphere(int(n)), SPH, int(n), ())


def plan_rottriangle(n, alpha, beta, gamma):
    return RotationPlan(_lib.ft_plan_rottriangle(int(n), float(alpha), float(beta), float(gamma)), TRI, int(n),
                        (float(alpha), float(beta), float(gamma)))


def plan_rotrectdisk(n, beta):
    return RotationPlan(_lib.ft_plan_rotrectdisk(int(n), float(beta)), RECTDISK, int(n), (float(beta),))


def plan_rotannulus(n, alpha, beta, gamma, rho):
    """ft_plan_rotannulus: the annulus of inner radius rho (gamma = 0); a disk-layout plan."""
    P = RotationPlan(_lib.ft_plan_rotannulus(int(n), float(alpha), float(beta), float(gamma), float(rho)), DISK,
                     int(n), (float(alpha), float(beta), float(gamma), float(rho)))
    P.family = "annulus"
    return P


def plan_rotdisk(n, alpha, beta):
    return RotationPlan(_lib.ft_plan_rotdisk(int(n), float(alpha), float(beta)), DISK, int(n),
                        (float(alpha), float(beta)))


def set_num_threads(n):
    _lib.ft_set_num_threads(int(n))


def set_tuning(L=0, W=0, wg_per_cu=0):
    rc = _lib.ft_gpu_set_tuning(int(L), int(W), int(wg_per_cu))
    if rc:
        raise ValueError("unsupported tuning (L=%r, W=%r, wg_per_cu=%r)" % (L, W, wg_per_cu))


ENGINE_AUTO, ENGINE_PIPELINE, ENGINE_SYSTOLIC = 0, 1, 2


def set_engine(engine=ENGINE_AUTO):
    """Engine of the whole-matrix executors: ENGINE_AUTO, ENGINE_PIPELINE or ENGINE_SYSTOLIC."""
    rc = _lib.ft_gpu_set_engine(int(engine))
    if rc:
        raise ValueError("unknown engine %r" % (engine,))


def set_systolic(maxpart=0):
    """Most bands of a band-systolic chain part (1..16; longer chains run in parts); 0 = built-in choice."""
    if _lib.ft_gpu_set_systolic(int(maxpart)):
        raise ValueError("maxpart must be in 0..16")


TUNE_COLS = 11
TUNE_FIELDS = ("kind", "precision", "direction", "batch", "nlo", "nhi", "engine", "maxpart", "L", "W", "wg_per_cu")


def set_tuning_table(rows):
    """Installs an autotuning table: a list of TUNE_COLS-int rows (TUNE_FIELDS),
    or "compiled" for the table built into the library, or None for none."""
    if rows is None or rows == "compiled":
        rc = _lib.ft_gpu_set_tuning_table(None, -1 if rows is None else 0)
    else:
        flat = [int(v) for r in rows for v in r]
        if any(len(r) != TUNE_COLS for r in rows):
            raise ValueError("tuning rows have %d fields" % TUNE_COLS)
        buf = (ctypes.c_int * max(1, len(flat)))(*flat)
        rc = _lib.ft_gpu_set_tuning_table(ctypes.cast(buf, _vp), len(rows)) if rows else \
            _lib.ft_gpu_set_tuning_table(None, -1)
    if rc:
        raise ValueError("invalid tuning table")


def tuning_table():
    """The active autotuning table as a list of TUNE_COLS-int rows."""
    n = int(_lib.ft_gpu_tuning_table(None, 0))
    buf = (ctypes.c_int * max(1, n * TUNE_COLS))()
    n = int(_lib.ft_gpu_tuning_table(ctypes.cast(buf, _vp), n))
    return [list(buf[k * TUNE_COLS:(k + 1) * TUNE_COLS]) for k in range(n)]


def tuned_choice(kind, precision, direction, batch, n):
    """(engine, maxpart, L, W, wg_per_cu) a launch of that shape uses now."""
    kind = _KIND.get(kind, kind)
    direction = {"hi2lo": 0, "lo2hi": 1}.get(direction, direction)
    out = (ctypes.c_int * 5)()
    if _lib.ft_gpu_tuned_choice(int(kind), int(precision), int(direction), int(batch), int(n), ctypes.cast(out, _vp)):
        raise ValueError("invalid shape")
    return tuple(out)


def set_spin_limit(limit=0):
    """Polls before a band-systolic wait gives up (tests of the failure path); 0 = default."""
    if _lib.ft_gpu_set_spin_limit(int(limit)):
        raise ValueError("spin limit must be >= 0")


class _HarmonicHead(ctypes.Structure):
    # the reference's ft_harmonic_plan members (fasttransforms.h:371-385), then the device state
    _fields_ = [("RP", ctypes.POINTER(ctypes.c_void_p)), ("MP", ctypes.c_void_p),
                ("B", ctypes.POINTER(ctypes.c_double)), ("P", ctypes.POINTER(ctypes.POINTER(ctypes.c_double))),
                ("Pinv", ctypes.POINTER(ctypes.POINTER(ctypes.c_double))), ("alpha", ctypes.c_double),
                ("beta", ctypes.c_double), ("gamma", ctypes.c_double), ("delta", ctypes.c_double),
                ("rho", ctypes.c_double), ("NRP", ctypes.c_int), ("NMP", ctypes.c_int), ("NP", ctypes.c_int),
                ("dev", ctypes.c_void_p)]


(SPH2FOURIER, FOURIER2SPH, SPHV2FOURIER, FOURIER2SPHV, DISK2CXF, CXF2DISK, TRI2CHEB, CHEB2TRI, RECTDISK2CHEB,
 CHEB2RECTDISK, ANN2CXF, CXF2ANN) = range(12)
_HARMONIC = {"sph2fourier": SPH2FOURIER, "fourier2sph": FOURIER2SPH, "sphv2fourier": SPHV2FOURIER,
             "fourier2sphv": FOURIER2SPHV, "disk2cxf": DISK2CXF, "cxf2disk": CXF2DISK, "tri2cheb": TRI2CHEB,
             "cheb2tri": CHEB2TRI, "rectdisk2cheb": RECTDISK2CHEB, "cheb2rectdisk": CHEB2RECTDISK,
             "ann2cxf": ANN2CXF, "cxf2ann": CXF2ANN}


class HarmonicPlan:
    """Owner of an ``ft_harmonic_plan *`` (ft_plan_sph2fourier)."""

    def __init__(self, ptr, n):
        if not ptr:
            _check()
            raise GPUError("harmonic plan construction failed")
        self._ptr = ptr
        self._destroy = _lib.ft_destroy_harmonic_plan
        self.n = n

    @property
    def ptr(self):
        return self._ptr

    def _head(self):
        return ctypes.cast(self._ptr, ctypes.POINTER(_HarmonicHead)).contents

    def P(self, k):
        """Host connection matrix P[k] (a copy, n x n)."""
        return np.ctypeslib.as_array(self._head().P[k], shape=(self.n * self.n,)).copy().reshape(self.n, self.n, order="F")

    def Pinv(self, k):
        return np.ctypeslib.as_array(self._head().Pinv[k], shape=(self.n * self.n,)).copy().reshape(self.n, self.n,
                                                                                                    order="F")

    def R(self):
        """The annulus plan's bidiagonal factor MP[1] as a dense upper
        bidiagonal n x n matrix (a copy)."""
        head = self._head()
        if head.NMP != 2:
            raise ValueError("not an annulus plan")
        mp = ctypes.cast(head.MP, ctypes.POINTER(ctypes.POINTER(_ModifiedHead)))[1].contents
        v = np.ctypeslib.as_array(mp.R, shape=(2 * self.n,)).copy()
        return np.diag(v[:self.n]) + np.diag(v[self.n:2 * self.n - 1], 1)

    def __del__(self):
        ptr, self._ptr = getattr(self, "_ptr", None), None
        if ptr:
            self._destroy(ptr)


class _ModifiedHead(ctypes.Structure):
    # the library's modified plan: n, nu, nv, bandwidth, {diagonal[n], superdiagonal[n-1]}
    _fields_ = [("n", ctypes.c_int), ("nu", ctypes.c_int), ("nv", ctypes.c_int), ("b", ctypes.c_int),
                ("R", ctypes.POINTER(ctypes.c_double))]


def plan_sph2fourier(n):
    return HarmonicPlan(_lib.ft_plan_sph2fourier(int(n)), int(n))


def plan_tri2cheb(n, alpha, beta, gamma):
    return HarmonicPlan(_lib.ft_plan_tri2cheb(int(n), float(alpha), float(beta), float(gamma)), int(n))


def plan_disk2cxf(n, alpha, beta):
    return HarmonicPlan(_lib.ft_plan_disk2cxf(int(n), float(alpha), float(beta)), int(n))


def plan_rectdisk2cheb(n, beta):
    return HarmonicPlan(_lib.ft_plan_rectdisk2cheb(int(n), float(beta)), int(n))


def plan_ann2cxf(n, alpha, beta, gamma, rho):
    """ft_plan_ann2cxf: annulus polynomials of inner radius rho (gamma = 0)."""
    return HarmonicPlan(_lib.ft_plan_ann2cxf(int(n), float(alpha), float(beta), float(gamma), float(rho)), int(n))


def _harmonic(name):
    fn = getattr(_lib, "ft_execute_" + name)

    def run(trans, plan, A, N=None, M=None):
        N = plan.n if N is None else int(N)
        M = _M_of(A, N, M)
        fn(trans.encode() if isinstance(trans, str) else trans, plan.ptr, _pointer(A, N, M), int(N), M)
        _check()
        return A

    run.__name__ = "execute_" + name
    run.__doc__ = "ft_execute_%s(TRANS, plan, A, N, M): in place on the column-major N x M matrix A." % name
    return run


execute_sph2fourier = _harmonic("sph2fourier")
execute_fourier2sph = _harmonic("fourier2sph")
execute_sphv2fourier = _harmonic("sphv2fourier")
execute_fourier2sphv = _harmonic("fourier2sphv")
execute_tri2cheb = _harmonic("tri2cheb")
execute_cheb2tri = _harmonic("cheb2tri")
execute_disk2cxf = _harmonic("disk2cxf")
execute_cxf2disk = _harmonic("cxf2disk")
execute_rectdisk2cheb = _harmonic("rectdisk2cheb")
execute_cheb2rectdisk = _harmonic("cheb2rectdisk")
execute_ann2cxf = _harmonic("ann2cxf")
execute_cxf2ann = _harmonic("cxf2ann")


def gpu_execute_harmonic(transform, trans, plan, A, M=None, stream=None):
    """ft_gpu_execute_harmonic on a device tensor (asynchronous; errors through synchronize())."""
    import torch
    transform = _HARMONIC.get(transform, transform)
    M = _M_of(A, plan.n, M)
    if stream is None:
        stream = torch.cuda.current_stream(A.device)
    rc = _lib.ft_gpu_execute_harmonic(int(transform), trans.encode() if isinstance(trans, str) else trans, plan.ptr,
                                      _pointer(A, plan.n, M), plan.n, M, _vp(stream.cuda_stream))
    if rc:
        raise GPUError("ft_gpu_execute_harmonic: %s" % _lib.ft_gpu_error_string(rc).decode())
    return A


def last_engine():
    """Engine of this thread's last whole-matrix launch (ENGINE_PIPELINE or ENGINE_SYSTOLIC; 0 before any)."""
    return int(_lib.ft_gpu_last_engine())


def _is_torch(A):
    return type(A).__module__.startswith("torch")


def _pointer(A, n, M, dtype=np.float64):
    """Address of a column-major n x M buffer (numpy Fortran array or torch tensor)."""
    if _is_torch(A):
        import torch
        want = torch.float64 if dtype == np.float64 else torch.float32
        if A.dtype != want:
            raise TypeError("expected a %s tensor" % want)
        if A.numel() < n * M:
            raise ValueError("tensor has fewer than n*M elements")
        if A.dim() == 2:
            ok = (tuple(A.shape) == (M, n) and A.is_contiguous()) or \
                 (tuple(A.shape) == (n, M) and A.stride() == (1, n))
            if not ok:
                raise ValueError("expected a contiguous (M, n) tensor or a column-major (n, M) one")
        elif not A.is_contiguous():
            raise ValueError("tensor must be contiguous")
        return _vp(A.data_ptr())
    if not isinstance(A, np.ndarray) or A.dtype != dtype:
        raise TypeError("expected a %s numpy array" % np.dtype(dtype).name)
    if A.ndim == 2:
        if A.shape[0] != n or not A.flags.f_contiguous:
            raise ValueError("expected a Fortran-ordered (n, M) array")
    elif not A.flags.c_contiguous:
        raise ValueError("expected a contiguous buffer")
    if A.size < n * M:
        raise ValueError("array has fewer than n*M elements")
    return A.ctypes.data_as(_vp)


def _M_of(A, n, M):
    if M is not None:
        return int(M)
    if A.ndim == 2:
        return int(A.shape[1]) if int(A.shape[0]) == n else int(A.shape[0])
    return A.size // n if not _is_torch(A) else A.numel() // n


def _execute(kind, direction):
    name = "ft_execute_%s_%s" % (kind, direction)
    fn = getattr(_lib, name)

    def run(plan, A, M=None):
        if _GEOMETRY[plan.kind] != _GEOMETRY[_KIND[kind]]:
            raise ValueError("plan's tables are not in the %s layout" % kind)
        M = _M_of(A, plan.n, M)
        fn(plan.ptr, _pointer(A, plan.n, M), None, M)
        _check()
        return A

    run.__name__ = name[3:]
    run.__doc__ = "%s(plan, A, M=None): in place on the column-major n x M matrix A." % name
    return run


def _kernel(kind, direction):
    name = "ft_kernel_%s_%s" % (kind, direction)
    fn = getattr(_lib, name)

    def run(plan, m1, m2, A, S=1):
        if _GEOMETRY[plan.kind] != _GEOMETRY[_KIND[kind]]:
            raise ValueError("plan's tables are not in the %s layout" % kind)
        span = (plan.n - 1) * int(S) + 1
        if (A.numel() if _is_torch(A) else A.size) < span:
            raise ValueError("buffer shorter than (n-1)*S+1")
        fn(plan.ptr, int(m1), int(m2), _pointer(A.reshape(-1), span, 1), int(S))
        _check()
        return A

    run.__name__ = name[3:]
    return run


execute_sph_hi2lo = _execute("sph", "hi2lo")
execute_sph_lo2hi = _execute("sph", "lo2hi")
execute_tri_hi2lo = _execute("tri", "hi2lo")
execute_tri_lo2hi = _execute("tri", "lo2hi")
execute_disk_hi2lo = _execute("disk", "hi2lo")
execute_disk_lo2hi = _execute("disk", "lo2hi")
execute_sphv_hi2lo = _execute("sphv", "hi2lo")
execute_sphv_lo2hi = _execute("sphv", "lo2hi")
execute_rectdisk_hi2lo = _execute("rectdisk", "hi2lo")
execute_rectdisk_lo2hi = _execute("rectdisk", "lo2hi")
kernel_sph_hi2lo = _kernel("sph", "hi2lo")
kernel_sph_lo2hi = _kernel("sph", "lo2hi")
kernel_tri_hi2lo = _kernel("tri", "hi2lo")
kernel_tri_lo2hi = _kernel("tri", "lo2hi")
kernel_disk_hi2lo = _kernel("disk", "hi2lo")
kernel_disk_lo2hi = _kernel("disk", "lo2hi")
kernel_rectdisk_hi2lo = _kernel("rectdisk", "hi2lo")
kernel_rectdisk_lo2hi = _kernel("rectdisk", "lo2hi")


def _tet_pointer(plan1, plan2, A, L, M):
    if plan1.kind != TRI or plan2.kind != TRI:
        raise ValueError("the tetrahedron takes triangle plans")
    n = plan1.n
    if _is_torch(A):
        if not A.is_contiguous() or A.numel() < n * L * M:
            raise ValueError("expected a contiguous tensor of n*L*M entries")
        import torch
        if A.dtype != torch.float64:
            raise TypeError("expected a torch.float64 tensor")
        return _vp(A.data_ptr())
    if not isinstance(A, np.ndarray) or A.dtype != np.float64:
        raise TypeError("expected a float64 numpy array")
    if not A.flags.f_contiguous or A.size < n * L * M:
        raise ValueError("expected a Fortran-ordered array of n*L*M entries")
    return A.ctypes.data_as(_vp)


def _execute_tet(direction):
    fn = getattr(_lib, "ft_execute_tet_" + direction)

    def run(plan1, plan2, A, L, M):
        fn(plan1.ptr, plan2.ptr, _tet_pointer(plan1, plan2, A, int(L), int(M)), int(L), int(M))
        _check()
        return A

    run.__name__ = "execute_tet_" + direction
    run.__doc__ = ("execute_tet_%s(plan1, plan2, A, L, M): in place on the n x L x M array A (Fortran order; "
                   "slab m is A[:, :, m])." % direction)
    return run


def _kernel_tet(direction):
    fn = getattr(_lib, "ft_kernel_tet_" + direction)

    def run(plan, L, m, A):
        fn(plan.ptr, int(L), int(m), _tet_pointer(plan, plan, A, int(L), 1))
        _check()
        return A

    run.__name__ = "kernel_tet_" + direction
    return run


execute_tet_hi2lo = _execute_tet("hi2lo")
execute_tet_lo2hi = _execute_tet("lo2hi")
kernel_tet_hi2lo = _kernel_tet("hi2lo")
kernel_tet_lo2hi = _kernel_tet("lo2hi")


def gpu_execute_tet(plan1, plan2, direction, A, L, M, stream=None):
    """Asynchronous ft_gpu_execute_tet on a contiguous torch float64 tensor of
    n*L*M entries (the n x L x M array in Fortran order)."""
    import torch
    direction = _DIR.get(direction, direction)
    if stream is None:
        stream = torch.cuda.current_stream(A.device)
    ptr = _tet_pointer(plan1, plan2, A, int(L), int(M))
    rc = _lib.ft_gpu_execute_tet(plan1.ptr, plan2.ptr, int(direction), ptr, int(L), int(M), _vp(stream.cuda_stream))
    if rc:
        raise GPUError("ft_gpu_execute_tet: %s" % _lib.ft_gpu_error_string(rc).decode())
    return A


def gpu_execute(plan, direction, A, M=None, precision=FP64, stream=None, kind=None):
    """Asynchronous device-resident execution (ft_gpu_execute) on a torch tensor.

    ``stream`` is a torch.cuda.Stream (default: the current stream); ``kind``
    the executor (default: the plan's own; "sphv" runs on a sphere plan).
    """
    import torch
    direction = _DIR.get(direction, direction)
    kind = plan.kind if kind is None else _KIND.get(kind, kind)
    M = _M_of(A, plan.n, M)
    if stream is None:
        stream = torch.cuda.current_stream(A.device)
    ptr = _pointer(A, plan.n, M, np.float64 if precision == FP64 else np.float32)
    rc = _lib.ft_gpu_execute(plan.ptr, kind, int(direction), int(precision), ptr, M, _vp(stream.cuda_stream))
    if rc:
        raise GPUError("ft_gpu_execute: %s" % _lib.ft_gpu_error_string(rc).decode())
    return A


def synchronize(stream=None):
    """ft_gpu_synchronize: wait for ``stream`` (a torch.cuda.Stream, default the
    current one) and raise GPUError if an asynchronous launch made on it ended
    with invalid results (a bounded wait expired, or work was left undone)."""
    import torch
    if stream is None:
        stream = torch.cuda.current_stream()
    rc = _lib.ft_gpu_synchronize(_vp(stream.cuda_stream))
    if rc:
        raise GPUError("ft_gpu_synchronize: %s (code %d)" % (_lib.ft_gpu_error_string(rc).decode(), rc))


def gpu_execute_batch(plans, direction, tensors, Ms=None, precision=FP64, stream=None, kind=None):
    """ft_gpu_execute_batch: many transforms of one kind and uneven degree in
    one persistent launch.  ``tensors[i]`` is a CUDA tensor holding the
    column-major plans[i].n x Ms[i] matrix; ``kind`` the executor (default:
    the plans' own)."""
    import torch
    direction = _DIR.get(direction, direction)
    count = len(plans)
    if len(tensors) != count:
        raise ValueError("one tensor per plan")
    if count == 0:
        return tensors
    kind = plans[0].kind if kind is None else _KIND.get(kind, kind)
    if Ms is None:
        Ms = [_M_of(T, P.n, None) for P, T in zip(plans, tensors)]
    dtype = np.float64 if precision == FP64 else np.float32
    ptrs = (_vp * count)(*[_pointer(T, P.n, M, dtype).value for P, T, M in zip(plans, tensors, Ms)])
    pls = (_vp * count)(*[P.ptr for P in plans])
    ms = (ctypes.c_int * count)(*[int(M) for M in Ms])
    if stream is None:
        stream = torch.cuda.current_stream(tensors[0].device)
    rc = _lib.ft_gpu_execute_batch(ctypes.cast(pls, _vp), kind, int(direction), int(precision),
                                   ctypes.cast(ptrs, _vp), ctypes.cast(ms, _vp), count, _vp(stream.cuda_stream))
    if rc:
        raise GPUError("ft_gpu_execute_batch: %s" % _lib.ft_gpu_error_string(rc).decode())
    return tensors


class BatchPlan:
    """ft_gpu_plan_batch: the schedule of one batch, owned by this object, so
    that its launch can be captured in a HIP graph (valid while the object
    lives).  ``execute(stream)`` runs ft_gpu_execute_batch_plan."""

    def __init__(self, plans, direction, tensors, Ms=None, precision=FP64, kind=None):
        direction = _DIR.get(direction, direction)
        kind = plans[0].kind if kind is None and plans else _KIND.get(kind, kind)
        count = len(plans)
        if count == 0 or len(tensors) != count:
            raise ValueError("one tensor per plan, at least one")
        if Ms is None:
            Ms = [_M_of(T, P.n, None) for P, T in zip(plans, tensors)]
        dtype = np.float64 if precision == FP64 else np.float32
        ptrs = (_vp * count)(*[_pointer(T, P.n, M, dtype).value for P, T, M in zip(plans, tensors, Ms)])
        pls = (_vp * count)(*[P.ptr for P in plans])
        ms = (ctypes.c_int * count)(*[int(M) for M in Ms])
        self._keep = (list(plans), list(tensors))
        self.ptr = _lib.ft_gpu_plan_batch(ctypes.cast(pls, _vp), kind, int(direction), int(precision),
                                          ctypes.cast(ptrs, _vp), ctypes.cast(ms, _vp), count)
        if not self.ptr:
            raise GPUError("ft_gpu_plan_batch: %s" % _lib.ft_gpu_error_string(_lib.ft_gpu_last_error()).decode())
        self._destroy = _lib.ft_gpu_destroy_batch

    def execute(self, stream=None):
        import torch
        if stream is None:
            stream = torch.cuda.current_stream()
        rc = _lib.ft_gpu_execute_batch_plan(self.ptr, _vp(stream.cuda_stream))
        if rc:
            raise GPUError("ft_gpu_execute_batch_plan: %s" % _lib.ft_gpu_error_string(rc).decode())

    def __del__(self):
        if getattr(self, "ptr", None):
            self._destroy(self.ptr)
            self.ptr = None


def valid_rows(kind, n, j):
    """Valid rows of column j in each full layout (src/ftutilities.c:67-129;
    sphv uses the sphere's, rectdisk with M = n keeps rows i < n - j)."""
    kind = _KIND.get(kind, kind)
    if kind in (SPH, SPHV):
        return n - (j + 1) // 2
    if kind in (TRI, RECTDISK):
        return n - j
    return n - (j + 3) // 4


def full_M(kind, n):
    """Columns of the full n-row layout: 2n-1 (sph, sphv), n (tri, rectdisk), 4n-3 (disk)."""
    kind = _KIND.get(kind, kind)
    return {SPH: 2 * n - 1, SPHV: 2 * n - 1, TRI: n, RECTDISK: n, DISK: 4 * n - 3}[kind]
