"""ctypes access to the CPU checkers (test infrastructure, not product code).

Only the test suite, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` import this package.  The product library never does.

* ``liboracle.so`` (built from ``ft_oracle.c``): the C restatement of the
  reference's plans, kernels and executors, its ``__float128`` version and the
  seeded valid-triangle fills.
* ``_ref/libftref.so`` (built from the reference's own kernel and permute
  sources plus ``ref_exec.c``): the reference's dispatched executors.  It is
  built where the reference tree exists and travels with the repository;
  ``have_ref()`` says whether it is present.

Matrices are column-major: a numpy array of shape ``(n, M)`` in Fortran order
(``A[i, j]`` is row i of column j), exactly the reference's ``A[i + n*j]``.
"""
import ctypes
import os

import numpy as np

SPH, TRI, DISK, RECTDISK, ANNULUS = 0, 1, 2, 4, 5
HI2LO, LO2HI = 0, 1
KINDS = {"sph": SPH, "tri": TRI, "disk": DISK, "rectdisk": RECTDISK, "annulus": ANNULUS}

_here = os.path.dirname(os.path.abspath(__file__))
_dp = ctypes.POINTER(ctypes.c_double)
_lib = None
_ref = None


def _load():
    global _lib
    if _lib is None:
        path = os.path.join(_here, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (run `make -C oracle`)")
        lib = ctypes.CDLL(path)
        lib.oracle_plan_size.restype = ctypes.c_long
        lib.oracle_plan_size.argtypes = [ctypes.c_int, ctypes.c_int]
        lib.oracle_plan_sph.argtypes = [ctypes.c_int, _dp, _dp]
        lib.oracle_plan_tri.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double, _dp, _dp]
        lib.oracle_plan_disk.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, _dp, _dp]
        lib.oracle_plan_rectdisk.argtypes = [ctypes.c_int, ctypes.c_double, _dp, _dp]
        lib.oracle_connection.restype = ctypes.c_int
        lib.oracle_connection.argtypes = [ctypes.c_int] * 4 + [ctypes.c_double, ctypes.c_double, _dp]
        lib.oracle_plan_annulus.restype = ctypes.c_int
        lib.oracle_plan_annulus.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double, _dp, _dp]
        for name in ("sphv", "rectdisk"):
            getattr(lib, "oracle_execute_" + name).argtypes = [ctypes.c_int, _dp, _dp, ctypes.c_int, _dp, ctypes.c_int]
        for name in ("sph", "tri", "disk"):
            fn = getattr(lib, "oracle_execute_" + name)
            fn.argtypes = [ctypes.c_int, _dp, _dp, ctypes.c_int, _dp, ctypes.c_int]
            for d in ("hi2lo", "lo2hi"):
                k = getattr(lib, "oracle_kernel_%s_%s" % (name, d))
                k.argtypes = [_dp, _dp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, ctypes.c_int]
        lib.oracle_execute_tet.argtypes = [ctypes.c_int, _dp, _dp, _dp, _dp] + [ctypes.c_int] * 3 + [_dp]
        lib.oracle_kernel_tet.argtypes = [ctypes.c_int, _dp, _dp] + [ctypes.c_int] * 3 + [_dp]
        lib.oracle_plan_spin.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp, _dp, _dp]
        lib.oracle_kernel_spinsph.argtypes = [ctypes.c_int, _dp, _dp, _dp, _dp, ctypes.c_int, ctypes.c_int,
                                              ctypes.c_int, _dp, ctypes.c_int]
        lib.oracle_execute_spinsph.argtypes = [ctypes.c_int, _dp, _dp, _dp, _dp, ctypes.c_int, ctypes.c_int, _dp,
                                               ctypes.c_int]
        lib.oracle_execute_q.restype = ctypes.c_int
        lib.oracle_execute_q.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp, ctypes.c_int, ctypes.c_double,
                                         ctypes.c_double, ctypes.c_double, ctypes.c_int, _dp, ctypes.c_int]
        _fp = ctypes.POINTER(ctypes.c_float)
        for name in ("sph", "tri", "disk"):
            for d in ("hi2lo", "lo2hi"):
                k = getattr(lib, "oracle_kernel_%s_%s_f32" % (name, d))
                k.argtypes = [_fp, _fp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _fp, ctypes.c_int]
        lib.oracle_execute_f32.restype = ctypes.c_int
        lib.oracle_execute_f32.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp, ctypes.c_int,
                                           ctypes.POINTER(ctypes.c_float), ctypes.c_int]
        lib.oracle_rand_entry.restype = ctypes.c_double
        lib.oracle_rand_entry.argtypes = [ctypes.c_uint64, ctypes.c_uint64]
        lib.oracle_valid_rows.restype = ctypes.c_int
        lib.oracle_valid_rows.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int]
        lib.oracle_fill.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_uint64, _dp]
        lib.oracle_set_fused.argtypes = [ctypes.c_int]
        _lib = lib
    return _lib


def _ptr(a):
    return a.ctypes.data_as(_dp)


def _kind(kind):
    return KINDS[kind] if isinstance(kind, str) else int(kind)


def _check_matrix(A, n):
    if A.dtype != np.float64 or not A.flags.f_contiguous or A.ndim != 2 or A.shape[0] != n:
        raise ValueError("expected a Fortran-ordered float64 array with n rows")


class Plan:
    """FP64 sine/cosine tables of rotations.c:113-190, 194-260 and 270-287, as
    the oracle builds them.  Plan("rectdisk", n, beta) takes the reference's
    one parameter (ft_plan_rotrectdisk(n, beta)); Plan("annulus", n, alpha,
    beta, rho=r) the annulus of inner radius r with gamma = 0
    (ft_plan_rotannulus(n, alpha, beta, 0, r))."""

    def __init__(self, kind, n, alpha=0.0, beta=0.0, gamma=0.0, rho=0.0):
        lib = _load()
        self.kind, self.n = _kind(kind), int(n)
        if self.kind == RECTDISK:
            alpha, beta = 0.0, alpha
        self.alpha, self.beta, self.gamma, self.rho = float(alpha), float(beta), float(gamma), float(rho)
        size = lib.oracle_plan_size(_geometry(self) if self.kind != RECTDISK else TRI, self.n)
        self.s = np.zeros(size)
        self.c = np.zeros(size)
        if self.kind == SPH:
            lib.oracle_plan_sph(self.n, _ptr(self.s), _ptr(self.c))
        elif self.kind == TRI:
            lib.oracle_plan_tri(self.n, self.alpha, self.beta, self.gamma, _ptr(self.s), _ptr(self.c))
        elif self.kind == RECTDISK:
            lib.oracle_plan_rectdisk(self.n, self.beta, _ptr(self.s), _ptr(self.c))
        elif self.kind == ANNULUS:
            if self.gamma != 0.0:
                raise ValueError("the reference's annulus plan has gamma = 0 only")
            if lib.oracle_plan_annulus(self.n, self.alpha, self.beta, self.rho, _ptr(self.s), _ptr(self.c)) != 0:
                raise MemoryError("oracle_plan_annulus failed")
        else:
            lib.oracle_plan_disk(self.n, self.alpha, self.beta, _ptr(self.s), _ptr(self.c))

    def index(self, l, m):
        n = self.n
        if _geometry(self) == DISK:
            return l + m * n - (m // 2) * ((m + 1) // 2)
        return l + m * (2 * n + 1 - m) // 2


def _geometry(plan):
    """The loop nests of the plan's kind (a rectangle-disk plan runs the
    sphere's, an annulus plan the disk's)."""
    return {RECTDISK: SPH, ANNULUS: DISK}.get(plan.kind, plan.kind)


def execute(plan, direction, A, M=None):
    """ft_execute_<kind>_<direction> restated (in place on A)."""
    lib = _load()
    _check_matrix(A, plan.n)
    M = A.shape[1] if M is None else M
    fn = getattr(lib, "oracle_execute_" + {SPH: "sph", TRI: "tri", DISK: "disk", RECTDISK: "rectdisk",
                                           ANNULUS: "disk"}[plan.kind])
    fn(int(direction), _ptr(plan.s), _ptr(plan.c), plan.n, _ptr(A), int(M))
    return A


def execute_sphv(plan, direction, A, M=None):
    """ft_execute_sphv_<direction> restated (a sphere plan; in place on A)."""
    lib = _load()
    _check_matrix(A, plan.n)
    if plan.kind != SPH:
        raise ValueError("sphv runs on a sphere plan")
    M = A.shape[1] if M is None else M
    lib.oracle_execute_sphv(int(direction), _ptr(plan.s), _ptr(plan.c), plan.n, _ptr(A), int(M))
    return A


class SpinPlan:
    """The spin-weighted sphere's tables as the oracle builds them
    (rotations.c:443-475): s1, c1 (n x n) and s2, c2 (|s| x n x n), in the
    reference's layouts."""

    def __init__(self, n, s):
        self.n, self.s = int(n), int(s)
        a = abs(self.s)
        self.s1, self.c1 = np.zeros(n * n), np.zeros(n * n)
        self.s2, self.c2 = np.zeros(max(a * n * n, 1)), np.zeros(max(a * n * n, 1))
        _load().oracle_plan_spin(self.n, self.s, _ptr(self.s1), _ptr(self.c1), _ptr(self.s2), _ptr(self.c2))

    def _args(self):
        return _ptr(self.s1), _ptr(self.c1), _ptr(self.s2), _ptr(self.c2)


def _check_complex(A, n, M):
    if A.dtype != np.complex128 or not A.flags["F_CONTIGUOUS"] or A.shape[0] != n or A.shape[1] < M:
        raise ValueError("A must be a Fortran-ordered complex128 array of n rows")


def execute_spinsph(plan, direction, A, M=None):
    """ft_execute_spinsph_<direction> restated (drivers.c:687-705; in place
    on the n x M complex array A)."""
    M = A.shape[1] if M is None else int(M)
    _check_complex(A, plan.n, M)
    _load().oracle_execute_spinsph(int(direction), *plan._args(), plan.n, plan.s,
                                   A.ctypes.data_as(_dp), M)
    return A


def kernel_spinsph(plan, direction, m, A, S=1):
    """ft_kernel_spinsph_<direction> restated (rotations_default.c:87-117) on
    the complex vector A with stride S."""
    if A.dtype != np.complex128 or not (A.flags["C_CONTIGUOUS"] or A.flags["F_CONTIGUOUS"]) or \
            A.size < (plan.n - 1) * S + 1:
        raise ValueError("A must be a contiguous complex128 array of (n-1)*S+1 entries")
    _load().oracle_kernel_spinsph(int(direction), *plan._args(), plan.n, plan.s, int(m), A.ctypes.data_as(_dp),
                                  int(S))
    return A


def _check_tet(plan1, plan2, A, L, M):
    if plan1.kind != TRI or plan2.kind != TRI or plan1.n != plan2.n:
        raise ValueError("the tetrahedron takes two triangle plans of one size")
    if not (0 <= L <= plan1.n) or M < 0:
        raise ValueError("L must be in [0, n] and M >= 0")
    if A.dtype != np.float64 or not A.flags["F_CONTIGUOUS"] or A.size < plan1.n * L * M:
        raise ValueError("A must be a Fortran-ordered float64 array of n*L*M entries")


def execute_tet(plan1, plan2, direction, A, L, M):
    """ft_execute_tet_<direction> restated (drivers.c:510-533; in place on the
    Fortran-ordered n x L x M array A)."""
    _check_tet(plan1, plan2, A, L, M)
    _load().oracle_execute_tet(int(direction), _ptr(plan1.s), _ptr(plan1.c), _ptr(plan2.s), _ptr(plan2.c), plan1.n,
                               int(L), int(M), _ptr(A))
    return A


def kernel_tet(plan, direction, L, m, A):
    """ft_kernel_tet_<direction> restated (rotations.c:297-321) on the n x L slab A."""
    _check_tet(plan, plan, A, L, 1)
    _load().oracle_kernel_tet(int(direction), _ptr(plan.s), _ptr(plan.c), plan.n, int(L), int(m), _ptr(A))
    return A


def kernel(plan, direction, m1, m2, A, S=1):
    """ft_kernel_<kind>_<direction> restated on a strided 1-D buffer (in place)."""
    lib = _load()
    if A.dtype != np.float64 or not A.flags.c_contiguous:
        raise ValueError("expected a contiguous float64 buffer")
    name = "oracle_kernel_%s_%s" % (("sph", "tri", "disk")[_geometry(plan)], ("hi2lo", "lo2hi")[direction])
    getattr(lib, name)(_ptr(plan.s), _ptr(plan.c), plan.n, int(m1), int(m2), _ptr(A), int(S))
    return A


def execute_f32(plan, direction, A, M=None):
    """The FP32 variant (float storage, s and c rounded once from the FP64
    plan, fused float rotations), in place on a Fortran-ordered float32 array."""
    lib = _load()
    if A.dtype != np.float32 or not A.flags.f_contiguous or A.ndim != 2 or A.shape[0] != plan.n:
        raise ValueError("expected a Fortran-ordered float32 array with n rows")
    M = A.shape[1] if M is None else M
    rc = lib.oracle_execute_f32(_geometry(plan), int(direction), _ptr(plan.s), _ptr(plan.c), plan.n,
                                A.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), int(M))
    if rc != 0:
        raise MemoryError("oracle_execute_f32 failed")
    return A


def kernel_f32(plan, direction, m1, m2, A, S=1):
    """The FP32 single-vector loop nest on a contiguous float32 buffer (in
    place), with s and c rounded once from the FP64 plan."""
    lib = _load()
    if A.dtype != np.float32 or not A.flags.c_contiguous:
        raise ValueError("expected a contiguous float32 buffer")
    sf = plan.s.astype(np.float32)
    cf = plan.c.astype(np.float32)
    fp = ctypes.POINTER(ctypes.c_float)
    name = "oracle_kernel_%s_%s_f32" % (("sph", "tri", "disk")[_geometry(plan)], ("hi2lo", "lo2hi")[direction])
    getattr(lib, name)(sf.ctypes.data_as(fp), cf.ctypes.data_as(fp), plan.n, int(m1), int(m2),
                       A.ctypes.data_as(fp), int(S))
    return A


def execute_q(plan, direction, A, hp_plan=False, M=None):
    """The same executor in __float128 (rounded to FP64 at the end)."""
    lib = _load()
    _check_matrix(A, plan.n)
    if plan.kind not in (SPH, TRI, DISK):
        raise ValueError("the quadruple-precision executor covers sphere, triangle and disk plans")
    M = A.shape[1] if M is None else M
    rc = lib.oracle_execute_q(plan.kind, int(direction), _ptr(plan.s), _ptr(plan.c), plan.n, plan.alpha,
                              plan.beta, plan.gamma, 1 if hp_plan else 0, _ptr(A), int(M))
    if rc != 0:
        raise MemoryError("oracle_execute_q failed")
    return A


CONNECTIONS = {"leg2cheb": 0, "cheb2leg": 1, "ultra2ultra": 2}


def connection(kind, flag1, flag2, n, lam=0.0, mu=0.0):
    """n x n connection matrix of transforms.c:50-93 (upper triangular):
    connection("leg2cheb", normleg, normcheb, n), ("cheb2leg", normcheb,
    normleg, n), ("ultra2ultra", norm1, norm2, n, lambda, mu)."""
    V = np.zeros((n, n), order="F")
    if _load().oracle_connection(CONNECTIONS[kind], int(flag1), int(flag2), int(n), float(lam), float(mu),
                                 _ptr(V)) != 0:
        raise MemoryError("oracle_connection failed")
    return V


def _tri_harmonic(transform, trans, plan, H, A, M, params):
    """tri2cheb / cheb2tri (drivers.c:921-957): left and right triangular
    products in long double, the 2D Chebyshev normalization (drivers.c:6-24)."""
    alpha, beta, gamma = params
    use0 = (beta + gamma != -1.5) or (alpha != -0.5)
    use1 = (gamma != -0.5) or (beta != -0.5)
    inverse = transform == "cheb2tri"

    def left(U, t):
        U = np.triu(U).astype(np.longdouble)
        A[:, :M] = ((U.T if t else U) @ A[:, :M].astype(np.longdouble)).astype(np.float64)

    def right(U, t):
        U = np.triu(U[:M, :M]).astype(np.longdouble)
        A[:, :M] = (A[:, :M].astype(np.longdouble) @ (U.T if t else U)).astype(np.float64)

    def normalize():
        s1, s2 = (np.sqrt(2.0), np.pi / 2) if inverse else (np.sqrt(0.5), 2 / np.pi)
        A[:, 0] *= s1
        A[0, :M] *= s1
        A[:, :M] *= s2

    if not inverse and trans == "N":
        execute(plan, HI2LO, A, M)
        if use0: left(H[0], False)
        if use1: right(H[1], True)
        normalize()
    elif not inverse:
        normalize()
        if use1: right(H[1], False)
        if use0: left(H[0], True)
        execute(plan, LO2HI, A, M)
    elif trans == "N":
        normalize()
        if use1: right(H[1], True)
        if use0: left(H[0], False)
        execute(plan, LO2HI, A, M)
    else:
        execute(plan, HI2LO, A, M)
        if use0: left(H[0], True)
        if use1: right(H[1], False)
        normalize()
    return A


def _rectdisk_harmonic(transform, trans, plan, H, A, M):
    """rectdisk2cheb / cheb2rectdisk (drivers.c:1143-1169): P[0], P[1] from
    the left on the even and odd columns, P[2] from the right; long double
    products."""
    inverse = transform == "cheb2rectdisk"

    def left(t):
        for c in (0, 1):
            cols = np.arange(c, M, 2)
            if cols.size:
                U = np.triu(H[c]).astype(np.longdouble)
                A[:, cols] = ((U.T if t else U) @ A[:, cols].astype(np.longdouble)).astype(np.float64)

    def right(t):
        U = np.triu(H[2][:M, :M]).astype(np.longdouble)
        A[:, :M] = (A[:, :M].astype(np.longdouble) @ (U.T if t else U)).astype(np.float64)

    if not inverse and trans == "N":
        execute(plan, HI2LO, A, M)
        left(False)
        right(True)
    elif not inverse:
        right(False)
        left(True)
        execute(plan, LO2HI, A, M)
    elif trans == "N":
        right(True)
        left(False)
        execute(plan, LO2HI, A, M)
    else:
        execute(plan, HI2LO, A, M)
        left(True)
        right(False)
    return A


def jacobi_operator(n, a, b):
    """(diagonal, off-diagonal) of the n x n normalized Jacobi(a, b)
    multiplication operator (banded_source.c:1797-1820 with norm = 1): x p_k =
    off_{k-1} p_{k-1} + diag_k p_k + off_k p_{k+1} for the orthonormal
    Jacobi polynomials p_k of weight (1 - x)^a (1 + x)^b."""
    k = np.arange(n, dtype=np.float64)
    s = 2 * k + a + b
    diag = np.empty(n)
    diag[0] = (b - a) / (a + b + 2)
    diag[1:] = (b - a) * (a + b) / (s[1:] * (s[1:] + 2))
    j = k[1:]
    sj = s[1:]
    with np.errstate(invalid="ignore", divide="ignore"):
        off = 2 * np.sqrt(j * (j + a) * (j + b) * (j + a + b) / ((sj - 1) * sj * sj * (sj + 1)))
    if n > 1:  # the first entry: j (j + a + b) / (2 j + a + b - 1) is 0/0 at a + b = -1
        off[0] = 2 * np.sqrt((a + 1) * (b + 1) / ((a + b + 2) ** 2 * (a + b + 3)))
    return diag, off[:n - 1]


def annulus_factor(n, alpha, beta, rho):
    """The upper bidiagonal Cholesky factor R of t I + J (drivers.c:1027-1046):
    J the normalized Jacobi(beta, alpha) operator, t = 1 + 2 rho^2 / (1 -
    rho^2); numpy's Cholesky of the dense matrix."""
    d, e = jacobi_operator(n, beta, alpha)
    t = 1.0 + 2.0 * rho * rho / (1.0 - rho * rho)
    T = np.diag(d + t) + np.diag(e, 1) + np.diag(e, -1)
    return np.triu(np.linalg.cholesky(T).T)


def _bidiag(R, X, solve, trans):
    """x <- R^{-1} x, R^{-T} x, R x or R^T x on every column of X for the upper
    bidiagonal R, in the order of banded_source.c:381-422."""
    n = R.shape[0]
    d, u = np.diag(R).copy(), np.diag(R, 1).copy()
    if solve and not trans:
        for i in range(n - 1, -1, -1):
            X[i] = (X[i] - (u[i] * X[i + 1] if i + 1 < n else 0.0)) / d[i]
    elif solve:
        for i in range(n):
            X[i] = (X[i] - (u[i - 1] * X[i - 1] if i > 0 else 0.0)) / d[i]
    elif not trans:
        for i in range(n):
            X[i] = d[i] * X[i] + (u[i] * X[i + 1] if i + 1 < n else 0.0)
    else:
        for i in range(n - 1, -1, -1):
            X[i] = (u[i - 1] * X[i - 1] if i > 0 else 0.0) + d[i] * X[i]
    return X


def _annulus_harmonic(transform, trans, plan, H, R, A, M):
    """ann2cxf / cxf2ann (drivers.c:1073-1121): the annulus's rotations, the
    bidiagonal R on the columns j = 1, 2 mod 4 (a solve for ann2cxf, a
    product for cxf2ann; the identity on j = 0, 3) and one matrix on every
    column (long double product)."""
    inverse = transform == "cxf2ann"
    cols = np.array([j for j in range(M) if j % 4 in (1, 2)], dtype=np.int64)

    def conv():
        if cols.size:
            A[:, cols] = _bidiag(R, A[:, cols].copy(), not inverse, trans == "T")

    def prod():
        U = np.triu(H).astype(np.longdouble)
        A[:, :M] = ((U.T if trans == "T" else U) @ A[:, :M].astype(np.longdouble)).astype(np.float64)

    rot_first = (trans == "T") if inverse else (trans == "N")
    if rot_first:
        execute(plan, HI2LO, A, M)
        conv()
        prod()
    else:
        prod()
        conv()
        execute(plan, LO2HI, A, M)
    return A


def harmonic(transform, trans, plan, P, Pinv, A, M=None, params=None, R=None):
    """The harmonic transforms of drivers.c:833-1021 on the n x M matrix A (in
    place): transform "sph2fourier", "fourier2sph", "sphv2fourier",
    "fourier2sphv" (plan: the sphere's rotation Plan), "disk2cxf", "cxf2disk"
    (the disk's), "tri2cheb", "cheb2tri" (the triangle's, params its
    (alpha, beta, gamma)) or "rectdisk2cheb", "cheb2rectdisk" (the
    rectangle-disk's; three matrices); trans "N" or "T"; P and Pinv the
    connection matrices and their inverses; "ann2cxf", "cxf2ann" (the
    annulus's Plan, P = (P[0],), R the bidiagonal factor of annulus_factor).
    The rotations are the oracle's; each triangular product is evaluated in
    long double and rounded once."""
    M = A.shape[1] if M is None else M
    if transform in ("ann2cxf", "cxf2ann"):
        return _annulus_harmonic(transform, trans, plan, (Pinv if transform == "cxf2ann" else P)[0], R, A, M)
    if transform in ("tri2cheb", "cheb2tri"):
        return _tri_harmonic(transform, trans, plan, Pinv if transform == "cheb2tri" else P, A, M, params)
    if transform in ("rectdisk2cheb", "cheb2rectdisk"):
        return _rectdisk_harmonic(transform, trans, plan, Pinv if transform == "cheb2rectdisk" else P, A, M)
    vec = transform.startswith("sphv") or transform.endswith("sphv")
    inverse = transform.startswith("fourier") or transform == "cxf2disk"
    H = Pinv if inverse else P
    rot_first = (trans == "T") if inverse else (trans == "N")
    run = execute_sphv if vec else execute
    if rot_first:
        run(plan, HI2LO, A, M)
    for c in range(4):
        cols = np.arange(c, M, 4)
        if cols.size == 0:
            continue
        even = c in (0, 3)
        U = np.triu(H[0] if even != vec else H[1]).astype(np.longdouble)
        if trans == "T":
            U = U.T
        A[:, cols] = (U @ A[:, cols].astype(np.longdouble)).astype(np.float64)
    if not rot_first:
        run(plan, LO2HI, A, M)
    return A


def set_fused(fused):
    """Select the fused (SIMD-path) or unfused (scalar default build) rotation form."""
    _load().oracle_set_fused(1 if fused else 0)


def valid_rows(kind, n, j):
    return _load().oracle_valid_rows(_kind(kind), int(n), int(j))


def fill(kind, n, M, seed):
    """Seeded U(-1,1) fill of each layout's valid triangle (zeros elsewhere)."""
    A = np.zeros((n, M), order="F")
    _load().oracle_fill(_kind(kind), int(n), int(M), ctypes.c_uint64(seed), _ptr(A))
    return A


def rand_entry(seed, k):
    return _load().oracle_rand_entry(ctypes.c_uint64(seed), ctypes.c_uint64(k))


# ---------------------------------------------------------------- reference --

def have_ref():
    return os.path.exists(os.path.join(_here, "_ref", "libftref.so"))


def _load_ref():
    global _ref
    if _ref is None:
        path = os.path.join(_here, "_ref", "libftref.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/_ref/libftref.so is not built (run `make -C oracle ref`)")
        lib = ctypes.CDLL(path)
        lib.ref_execute.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, _dp, ctypes.c_int, _dp, _dp,
                                    ctypes.c_int]
        lib.ref_kernel_column.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp, ctypes.c_int, ctypes.c_int,
                                          ctypes.c_int, _dp, ctypes.c_int]
        for name in ("sphv", "rectdisk"):
            getattr(lib, "ref_execute_" + name).argtypes = [ctypes.c_int, _dp, _dp, ctypes.c_int, _dp, ctypes.c_int]
        lib.ref_select_isa.restype = ctypes.c_int
        lib.ref_execute_tet.argtypes = [ctypes.c_int, _dp, _dp, _dp, _dp] + [ctypes.c_int] * 3 + [_dp]
        lib.ref_execute_spinsph.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp, _dp, _dp, ctypes.c_int,
                                            ctypes.c_int, _dp, _dp, ctypes.c_int]
        lib.ref_set_num_threads.argtypes = [ctypes.c_int]
        lib.ref_max_threads.restype = ctypes.c_int
        _ref = lib
    return _ref


REF_ISA_NAMES = {0: "default", 1: "SSE2", 2: "AVX", 3: "AVX_FMA", 4: "AVX512F"}


def ref_isa():
    return _load_ref().ref_select_isa()


def ref_set_num_threads(t):
    _load_ref().ref_set_num_threads(int(t))


def ref_execute(plan, direction, A, M=None, isa=-1, B=None):
    """The reference's ft_execute_* (its compiled kernels, dispatched as drivers.c does)."""
    lib = _load_ref()
    _check_matrix(A, plan.n)
    M = A.shape[1] if M is None else M
    if B is None:
        B = np.zeros(plan.n * max(M, 1))
    lib.ref_execute(_geometry(plan), int(direction), int(isa), _ptr(plan.s), _ptr(plan.c), plan.n, _ptr(A), _ptr(B),
                    int(M))
    return A


def ref_execute_sphv(plan, direction, A, M=None):
    """The reference's scalar ft_execute_sphv_* (its compiled sphere kernels)."""
    lib = _load_ref()
    _check_matrix(A, plan.n)
    M = A.shape[1] if M is None else M
    lib.ref_execute_sphv(int(direction), _ptr(plan.s), _ptr(plan.c), plan.n, _ptr(A), int(M))
    return A


def ref_execute_rectdisk(plan, direction, A, M=None):
    """The reference's ft_execute_rectdisk_* (its compiled rectangle-disk kernels)."""
    lib = _load_ref()
    _check_matrix(A, plan.n)
    M = A.shape[1] if M is None else M
    lib.ref_execute_rectdisk(int(direction), _ptr(plan.s), _ptr(plan.c), plan.n, _ptr(A), int(M))
    return A


def ref_execute_tet(plan1, plan_L, direction, A, L, M):
    """The reference's scalar ft_execute_tet_* through its compiled triangle
    kernels: plan1 of size n for the slab columns, plan_L the triangle plan of
    size L with the block rotations' parameters (oracle/ref_exec.c)."""
    lib = _load_ref()
    if plan_L.n != L:
        raise ValueError("plan_L must have size L")
    _check_tet(plan1, plan1, A, L, M)
    lib.ref_execute_tet(int(direction), _ptr(plan1.s), _ptr(plan1.c), _ptr(plan_L.s), _ptr(plan_L.c), plan1.n,
                        int(L), int(M), _ptr(A))
    return A


def ref_execute_spinsph(plan, direction, A, M=None, isa="default"):
    """The reference's ft_execute_spinsph_* through its compiled spin kernels:
    isa "default" the scalar executor, "avx_fma" the dispatched one (scalar
    orders 0, +-1, fused twisted kernels on the rest)."""
    lib = _load_ref()
    M = A.shape[1] if M is None else int(M)
    _check_complex(A, plan.n, M)
    B = np.zeros(2 * plan.n * max(M, 1))
    lib.ref_execute_spinsph(int(direction), 0 if isa == "default" else 3, *plan._args(), plan.n, plan.s,
                            A.ctypes.data_as(_dp), _ptr(B), M)
    return A


def ref_kernel(plan, direction, m1, m2, A, S=1):
    lib = _load_ref()
    lib.ref_kernel_column(_geometry(plan), int(direction), _ptr(plan.s), _ptr(plan.c), plan.n, int(m1), int(m2),
                          _ptr(A), int(S))
    return A


def rel2(A, B):
    """Relative 2-norm ||A-B|| / ||B|| (ftutilities_source.c:6-20)."""
    d = np.linalg.norm((np.asarray(A) - np.asarray(B)).ravel())
    b = np.linalg.norm(np.asarray(B).ravel())
    return d / b if b else d


def relinf(A, B):
    """Relative inf-norm max|A-B| / max|B| (ftutilities_source.c:22-36)."""
    d = np.max(np.abs(np.asarray(A) - np.asarray(B)))
    b = np.max(np.abs(np.asarray(B)))
    return d / b if b else d


# ---------------------------------------------------------------- grids --
# fftw.c:40-270 restated with scipy.fft / numpy.fft (pocketfft) in place of
# FFTW: the r2r kinds are FFTW's definitions (REDFT01 = scipy DCT-III,
# RODFT01 = DST-III, REDFT10 = DCT-II, RODFT10 = DST-II, all unnormalized),
# the phi DFTs unnormalized real transforms.  Pinned in tests against the
# definitions summed directly and against closed forms.

_GRID_C = {"sqrt2": 1.4142135623730950488, "sqrt1_2": 0.70710678118654752440,
           "1_4_sqrt_pi": 0.141047395886939072, "1_2_sqrt_pi": 0.282094791773878143,
           "4_sqrt_pi": 7.089815403622064109, "2_sqrt_pi": 3.544907701811032055}


def _grid_classes(M, sphv):
    """Cosine-class columns of a plan (fftw.c:84-96: columns j mod 4 in {0, 3}
    take kind[0], REDFT for the sphere's plans, RODFT for the vector field's;
    sphv == "annulus": every column, fftw.c:560-575)."""
    j = np.arange(M)
    if sphv == "annulus":
        return np.ones(M, dtype=bool)
    c = (j % 4 == 0) | (j % 4 == 3)
    return ~c if sphv else c


def annulus_weights(N, rho):
    """fftw.c:540-548."""
    w = np.empty(N)
    for n in range(N):
        t = (N - n - 0.5) * np.pi / (2.0 * N)
        ct, st = np.sin(t), np.cos(t)
        w[n] = np.sqrt(2.0 * (ct * ct + rho * rho * st * st) / (1.0 - rho * rho))
    return w


def _theta(X, synthesis, plan_sphv, disk=False):
    import scipy.fft as sf
    cos = _grid_classes(X.shape[1], plan_sphv)
    t = 3 if synthesis else 2
    X[:, cos] = sf.dct(X[:, cos], type=t, axis=0)
    if disk:   # fftw.c:450-458: REDFT11 (scipy's DCT-IV) in both plan kinds
        X[:, ~cos] = sf.dct(X[:, ~cos], type=4, axis=0)
    else:
        X[:, ~cos] = sf.dst(X[:, ~cos], type=t, axis=0)


def _data_r2c(X):
    """fftw.c:40-49: column 0 the DC term, columns 2j, 2j-1 (re, -im) of frequency j."""
    N, M = X.shape
    Y = np.zeros((N, M // 2 + 1), dtype=np.complex128)
    Y[:, 0] = X[:, 0]
    for j in range(1, M // 2 + 1):
        Y[:, j] = X[:, 2 * j] - 1j * X[:, 2 * j - 1]
    return Y


def _data_c2r(Y, M):
    N = Y.shape[0]
    X = np.zeros((N, M))
    X[:, 0] = Y[:, 0].real
    for j in range(1, M // 2 + 1):
        X[:, 2 * j] = Y[:, j].real
        X[:, 2 * j - 1] = -Y[:, j].imag
    return X


def grid(which, trans, synthesis_plan, plan_sphv, X, rho=None):
    """ft_execute_{sph,sphv,disk}_{synthesis,analysis}(TRANS, P, X, N, M)
    restated (fftw.c:126-270, 460-520) on a copy of the N x M array X: which
    is "sph_synthesis", "sph_analysis", "sphv_synthesis", "sphv_analysis",
    "disk_synthesis" or "disk_analysis"; the plan is a synthesis plan
    (DCT-III / DST-III, c2r) or not (DCT-II / DST-II, r2c), plan_sphv its
    column classes (the disk's plans: the sphere's classes, DCT-IV on the
    sine class)."""
    X = np.array(X, dtype=np.float64, order="F")
    N, M = X.shape
    analysis, exec_sphv = "analysis" in which, which.startswith("sphv")
    disk = which.startswith("disk")   # fftw.c:460-520: no last-row fix-ups
    ann = which.startswith("annulus")  # fftw.c:583-655: one class, radial weights
    if ann:
        plan_sphv = exec_sphv = "annulus"
        wcols = (np.arange(M) % 4 == 1) | (np.arange(M) % 4 == 2)
        w = annulus_weights(N, rho)[:, None]
    cos = _grid_classes(M, exec_sphv)
    c = _GRID_C
    pre3 = (not analysis and trans == "N") or (analysis and trans == "T")
    if pre3 != bool(synthesis_plan):
        raise ValueError("this executor and TRANS take the other plan")
    if pre3:
        if trans == "N":   # synthesis: x2 on row 0 of the cosine columns
            X[0, cos] *= 2.0
            a = c["1_4_sqrt_pi"]
        else:              # analysis 'T': x2 on the last row of the sine columns
            if not disk:
                X[N - 1, ~cos] *= 2.0
            a = c["2_sqrt_pi"] / (2 * N * M)
        _theta(X, True, plan_sphv, disk)
        if ann:
            X[:, wcols] = X[:, wcols] * w if trans == "N" else X[:, wcols] / w
        X *= a
        X[:, 0] *= c["sqrt2"]
        return np.fft.irfft(_data_r2c(X), n=M, axis=1) * M
    Y = np.fft.rfft(X, axis=1)
    X = _data_c2r(Y, M)
    if trans == "T":       # synthesis 'T'
        X *= c["1_2_sqrt_pi"]
    else:                  # analysis 'N'
        X *= c["4_sqrt_pi"] / (2 * N * M)
    X[:, 0] *= c["sqrt1_2"]
    if ann:
        X[:, wcols] = X[:, wcols] * w if trans == "T" else X[:, wcols] / w
    _theta(X, False, plan_sphv, disk)
    if trans == "T":
        if not disk:
            X[N - 1, ~cos] *= 0.5
    else:
        X[0, cos] *= 0.5
    return X


def tri_grid(analysis, trans, synthesis_plan, X):
    """ft_execute_tri_{synthesis,analysis}(TRANS, P, X, N, M) restated
    (fftw.c:280-330) on a copy of the N x M column-major array X: FFTW's
    r2r_2d(N, M) reads the buffer as an N x M row-major array; the plan's
    kinds are REDFT01 on both axes (synthesis plan) or REDFT10."""
    import scipy.fft as sf
    X = np.array(X, dtype=np.float64, order="F")
    N, M = X.shape
    if N < 2 or M < 2:
        return X
    buf = X.reshape(-1, order="F").copy()
    e = np.arange(N * M)
    first = (e < N).astype(float) + 1.0   # 2 where e < N
    zero = (e % N == 0).astype(float) + 1.0

    def r2r(b):
        B = b.reshape((N, M))             # row-major view
        t = 3 if synthesis_plan else 2
        return sf.dct(sf.dct(B, type=t, axis=1), type=t, axis=0).reshape(-1)

    nm = float(N) * float(M)
    if not analysis and trans == "N":
        buf = buf * first * zero * 0.25
        buf = r2r(buf)
    elif not analysis:
        buf = r2r(buf * 0.25)
    elif trans == "N":
        buf = r2r(buf)
        buf = buf / first / zero / nm
    else:
        buf = r2r(buf / nm)
    return buf.reshape((N, M), order="F")


def tet_grid(analysis, trans, synthesis_plan, X):
    """ft_execute_tet_{synthesis,analysis}(TRANS, P, X, N, L, M) restated
    (fftw.c:333-395) on a copy of the N x L x M column-major array X: FFTW's
    r2r_3d(N, L, M) reads the buffer as a row-major N x L x M array; the face
    scalings act where a coordinate of the column-major view is 0."""
    import scipy.fft as sf
    X = np.array(X, dtype=np.float64, order="F")
    N, L, M = X.shape
    if N < 2 or L < 2 or M < 2:
        return X
    buf = X.reshape(-1, order="F").copy()
    e = np.arange(N * L * M)
    i, j, k = e % N, (e // N) % L, e // (N * L)
    faces = (i == 0).astype(int) + (j == 0).astype(int) + (k == 0).astype(int)

    def r2r(b):
        B = b.reshape((N, L, M))          # row-major view
        t = 3 if synthesis_plan else 2
        for ax in (2, 1, 0):
            B = sf.dct(B, type=t, axis=ax)
        return B.reshape(-1)

    nlm = float(N) * float(L) * float(M)
    if not analysis and trans == "N":
        buf = r2r(buf * 2.0 ** faces * 0.125)
    elif not analysis:
        buf = r2r(buf) * 0.125
    elif trans == "N":
        buf = r2r(buf) * 0.5 ** faces / nlm
    else:
        buf = r2r(buf / nlm)
    return buf.reshape((N, L, M), order="F")


def rect_grid(analysis, trans, synthesis_plan, X):
    """ft_execute_rectdisk_{synthesis,analysis}(TRANS, P, X, N, M) restated
    (fftw.c:657-740) on a copy of the N x M column-major array X: x, the
    column transforms (REDFT on the even columns, RODFT on the odd ones), and
    y, REDFT along the rows, of kind 01 (synthesis plan) or 10."""
    import scipy.fft as sf
    X = np.array(X, dtype=np.float64, order="F")
    N, M = X.shape
    t = 3 if synthesis_plan else 2
    nm = float(N) * float(M)

    def x(A):
        A[:, 0::2] = sf.dct(A[:, 0::2], type=t, axis=0)
        A[:, 1::2] = sf.dst(A[:, 1::2], type=t, axis=0)

    def y(A):
        A[:, :] = sf.dct(A, type=t, axis=1)

    if not analysis and trans == "N":
        X[0, 0::2] *= 2.0
        x(X)
        X[:, 0] *= 2.0
        X *= 0.25
        y(X)
    elif not analysis:
        y(X)
        X *= 0.25
        x(X)
        X[N - 1, 1::2] *= 0.5
    elif trans == "N":
        y(X)
        X /= nm
        X[:, 0] *= 0.5
        x(X)
        X[0, 0::2] *= 0.5
    else:
        X[N - 1, 1::2] *= 2.0
        x(X)
        y(X)
        X /= nm
    return X


def spin_grid(analysis, trans, synthesis_plan, S, X):
    """ft_execute_spinsph_{synthesis,analysis}(TRANS, P, X, N, M) restated
    (fftw.c:742-937) on a copy of the N x M complex array X: the theta
    transforms of the sphere's (even S) or the vector field's (odd S) column
    classes on the real and imaginary parts, colswap (fftw.c:6-21), a complex
    DFT along the rows (synthesis plans backward, analysis plans forward)."""
    X = np.array(X, dtype=np.complex128, order="F")
    N, M = X.shape
    sphv = S % 2 != 0
    cos = _grid_classes(M, sphv)
    perm = np.zeros(M, dtype=int)          # X column j goes to Y column perm[j]
    for j in range(1, M):
        perm[j] = j // 2 if j % 2 == 0 else M - (j + 1) // 2

    def theta(A, syn):
        re, im = np.array(A.real, order="F"), np.array(A.imag, order="F")
        _theta(re, syn, sphv)
        _theta(im, syn, sphv)
        return re + 1j * im

    s12, s2 = 0.199471140200716338, 5.013256549262001005
    pre3 = (not analysis and trans == "N") or (analysis and trans == "C")
    if pre3 != bool(synthesis_plan):
        raise ValueError("this executor and TRANS take the other plan")
    if pre3:
        if trans == "N":
            X[0, cos] *= 2.0
            a = s12
        else:
            X[N - 1, ~cos] *= 2.0
            a = s2 / (2 * N * M)
        X = theta(X, True) * a
        Y = np.zeros_like(X)
        Y[:, perm] = X
        return np.fft.ifft(Y, axis=1) * M
    Y = np.fft.fft(X, axis=1)
    X = Y[:, perm]
    X = X * (s12 if trans == "C" else s2 / (2 * N * M))
    X = theta(X, False)
    if trans == "C":
        X[N - 1, ~cos] *= 0.5
    else:
        X[0, cos] *= 0.5
    return X
