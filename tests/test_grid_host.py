"""Synthesis and analysis on the sphere's grid (fftw.c:40-270) on the CPU.

The oracle (oracle.grid: scipy / numpy FFTs in place of FFTW) is pinned by:

* its r2r transforms against FFTW's definitions of REDFT01, RODFT01,
  REDFT10 and RODFT10 summed directly in long double, and its real DFTs
  against the unnormalized DFT sums (FFTW's r2c / c2r conventions);
* synthesis 'N' of every unit coefficient against the closed form it must
  produce on the grid theta_n = (2n+1) pi / 2N, phi_j = 2 pi j / M:
  a * w * {cos(l theta), sin((l+1) theta)} * {sqrt 2, 2 cos(q phi), 2 sin(q phi)}
  for the column classes of the sphere's and the vector field's plans;
* 'T' is the transpose of 'N' (dense operators at small sizes), for both
  executors and both plan families;
* analysis 'N' inverts synthesis 'N'.

The library's grid plans need a device; here only their argument checks run.
"""
import numpy as np
import pytest

import oracle as O

EPS = 2.0 ** -52


def _direct_r2r(x, kind):
    """FFTW's r2r definitions (FFTW manual, 'What FFTW Really Computes'), in long double."""
    x = np.asarray(x, dtype=np.longdouble)
    n = len(x)
    k = np.arange(n, dtype=np.longdouble)
    pi = np.longdouble(np.pi)
    y = np.zeros(n, dtype=np.longdouble)
    for i in range(n):
        if kind == "REDFT01":
            y[i] = x[0] + 2 * np.sum(x[1:] * np.cos(pi * k[1:] * (i + np.longdouble(0.5)) / n))
        elif kind == "RODFT01":
            y[i] = (-1) ** i * x[n - 1] + 2 * np.sum(x[:n - 1] * np.sin(pi * (k[:n - 1] + 1) * (i + np.longdouble(0.5)) / n))
        elif kind == "REDFT10":
            y[i] = 2 * np.sum(x * np.cos(pi * (k + np.longdouble(0.5)) * i / n))
        elif kind == "REDFT11":
            y[i] = 2 * np.sum(x * np.cos(pi * (k + np.longdouble(0.5)) * (i + np.longdouble(0.5)) / n))
        else:  # RODFT10
            y[i] = 2 * np.sum(x * np.sin(pi * (k + np.longdouble(0.5)) * (i + 1) / n))
    return y.astype(np.float64)


@pytest.mark.parametrize("n", [1, 2, 5, 8, 17, 64])
def test_r2r_kinds_are_fftw_definitions(n):
    import scipy.fft as sf
    x = np.random.default_rng(n).standard_normal(n)
    for kind, got in (("REDFT01", sf.dct(x, type=3)), ("RODFT01", sf.dst(x, type=3)),
                      ("REDFT10", sf.dct(x, type=2)), ("RODFT10", sf.dst(x, type=2)),
                      ("REDFT11", sf.dct(x, type=4))):
        ref = _direct_r2r(x, kind)
        assert np.max(np.abs(got - ref)) <= 64 * n * EPS * max(1.0, np.max(np.abs(ref))), kind


@pytest.mark.parametrize("M", [1, 3, 7, 15])
def test_real_dfts_are_unnormalized_sums(M):
    x = np.random.default_rng(M).standard_normal(M)
    j = np.arange(M)
    Y = np.fft.rfft(x)
    ref = np.array([np.sum(x * np.exp(-2j * np.pi * j * k / M)) for k in range(M // 2 + 1)])
    assert np.max(np.abs(Y - ref)) <= 64 * M * EPS
    back = np.fft.irfft(Y, n=M) * M
    ref2 = np.array([Y[0].real + 2 * np.sum((Y[1:] * np.exp(2j * np.pi * np.arange(1, M // 2 + 1) * t / M)).real)
                     for t in j])
    assert np.max(np.abs(back - ref2)) <= 64 * M * EPS * max(1.0, np.max(np.abs(ref2)))


@pytest.mark.parametrize("sphv", [False, True])
@pytest.mark.parametrize("N,M", [(1, 1), (4, 7), (6, 11)])
def test_synthesis_of_unit_coefficients(sphv, N, M):
    which = "sphv_synthesis" if sphv else "sph_synthesis"
    cos = O._grid_classes(M, sphv)
    theta = (2 * np.arange(N) + 1) * np.pi / (2 * N)
    phi = 2 * np.pi * np.arange(M) / M
    a = 1.0 / (4.0 * np.sqrt(np.pi))
    for l in range(N):
        for col in range(M):
            X = np.zeros((N, M))
            X[l, col] = 1.0
            G = O.grid(which, "N", True, sphv, X)
            if cos[col]:
                t = 2 * np.cos(l * theta)
            else:
                t = (2.0 if l < N - 1 else 1.0) * np.sin((l + 1) * theta)
            if col == 0:
                p = np.full(M, np.sqrt(2.0))
            elif col % 2 == 0:
                p = 2 * np.cos(col // 2 * phi)
            else:
                p = 2 * np.sin((col + 1) // 2 * phi)
            assert np.max(np.abs(G - a * np.outer(t, p))) <= 32 * (N + M) * EPS, (l, col)


def _dense(which, trans, syn_plan, sphv, N, M):
    D = np.zeros((N * M, N * M))
    for i in range(N * M):
        e = np.zeros(N * M)
        e[i] = 1.0
        D[:, i] = O.grid(which, trans, syn_plan, sphv, e.reshape((N, M), order="F")).reshape(-1, order="F")
    return D


@pytest.mark.parametrize("sphv", [False, True])
@pytest.mark.parametrize("N,M", [(3, 5), (4, 7)])
def test_transposes_and_inverse(sphv, N, M):
    pre = "sphv_" if sphv else "sph_"
    SN = _dense(pre + "synthesis", "N", True, sphv, N, M)
    ST = _dense(pre + "synthesis", "T", False, sphv, N, M)
    AN = _dense(pre + "analysis", "N", False, sphv, N, M)
    AT = _dense(pre + "analysis", "T", True, sphv, N, M)
    assert np.max(np.abs(ST - SN.T)) <= 64 * EPS * N * M
    assert np.max(np.abs(AT - AN.T)) <= 64 * EPS * N * M
    assert np.max(np.abs(AN @ SN - np.eye(N * M))) <= 64 * EPS * N * M


def test_library_rejects_bad_shapes(ft):
    with pytest.raises(ft.GPUError):
        ft.plan_sph_synthesis(8, 14)
    with pytest.raises(ft.GPUError):
        ft.plan_sph_analysis(0, 15)


@pytest.mark.parametrize("N", [2, 3, 6])
def test_tri_synthesis_of_unit_coefficients(N):
    """fftw.c:290-300 with N = M: a unit coefficient at (l, c) gives
    cos(l theta_n) cos(c theta_m) on the grid theta_n = (2n+1) pi / 2N."""
    th = (2 * np.arange(N) + 1) * np.pi / (2 * N)
    for l in range(N):
        for c in range(N):
            X = np.zeros((N, N))
            X[l, c] = 1.0
            G = O.tri_grid(False, "N", True, X)
            assert np.max(np.abs(G - np.outer(np.cos(l * th), np.cos(c * th)))) <= 32 * N * EPS, (l, c)


@pytest.mark.parametrize("N,M", [(2, 2), (3, 3), (4, 4), (7, 7)])
def test_tri_transposes_and_inverse(N, M):
    """For N = M (the reference's tests' shape).  For N != M the reference's
    scalings (on the column-major array's first column and row) and FFTW's
    r2r_2d (over the buffer as an N x M row-major array) address different
    elements, so these identities do not hold there; the library reproduces
    that behaviour as the oracle restates it."""
    def dense(analysis, trans, syn):
        D = np.zeros((N * M, N * M))
        for i in range(N * M):
            e = np.zeros(N * M)
            e[i] = 1.0
            D[:, i] = O.tri_grid(analysis, trans, syn, e.reshape((N, M), order="F")).reshape(-1, order="F")
        return D
    SN, ST = dense(False, "N", True), dense(False, "T", False)
    AN, AT = dense(True, "N", False), dense(True, "T", True)
    assert np.max(np.abs(ST - SN.T)) <= 64 * EPS * N * M
    assert np.max(np.abs(AT - AN.T)) <= 64 * EPS * N * M
    assert np.max(np.abs(AN @ SN - np.eye(N * M))) <= 64 * EPS * N * M


def test_tri_small_sizes_are_untouched():
    X = np.random.default_rng(1).standard_normal((1, 5))
    assert np.array_equal(O.tri_grid(False, "N", True, X), X)


@pytest.mark.parametrize("N,M", [(1, 1), (4, 9), (5, 13)])
def test_disk_synthesis_of_unit_coefficients(N, M):
    """fftw.c:460-478: the disk's cosine class as the sphere's, its sine class
    through REDFT11: a unit coefficient at row l gives 2 cos((l + 1/2) theta)."""
    cos = O._grid_classes(M, False)
    theta = (2 * np.arange(N) + 1) * np.pi / (2 * N)
    phi = 2 * np.pi * np.arange(M) / M
    a = 1.0 / (4.0 * np.sqrt(np.pi))
    for l in range(N):
        for col in range(M):
            X = np.zeros((N, M))
            X[l, col] = 1.0
            G = O.grid("disk_synthesis", "N", True, False, X)
            t = 2 * np.cos(l * theta) if cos[col] else 2 * np.cos((l + 0.5) * theta)
            if col == 0:
                p = np.full(M, np.sqrt(2.0))
            elif col % 2 == 0:
                p = 2 * np.cos(col // 2 * phi)
            else:
                p = 2 * np.sin((col + 1) // 2 * phi)
            assert np.max(np.abs(G - a * np.outer(t, p))) <= 32 * (N + M) * EPS, (l, col)


@pytest.mark.parametrize("N,M", [(3, 5), (4, 9)])
def test_disk_transposes_and_inverse(N, M):
    SN = _dense("disk_synthesis", "N", True, False, N, M)
    ST = _dense("disk_synthesis", "T", False, False, N, M)
    AN = _dense("disk_analysis", "N", False, False, N, M)
    AT = _dense("disk_analysis", "T", True, False, N, M)
    assert np.max(np.abs(ST - SN.T)) <= 64 * EPS * N * M
    assert np.max(np.abs(AT - AN.T)) <= 64 * EPS * N * M
    assert np.max(np.abs(AN @ SN - np.eye(N * M))) <= 64 * EPS * N * M


@pytest.mark.parametrize("N", [2, 3, 4])
def test_tet_synthesis_of_unit_coefficients(N):
    """fftw.c:349-362 with N = L = M: a unit coefficient at (l, m, p) gives
    cos(l theta_i) cos(m theta_j) cos(p theta_k) on the grid."""
    th = (2 * np.arange(N) + 1) * np.pi / (2 * N)
    for l in range(N):
        for m in range(N):
            for p in range(N):
                X = np.zeros((N, N, N), order="F")
                X[l, m, p] = 1.0
                G = O.tet_grid(False, "N", True, X)
                ref = np.einsum("i,j,k->ijk", np.cos(l * th), np.cos(m * th), np.cos(p * th))
                assert np.max(np.abs(G - ref)) <= 64 * N * EPS, (l, m, p)


@pytest.mark.parametrize("N", [2, 3])
def test_tet_transposes_and_inverse(N):
    L = M = N
    size = N * L * M

    def dense(analysis, trans, syn):
        D = np.zeros((size, size))
        for c in range(size):
            e = np.zeros(size)
            e[c] = 1.0
            D[:, c] = O.tet_grid(analysis, trans, syn, e.reshape((N, L, M), order="F")).reshape(-1, order="F")
        return D
    SN, ST = dense(False, "N", True), dense(False, "T", False)
    AN, AT = dense(True, "N", False), dense(True, "T", True)
    assert np.max(np.abs(ST - SN.T)) <= 64 * EPS * size
    assert np.max(np.abs(AT - AN.T)) <= 64 * EPS * size
    assert np.max(np.abs(AN @ SN - np.eye(size))) <= 64 * EPS * size


@pytest.mark.parametrize("rho", [0.0, 0.5])
def test_annulus_synthesis_of_unit_coefficients(rho):
    """fftw.c:583-600: every column through REDFT01, the radial weights on
    columns 1 and 2 mod 4: a unit coefficient at row l gives
    2 cos(l theta) w (those columns) or 2 cos(l theta)."""
    N, M = 5, 9
    theta = (2 * np.arange(N) + 1) * np.pi / (2 * N)
    phi = 2 * np.pi * np.arange(M) / M
    a = 1.0 / (4.0 * np.sqrt(np.pi))
    w = O.annulus_weights(N, rho)
    # the weights are sqrt(2/(1-rho^2)) r at the grid's radii, (2r^2-1-rho^2)/(1-rho^2) = cos(theta)
    # (fftw.c:535-538)
    r = np.sqrt((1 + rho ** 2 + (1 - rho ** 2) * np.cos(theta)) / 2)
    assert np.max(np.abs(w - np.sqrt(2 / (1 - rho ** 2)) * r)) <= 8 * EPS
    for l in range(N):
        for col in range(M):
            X = np.zeros((N, M))
            X[l, col] = 1.0
            G = O.grid("annulus_synthesis", "N", True, False, X, rho=rho)
            t = 2 * np.cos(l * theta) * (w if col % 4 in (1, 2) else 1.0)
            if col == 0:
                p = np.full(M, np.sqrt(2.0))
            elif col % 2 == 0:
                p = 2 * np.cos(col // 2 * phi)
            else:
                p = 2 * np.sin((col + 1) // 2 * phi)
            assert np.max(np.abs(G - a * np.outer(t, p))) <= 64 * (N + M) * EPS, (l, col)


def test_annulus_transposes_and_inverse():
    N, M, rho = 4, 7, 0.3

    def dense(which, trans, syn):
        D = np.zeros((N * M, N * M))
        for i in range(N * M):
            e = np.zeros(N * M)
            e[i] = 1.0
            D[:, i] = O.grid(which, trans, syn, False, e.reshape((N, M), order="F"), rho=rho).reshape(-1, order="F")
        return D
    SN, ST = dense("annulus_synthesis", "N", True), dense("annulus_synthesis", "T", False)
    AN, AT = dense("annulus_analysis", "N", False), dense("annulus_analysis", "T", True)
    assert np.max(np.abs(ST - SN.T)) <= 64 * EPS * N * M
    assert np.max(np.abs(AT - AN.T)) <= 64 * EPS * N * M
    assert np.max(np.abs(AN @ SN - np.eye(N * M))) <= 64 * EPS * N * M


@pytest.mark.parametrize("N,M", [(3, 4), (4, 5), (1, 3)])
def test_rectdisk_synthesis_of_unit_coefficients(N, M):
    """fftw.c:690-700: a unit coefficient at (l, c) gives
    {cos(l theta), sin((l+1) theta) (1/2 at l = N-1)} x cos(c phi) with
    theta_n = (2n+1) pi / 2N, phi_m = (2m+1) pi / 2M (even c: cosines, odd c:
    sines along the columns)."""
    th = (2 * np.arange(N) + 1) * np.pi / (2 * N)
    ph = (2 * np.arange(M) + 1) * np.pi / (2 * M)
    for l in range(N):
        for c in range(M):
            X = np.zeros((N, M))
            X[l, c] = 1.0
            G = O.rect_grid(False, "N", True, X)
            if c % 2 == 0:
                t = np.cos(l * th)
            else:
                t = (0.5 if l == N - 1 else 1.0) * np.sin((l + 1) * th)
            ref = np.outer(t, np.cos(c * ph))
            assert np.max(np.abs(G - ref)) <= 32 * (N + M) * EPS, (l, c)


@pytest.mark.parametrize("N,M", [(3, 4), (4, 5)])
def test_rectdisk_transposes_and_inverse(N, M):
    def dense(analysis, trans, syn):
        D = np.zeros((N * M, N * M))
        for i in range(N * M):
            e = np.zeros(N * M)
            e[i] = 1.0
            D[:, i] = O.rect_grid(analysis, trans, syn, e.reshape((N, M), order="F")).reshape(-1, order="F")
        return D
    SN, ST = dense(False, "N", True), dense(False, "T", False)
    AN, AT = dense(True, "N", False), dense(True, "T", True)
    assert np.max(np.abs(ST - SN.T)) <= 64 * EPS * N * M
    assert np.max(np.abs(AT - AN.T)) <= 64 * EPS * N * M
    assert np.max(np.abs(AN @ SN - np.eye(N * M))) <= 64 * EPS * N * M


@pytest.mark.parametrize("S", [0, 1, -2])
def test_spin_synthesis_of_unit_coefficients(S):
    """fftw.c:800-828: a unit coefficient in row l of column 2q (2q-1) gives
    a w {cos(l theta), sin((l+1) theta)} e^{+i q phi} (e^{-i q phi}) with
    a = 1 / (2 sqrt(2 pi)), the sphere's column classes for even S and the
    vector field's for odd S."""
    N, M = 4, 7
    cos = O._grid_classes(M, S % 2 != 0)
    theta = (2 * np.arange(N) + 1) * np.pi / (2 * N)
    phi = 2 * np.pi * np.arange(M) / M
    a = 1.0 / (2.0 * np.sqrt(2.0 * np.pi))
    for l in range(N):
        for col in range(M):
            X = np.zeros((N, M), dtype=np.complex128)
            X[l, col] = 1.0
            G = O.spin_grid(False, "N", True, S, X)
            t = 2 * np.cos(l * theta) if cos[col] else (2.0 if l < N - 1 else 1.0) * np.sin((l + 1) * theta)
            q = (col + 1) // 2
            p = np.exp((1j if col % 2 == 0 else -1j) * q * phi)
            assert np.max(np.abs(G - a * np.outer(t, p))) <= 64 * (N + M) * EPS, (l, col)


@pytest.mark.parametrize("S", [0, 3])
def test_spin_adjoints_and_inverse(S):
    N, M = 3, 5

    def dense(analysis, trans, syn):
        D = np.zeros((N * M, N * M), dtype=np.complex128)
        for i in range(N * M):
            e = np.zeros(N * M, dtype=np.complex128)
            e[i] = 1.0
            D[:, i] = O.spin_grid(analysis, trans, syn, S, e.reshape((N, M), order="F")).reshape(-1, order="F")
        return D
    SN, SC = dense(False, "N", True), dense(False, "C", False)
    AN, AC = dense(True, "N", False), dense(True, "C", True)
    assert np.max(np.abs(SC - SN.conj().T)) <= 64 * EPS * N * M
    assert np.max(np.abs(AC - AN.conj().T)) <= 64 * EPS * N * M
    assert np.max(np.abs(AN @ SN - np.eye(N * M))) <= 64 * EPS * N * M
