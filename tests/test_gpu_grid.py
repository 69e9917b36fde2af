"""GPU synthesis and analysis on the sphere's grid (ft_execute_{sph,sphv}_
{synthesis,analysis}, fftw.c:126-270) against the oracle (oracle.grid,
pinned on the CPU): the FFTs sum in another order than the oracle's, so the
bar is the relative 2-norm 2e-14 (measured ~1e-16 to 1e-15).  The whole
spherical harmonic transform on the device -- sph2fourier, synthesis,
analysis, fourier2sph -- returns its input within the reference's bounds
(test_fftw.c:44-67: 4 sqrt(N) eps rel-2, 2N eps rel-inf).
"""
import sys

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

EPS = 2.0 ** -52
SIZES = [(1, 1), (2, 3), (5, 9), (8, 5), (64, 127), (65, 129), (256, 511), (1024, 2047)]
CASES = [(w, t) for w in ("sph_synthesis", "sph_analysis", "sphv_synthesis", "sphv_analysis", "disk_synthesis",
                          "disk_analysis") for t in ("N", "T")]


def _say(*args):
    print(*args, flush=True)
    sys.stdout.flush()


def _plan(ft, which, trans, N, M):
    """The plan the executor takes for TRANS: synthesis 'N' and analysis 'T'
    a synthesis plan, the others an analysis plan (test_fftw.c:52-64)."""
    geo = which.split("_")[0]
    syn = (which.endswith("synthesis")) == (trans == "N")
    name = geo + "_" + ("synthesis" if syn else "analysis")
    return getattr(ft, "plan_" + name)(N, M), syn, geo == "sphv"


@pytest.mark.parametrize("which,trans", CASES)
def test_against_oracle(ft, which, trans):
    for N, M in SIZES:
        P, syn, sphv = _plan(ft, which, trans, N, M)
        X = np.asfortranarray(np.random.default_rng(N * 7 + M).standard_normal((N, M)))
        _say("grid", which, trans, "N=%d M=%d" % (N, M))
        got = getattr(ft, "execute_" + which)(trans, P, X.copy(order="F"))
        ref = O.grid(which, trans, syn, sphv, X)
        assert O.rel2(got, ref) <= 2e-14, (which, trans, N, M, O.rel2(got, ref))


def test_device_resident_and_wrong_plans(ft):
    import torch
    N, M = 300, 599
    S, A = ft.plan_sph_synthesis(N, M), ft.plan_sph_analysis(N, M)
    X = np.asfortranarray(np.random.default_rng(3).standard_normal((N, M)))
    T = torch.from_numpy(np.ascontiguousarray(X.T)).cuda()
    _say("gpu_execute_sph_grid N=%d M=%d" % (N, M))
    ft.gpu_execute_sph_grid("sph_synthesis", "N", S, T)
    ft.gpu_execute_sph_grid("sph_analysis", "N", A, T)
    ft.synchronize()
    assert O.rel2(T.cpu().numpy().T, X) <= 1e-14
    with pytest.raises(ft.GPUError):
        ft.execute_sph_synthesis("N", A, X.copy(order="F"))   # an analysis plan
    with pytest.raises(ft.GPUError):
        ft.execute_sph_analysis("N", S, X.copy(order="F"))    # a synthesis plan


@pytest.mark.parametrize("which", ["sph_synthesis", "sph_analysis", "sphv_synthesis", "annulus_analysis",
                                   "disk_synthesis", "disk_analysis"])
@pytest.mark.parametrize("trans", ["N", "T"])
def test_paired_and_real_phi_dfts_agree(ft, which, trans):
    """With N even the phi DFTs of rows 2p and 2p+1 run as one complex DFT
    over X read as complex; a device array that is not 16-byte aligned takes
    the real DFTs instead.  Both against the oracle."""
    import torch
    N, M = 64, 127
    geo, kind = which.split("_")
    syn = (kind == "synthesis") == (trans == "N")
    if geo == "annulus":
        rho = 0.5
        P = (ft.plan_annulus_synthesis if syn else ft.plan_annulus_analysis)(N, M, rho)
    else:
        rho = None
        P = getattr(ft, "plan_%s_%s" % (geo, "synthesis" if syn else "analysis"))(N, M)
    X = np.asfortranarray(np.random.default_rng(77).standard_normal((N, M)))
    if geo == "annulus":
        ref = O.grid(which, trans, syn, False, X, rho=rho)
    else:
        ref = O.grid(which, trans, syn, geo == "sphv", X)
    flat = np.ascontiguousarray(X.T).reshape(-1)
    T1 = torch.from_numpy(flat.copy()).cuda().view(M, N)
    buf = torch.zeros(N * M + 1, dtype=torch.float64, device="cuda")
    T2 = buf[1:].view(M, N)
    T2.copy_(T1)
    assert T1.data_ptr() % 16 == 0 and T2.data_ptr() % 16 == 8
    _say("grid", which, trans, "aligned and unaligned X, N=%d M=%d" % (N, M))
    ft.gpu_execute_sph_grid(which, trans, P, T1)
    ft.gpu_execute_sph_grid(which, trans, P, T2)
    ft.synchronize()
    g1, g2 = T1.cpu().numpy().T, T2.cpu().numpy().T
    assert O.rel2(g1, ref) <= 2e-14 and O.rel2(g2, ref) <= 2e-14, (O.rel2(g1, ref), O.rel2(g2, ref))


@pytest.mark.parametrize("n", [64, 1024])
def test_spherical_harmonic_transform_round_trip(ft, n):
    """test_fftw.c:44-67 on the device: sph2fourier, synthesis, analysis,
    fourier2sph ('N'), then the same with 'T', against the reference's bounds."""
    import torch
    M = 2 * n - 1
    H = ft.plan_sph2fourier(n)
    PS, PA = ft.plan_sph_synthesis(n, M), ft.plan_sph_analysis(n, M)
    A = O.fill("sph", n, M, 99)
    T = torch.from_numpy(np.ascontiguousarray(A.T)).cuda()
    _say("SHT round trip n=%d" % n)
    for trans, syn, ana in (("N", PS, PA), ("T", PA, PS)):
        ft.gpu_execute_harmonic("sph2fourier", trans, H, T)
        ft.gpu_execute_sph_grid("sph_synthesis", trans, syn, T)
        ft.gpu_execute_sph_grid("sph_analysis", trans, ana, T)
        ft.gpu_execute_harmonic("fourier2sph", trans, H, T)
    ft.synchronize()
    X = T.cpu().numpy().T
    e2, einf = O.rel2(X, A), O.relinf(X, A)
    _say("SHT n=%d round trip rel2 %.2e relinf %.2e" % (n, e2, einf))
    assert e2 < 4 * np.sqrt(n) * EPS and einf < 2 * n * EPS


@pytest.mark.parametrize("analysis", [False, True])
@pytest.mark.parametrize("trans", ["N", "T"])
def test_tri_against_oracle(ft, analysis, trans):
    for N, M in [(1, 4), (2, 2), (5, 5), (5, 8), (64, 64), (65, 65), (300, 300), (1024, 1024)]:
        for syn_plan in (True, False):
            P = (ft.plan_tri_synthesis if syn_plan else ft.plan_tri_analysis)(N, M)
            X = np.asfortranarray(np.random.default_rng(N + 3 * M).standard_normal((N, M)))
            name = "tri_analysis" if analysis else "tri_synthesis"
            _say("grid", name, trans, "synthesis plan" if syn_plan else "analysis plan", "N=%d M=%d" % (N, M))
            got = getattr(ft, "execute_" + name)(trans, P, X.copy(order="F"))
            ref = O.tri_grid(analysis, trans, syn_plan, X)
            assert O.rel2(got, ref) <= 2e-14, (name, trans, syn_plan, N, M, O.rel2(got, ref))


@pytest.mark.parametrize("n", [64, 1024])
def test_triangle_transform_round_trip(ft, n):
    """test_fftw.c:164-196 on the device: tri2cheb, synthesis, analysis,
    cheb2tri with 'N' and with 'T'.  At n = 64 (the reference's test size)
    both chains in a row against its bounds (32 sqrt(N) eps rel-2, 8N eps
    rel-inf).  At n = 1024 the chains' conditioning sets the error: the
    oracle's own chains (same input, connection products in long double)
    return it within 1.9e-14 ('N') and 2.9e-13 ('T') rel-2; here the products
    are FP64, as the reference's BLAS products are (measured 4.6e-14 and
    1.1e-12), so each chain is checked on its own against 1e-13 and 4e-12."""
    import torch
    H = ft.plan_tri2cheb(n, 0.0, 0.0, 0.0)
    QS, QA = ft.plan_tri_synthesis(n, n), ft.plan_tri_analysis(n, n)
    A = O.fill("tri", n, n, 41)

    def chain(T, trans):
        syn, ana = (QS, QA) if trans == "N" else (QA, QS)
        ft.gpu_execute_harmonic("tri2cheb", trans, H, T)
        ft.gpu_execute_tri_grid("tri_synthesis", trans, syn, T)
        ft.gpu_execute_tri_grid("tri_analysis", trans, ana, T)
        ft.gpu_execute_harmonic("cheb2tri", trans, H, T)

    _say("triangle transform round trip n=%d" % n)
    if n <= 64:
        T = torch.from_numpy(np.ascontiguousarray(A.T)).cuda()
        chain(T, "N")
        chain(T, "T")
        ft.synchronize()
        X = T.cpu().numpy().T
        e2, einf = O.rel2(X, A), O.relinf(X, A)
        _say("triangle n=%d round trip rel2 %.2e relinf %.2e" % (n, e2, einf))
        assert e2 < 32 * np.sqrt(n) * EPS and einf < 8 * n * EPS
        return
    for trans, bound in (("N", 1e-13), ("T", 4e-12)):
        T = torch.from_numpy(np.ascontiguousarray(A.T)).cuda()
        chain(T, trans)
        ft.synchronize()
        X = T.cpu().numpy().T
        _say("triangle n=%d %s chain rel2 %.2e relinf %.2e" % (n, trans, O.rel2(X, A), O.relinf(X, A)))
        assert O.rel2(X, A) < bound


@pytest.mark.parametrize("n", [64, 1024])
def test_disk_transform_round_trip(ft, n):
    """test_fftw.c:226-260 on the device: disk2cxf, synthesis, analysis,
    cxf2disk with 'N', then with 'T', against the reference's bounds
    (8 sqrt(N) eps rel-2, 4N eps rel-inf)."""
    import torch
    M = 4 * n - 3
    H = ft.plan_disk2cxf(n, 0.0, 0.0)
    RS, RA = ft.plan_disk_synthesis(n, M), ft.plan_disk_analysis(n, M)
    A = O.fill("disk", n, M, 43)
    T = torch.from_numpy(np.ascontiguousarray(A.T)).cuda()
    _say("disk transform round trip n=%d" % n)
    for trans, syn, ana in (("N", RS, RA), ("T", RA, RS)):
        ft.gpu_execute_harmonic("disk2cxf", trans, H, T)
        ft.gpu_execute_sph_grid("disk_synthesis", trans, syn, T)
        ft.gpu_execute_sph_grid("disk_analysis", trans, ana, T)
        ft.gpu_execute_harmonic("cxf2disk", trans, H, T)
    ft.synchronize()
    X = T.cpu().numpy().T
    e2, einf = O.rel2(X, A), O.relinf(X, A)
    _say("disk n=%d round trip rel2 %.2e relinf %.2e" % (n, e2, einf))
    assert e2 < 8 * np.sqrt(n) * EPS and einf < 4 * n * EPS


@pytest.mark.parametrize("analysis", [False, True])
@pytest.mark.parametrize("trans", ["N", "T"])
def test_tet_against_oracle(ft, analysis, trans):
    for N, L, M in [(1, 3, 3), (2, 2, 2), (5, 5, 5), (5, 3, 4), (64, 64, 64), (128, 128, 128)]:
        for syn_plan in (True, False):
            P = (ft.plan_tet_synthesis if syn_plan else ft.plan_tet_analysis)(N, L, M)
            X = np.asfortranarray(np.random.default_rng(N + 7 * L + M).standard_normal((N, L, M)))
            name = "tet_analysis" if analysis else "tet_synthesis"
            _say("grid", name, trans, "synthesis plan" if syn_plan else "analysis plan", "N=%d L=%d M=%d" % (N, L, M))
            got = getattr(ft, "execute_" + name)(trans, P, X.copy(order="F"))
            ref = O.tet_grid(analysis, trans, syn_plan, X)
            assert O.rel2(got, ref) <= 2e-14, (name, trans, syn_plan, N, L, M, O.rel2(got, ref))


@pytest.mark.parametrize("which,trans", [(w, t) for w in ("annulus_synthesis", "annulus_analysis") for t in ("N", "T")])
def test_annulus_against_oracle(ft, which, trans):
    for N, M in [(1, 1), (5, 9), (64, 253), (256, 1021), (1024, 4093)]:
        for rho in (0.0, float(np.sqrt(1.0 / 3.0))):
            syn = which.endswith("synthesis") == (trans == "N")
            P = (ft.plan_annulus_synthesis if syn else ft.plan_annulus_analysis)(N, M, rho)
            X = np.asfortranarray(np.random.default_rng(N + M).standard_normal((N, M)))
            _say("grid", which, trans, "rho=%.3f N=%d M=%d" % (rho, N, M))
            got = getattr(ft, "execute_" + which)(trans, P, X.copy(order="F"))
            ref = O.grid(which, trans, syn, False, X, rho=rho)
            assert O.rel2(got, ref) <= 2e-14, (which, trans, rho, N, M, O.rel2(got, ref))


@pytest.mark.parametrize("analysis", [False, True])
@pytest.mark.parametrize("trans", ["N", "T"])
def test_rectdisk_against_oracle(ft, analysis, trans):
    for N, M in [(1, 1), (3, 4), (5, 5), (64, 64), (300, 300), (1024, 1024)]:
        for syn_plan in (True, False):
            P = (ft.plan_rectdisk_synthesis if syn_plan else ft.plan_rectdisk_analysis)(N, M)
            X = np.asfortranarray(np.random.default_rng(N + 5 * M).standard_normal((N, M)))
            name = "rectdisk_analysis" if analysis else "rectdisk_synthesis"
            _say("grid", name, trans, "synthesis plan" if syn_plan else "analysis plan", "N=%d M=%d" % (N, M))
            got = getattr(ft, "execute_" + name)(trans, P, X.copy(order="F"))
            ref = O.rect_grid(analysis, trans, syn_plan, X)
            assert O.rel2(got, ref) <= 2e-14, (name, trans, syn_plan, N, M, O.rel2(got, ref))


@pytest.mark.parametrize("n", [64, 1024])
def test_rectdisk_transform_round_trip(ft, n):
    """test_fftw.c:352-380 through the host-pointer executors: rectdisk2cheb,
    synthesis, analysis, cheb2rectdisk ('N', then 'T'), against the
    reference's bounds (8 sqrt(N) eps rel-2, 4N eps rel-inf)."""
    H = ft.plan_rectdisk2cheb(n, 0.0)
    SS, SA = ft.plan_rectdisk_synthesis(n, n), ft.plan_rectdisk_analysis(n, n)
    # rectdiskrand (ftutilities.c:123-130): U(-1, 1) where i + j < M
    A = np.random.default_rng(47).uniform(-1, 1, (n, n))
    A[np.add.outer(np.arange(n), np.arange(n)) >= n] = 0.0
    A = np.asfortranarray(A)
    X = A.copy(order="F")
    _say("rectdisk transform round trip n=%d" % n)
    for trans, syn, ana in (("N", SS, SA), ("T", SA, SS)):
        ft.execute_rectdisk2cheb(trans, H, X)
        ft.execute_rectdisk_synthesis(trans, syn, X)
        ft.execute_rectdisk_analysis(trans, ana, X)
        ft.execute_cheb2rectdisk(trans, H, X)
    e2, einf = O.rel2(X, A), O.relinf(X, A)
    _say("rectdisk n=%d round trip rel2 %.2e relinf %.2e" % (n, e2, einf))
    assert e2 < 8 * np.sqrt(n) * EPS and einf < 4 * n * EPS


@pytest.mark.parametrize("analysis", [False, True])
@pytest.mark.parametrize("trans", ["N", "C"])
def test_spinsph_grid_against_oracle(ft, analysis, trans):
    for N, M in [(1, 1), (4, 7), (65, 129), (256, 511), (1024, 2047)]:
        for S in (0, 1, -2):
            syn = (not analysis) == (trans == "N")
            P = (ft.plan_spinsph_synthesis if syn else ft.plan_spinsph_analysis)(N, M, S)
            g = np.random.default_rng(N + M + S + 9)
            X = np.asfortranarray(g.standard_normal((N, M)) + 1j * g.standard_normal((N, M)))
            name = "spinsph_analysis" if analysis else "spinsph_synthesis"
            _say("grid", name, trans, "S=%d N=%d M=%d" % (S, N, M))
            got = getattr(ft, "execute_" + name)(trans, P, X.copy(order="F"))
            ref = O.spin_grid(analysis, trans, syn, S, X)
            assert O.rel2(got, ref) <= 2e-14, (name, trans, S, N, M, O.rel2(got, ref))


@pytest.mark.parametrize("S", [0, 1, -2])
def test_spin_transform_round_trip_1024(ft, S):
    """The spin-weighted rotations and the grid on the device at n = 1024:
    hi2lo, synthesis, analysis, lo2hi returns the input within the sphere's
    bounds (4 sqrt(N) eps rel-2, 2N eps rel-inf)."""
    import torch
    n = 1024
    M = 2 * n - 1
    R = ft.plan_rotspinsphere(n, S)
    PS, PA = ft.plan_spinsph_synthesis(n, M, S), ft.plan_spinsph_analysis(n, M, S)
    g = np.random.default_rng(5 + S)
    A = np.asfortranarray(g.uniform(-1, 1, (n, M)) + 1j * g.uniform(-1, 1, (n, M)))
    T = torch.from_numpy(np.ascontiguousarray(A.T)).cuda()
    _say("spin transform round trip S=%d" % S)
    ft.gpu_execute_spinsph(R, "hi2lo", T)
    ft.synchronize()
    X = T.cpu().numpy().T.copy(order="F")
    ft.execute_spinsph_synthesis("N", PS, X)
    ft.execute_spinsph_analysis("N", PA, X)
    T = torch.from_numpy(np.ascontiguousarray(X.T)).cuda()
    ft.gpu_execute_spinsph(R, "lo2hi", T)
    ft.synchronize()
    X = T.cpu().numpy().T
    e2, einf = O.rel2(X, A), O.relinf(X, A)
    _say("spin S=%d n=1024 round trip rel2 %.2e relinf %.2e" % (S, e2, einf))
    assert e2 < 4 * np.sqrt(n) * EPS and einf < 2 * n * EPS
