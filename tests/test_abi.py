"""The C ABI without a GPU: the library loads, exports every declared symbol,
builds bit-exact plans on the host and rejects invalid arguments before any
device work."""
import os
import re

import numpy as np
import pytest

import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    with open(os.path.join(ROOT, "include", "fasttransforms_amd.h")) as f:
        text = re.sub(r"/\*.*?\*/", "", f.read(), flags=re.S)
    return sorted(set(re.findall(r"\b(ft_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_reference_entry_points():
    names = set(declared_functions())
    for kind in ("sph", "tri", "disk", "rectdisk"):
        for d in ("hi2lo", "lo2hi"):
            assert "ft_kernel_%s_%s" % (kind, d) in names
    for kind in ("sph", "sphv", "tri", "disk", "rectdisk"):
        for d in ("hi2lo", "lo2hi"):
            assert "ft_execute_%s_%s" % (kind, d) in names
    for fn in ("ft_plan_rotsphere", "ft_plan_rottriangle", "ft_plan_rotdisk", "ft_plan_rotrectdisk",
               "ft_destroy_rotation_plan", "ft_set_num_threads"):
        assert fn in names


def test_library_exports_every_declared_symbol(ft):
    import ctypes
    lib = ctypes.CDLL(ft.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    assert set(declared_functions()) <= set(ft.SIGNATURES)


@pytest.mark.parametrize("kind,params", [("sph", ()), ("tri", (0.0, 0.0, 0.0)), ("tri", (-0.5, -0.5, -0.5)),
                                         ("tri", (0.25, -0.5, 1.5)), ("disk", (0.0, 0.0)), ("disk", (-0.5, 2.0))])
@pytest.mark.parametrize("n", [1, 2, 3, 17, 256])
def test_product_plans_bitwise_equal_oracle(ft, kind, params, n):
    maker = {"sph": ft.plan_rotsphere, "tri": ft.plan_rottriangle, "disk": ft.plan_rotdisk}[kind]
    P = maker(n, *params)
    Q = O.Plan(kind, n, *params)
    assert P.n == n
    assert np.array_equal(P.s, Q.s) and np.array_equal(P.c, Q.c)


def test_invalid_arguments_rejected_without_device(ft):
    P = ft.plan_rotsphere(8)
    A = np.zeros((8, 16), order="F")  # even M is not a sph layout
    with pytest.raises(ft.GPUError):
        ft.execute_sph_hi2lo(P, A)
    A = np.zeros((8, 17), order="F")  # M > 2n-1
    with pytest.raises(ft.GPUError):
        ft.execute_sph_hi2lo(P, A)
    T = ft.plan_rottriangle(8, 0.0, 0.0, 0.0)
    with pytest.raises(ft.GPUError):
        ft.execute_tri_lo2hi(T, np.zeros((8, 9), order="F"))  # M > n
    with pytest.raises(ValueError):
        ft.execute_tri_hi2lo(P, np.zeros((8, 8), order="F"))  # wrong plan kind
    with pytest.raises(ft.GPUError):
        ft.plan_rotsphere(0)
    assert ft.last_error() == 0


def test_zero_columns_is_a_no_op(ft):
    P = ft.plan_rotsphere(8)
    A = np.ones((8, 1), order="F")
    ft.execute_sph_hi2lo(P, A, M=0)
    assert np.all(A == 1.0)


def test_rotation_count_matches_survey_table(ft):
    from fasttransforms_amd.utilities import rotations
    # exact rotation counts quoted for these sizes (sph n=256, 1024; tri and disk n=1024)
    assert rotations("sph", 256) == 5510784
    assert rotations("sph", 1024) == 356604416
    assert rotations("tri", 1024) == 357389824
    assert rotations("disk", 1024) == 1427466238
