"""Host logic without a GPU: the triangulations of the workloads, BEST's file format, COMPON, and
the C ABI of libbesthip.so (it loads and exports every function include/best_hip.h declares)."""
import math
import os
import re

import numpy as np
import pytest
from conftest import GOLDEN, ROOT


def test_sphere_header_and_orientation(bh):
    from best_hydro_amd import meshes as M
    t = M.sphere_mesh(3, 15.0)
    assert t.nar == 1280
    assert repr(t.truearea) == "2827.4333882308138" and repr(t.truevol) == "14137.166941154068"
    assert M.closed_check(t) == (0, 0, 0)
    assert t.signed_volume_best_normal() < 0  # the reference convention (its example meshes)
    # vertices on the sphere to the 8 digits of the file format
    r = np.linalg.norm(t.v1, axis=1)
    assert np.max(np.abs(r - 15.0)) < 1e-6


def test_reference_examples_share_the_orientation(bh):
    from best_hydro_amd import meshes as M
    for f in ("2eu7_02562", "1aki49_atom_01806"):
        t = M.read_best(os.path.join(GOLDEN, f))
        assert t.signed_volume_best_normal() < 0


def test_format_roundtrip_is_exact(bh):
    from best_hydro_amd import meshes as M
    t = M.read_best(os.path.join(GOLDEN, "2eu7_02562"))
    t2 = M.parse_best(M.format_best(t))
    assert np.array_equal(t.v1, t2.v1) and np.array_equal(t.v3, t2.v3)
    line = M.format_best(t).splitlines()[1]
    assert len(line) == 9 * 17 and line.count("E") == 9


def test_fortran_e_format(bh):
    from best_hydro_amd.meshes import _fortran_e
    assert _fortran_e(16.902591) == "   0.16902591E+02"
    assert _fortran_e(-0.0001234567891) == "  -0.12345679E-03"
    assert _fortran_e(0.0) == "   0.00000000E+00"


def test_exact_area_volume_ellipsoid(bh):
    from best_hydro_amd import meshes as M
    a, v = M.exact_area_volume(M.RadialShape((30.0, 20.0, 15.0)))
    assert abs(v - 4.0 / 3.0 * math.pi * 30 * 20 * 15) < 1e-9 * v
    # Knud Thomsen's approximation of the ellipsoid area is good to ~1%
    p = 1.6075
    approx = 4 * math.pi * (((30 * 20) ** p + (30 * 15) ** p + (20 * 15) ** p) / 3) ** (1 / p)
    assert abs(a - approx) < 0.012 * approx


def test_workload_meshes(bh):
    from best_hydro_amd import meshes as M
    c2 = M.bumpy_ellipsoid_mesh()
    assert c2.nar == 5120 and M.closed_check(c2) == (0, 0, 0) and c2.signed_volume_best_normal() < 0
    assert 0.99 < c2.truearea / c2.polyhedral_area() < 1.01
    two = M.two_surface_mesh()
    assert two.nar == 5200 and M.closed_check(two) == (0, 0, 0)
    ser = M.series_meshes(sizes=(2500, 3143))
    assert [t.nar for t in ser] == [2500, 3144]
    for t in ser:
        assert M.closed_check(t) == (0, 0, 0) and t.signed_volume_best_normal() < 0


def test_series_is_deterministic(bh):
    from best_hydro_amd import meshes as M
    a = M.series_meshes(sizes=(2500,))[0]
    b = M.series_meshes(sizes=(2500,))[0]
    assert np.array_equal(a.v1, b.v1)


def _declared_functions():
    text = open(os.path.join(ROOT, "include", "best_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*[A-Za-z_][\w\s\*]*?\b(bh[mg]_\w+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_library_exports_every_declared_symbol(bh):
    L = bh.lib()
    names = _declared_functions()
    assert len(names) >= 25
    for name in names:
        assert hasattr(L, name), name
    assert L.bhm_available() == 1


def test_hmatrix_abi_names_match_the_reference_header(bh):
    """The nine names of hmatrix/best_hmatrix.h (as best/hmstub.f90 defines them)."""
    want = {"bhm_available", "bhm_create", "bhm_incenters", "bhm_entries", "bhm_diag_mean", "bhm_build", "bhm_solve",
            "bhm_log_line", "bhm_destroy"}
    assert want <= set(_declared_functions())


def test_compon_in_the_library_matches_the_oracle(bh, oracle):
    from best_hydro_amd import meshes as M
    import ctypes as C
    t = M.two_surface_mesh(level=2)
    col = lambda v: np.ascontiguousarray(v.T).ravel()
    n = t.nar
    ic = np.zeros(n, dtype=np.int32)
    nc, no = C.c_int(), C.c_int()
    assert bh.lib().bhg_compon(n, col(t.v1), col(t.v2), col(t.v3), ic, C.byref(nc), C.byref(no)) == 0
    m = oracle.prep(t.v1, t.v2, t.v3, 1, t.truearea)
    ic2, nc2, no2 = oracle.compon(m)
    assert (nc.value, no.value) == (nc2, no2) == (2, 0)
    assert np.array_equal(ic, ic2)


def test_no_device_fails_loudly(bh):
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    with pytest.raises(bh.HipError):
        bh.Solver(0)
