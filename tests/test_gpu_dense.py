"""Parity of the HIP path with the oracle, through the C ABI (include/best_hip.h).

Bars: incenters and the mean diagonal bit for bit; GR entries within 1e-12 of each 3x3 block's
largest entry (the device log/acos/sin/cos may differ from glibc's by an ulp), and most of them
bit for bit; D_t, D_r, CHI*V and the centre vectors within 1e-10 relative of the oracle's dense
FP64 path (LAPACK) for both the FP64 LU and the FP32 LU + FP64 refinement.
"""
import os

import numpy as np
import pytest
from conftest import GOLDEN, golden_expected, rel

pytestmark = pytest.mark.gpu

TOL = 1e-10


@pytest.fixture(scope="module")
def solver(bh):
    s = bh.Solver(0)
    yield s
    s.close()


def _oracle_run(oracle, path, **kw):
    v1, v2, v3, ta, tv = oracle.read_best(path)
    return oracle.best_run(v1, v2, v3, ta, tv, **kw)


def _block_scaled_error(a, b):
    """max |a - b| / max|b over the 3x3 block| over all entries."""
    n3 = a.shape[0]
    bb = np.abs(b).reshape(n3 // 3, 3, n3 // 3, 3).max(axis=(1, 3))
    scale = np.repeat(np.repeat(bb, 3, axis=0), 3, axis=1)
    return float(np.max(np.abs(a - b) / scale))


@pytest.mark.parametrize("fname", ["2eu7_02562", "1aki49_atom_01806"])
def test_entries_incenters_diag_mean(bh, oracle, solver, fname):
    v1, v2, v3, ta, tv = oracle.read_best(os.path.join(GOLDEN, fname))
    m = oracle.prep(v1, v2, v3, 1, ta)
    print(f"case {fname}: N={m['nar']}", flush=True)
    solver.load_prepared(m["nar"], m["v1"], m["v2"], m["v3"], m["area"], m["afac"], 1.0)
    x = solver.incenters()
    xo, sb = oracle.self_blocks(m)
    assert np.array_equal(x.view(np.int64), xo.view(np.int64))
    x2, gr = solver.tgetgr()
    xo2, gro = oracle.tgetgr(m)
    assert np.array_equal(x2.view(np.int64), xo.view(np.int64))
    err = _block_scaled_error(gr, gro)
    same = float(np.mean(gr.view(np.int64) == gro.view(np.int64)))
    print(f"  GR block-scaled max error {err:.2e}, bit-identical fraction {same:.4f}", flush=True)
    assert err < 1e-12
    assert same > 0.95
    _, diagm, _ = oracle.deflate(m, gro, np.ones(m["nar"], dtype=np.int32), 1)
    # the mean diagonal is the reference's running sum (best/best.f:386-403) of the device's own
    # diagonal entries, bit for bit; those entries may differ from glibc's by an ulp
    dsum = 0.0
    for i in range(0, 3 * m["nar"], 3):
        dsum = dsum + gr[i, i] + gr[i + 1, i + 1] + gr[i + 2, i + 2]
    assert solver.diag_mean() == dsum / (3 * m["nar"])
    assert rel(solver.diag_mean(), diagm) < 1e-14
    sub = solver.entries(7, 40, 3000, 33)
    assert np.array_equal(sub, gr[7:47, 3000:3033])


def test_entries_match_reference_kernel(bh, oracle, solver):
    """Against the reference's own kernel (oracle/_ref) where it was built and shipped."""
    R = oracle.reference_lib()
    if R is None:
        pytest.skip("oracle/_ref not present")
    from best_hydro_amd import meshes as M
    t = M.sphere_mesh(3, 15.0)
    m = oracle.prep(t.v1, t.v2, t.v3, 1, t.truearea)
    n = m["nar"]
    for cg in (1.0, 0.0):
        assert R.bhm_create(n, m["v1"], m["v2"], m["v3"], m["area"], m["afac"], cg, 4) == 0
        solver.load_prepared(n, m["v1"], m["v2"], m["v3"], m["area"], m["afac"], cg)
        n3 = 3 * n
        a = np.zeros(n3 * n3)
        R.bhm_entries(0, n3, 0, n3, a, n3)
        a = a.reshape(n3, n3).T
        _, gr = solver.tgetgr()
        assert _block_scaled_error(gr, a) < 1e-12
        xr = np.zeros(3 * n)
        R.bhm_incenters(xr)
        assert np.array_equal(solver.incenters().view(np.int64), xr.view(np.int64))
        assert rel(solver.diag_mean(), R.bhm_diag_mean()) < 1e-14


CASES = {
    "2eu7_v": dict(file="2eu7_02562", iarea=1, icov=0, wm=30000.0, golden="2eu7_v_mw30000"),
    "2eu7_vc": dict(file="2eu7_02562", iarea=1, icov=1, wm=30000.0, golden="2eu7_vc_mw30000"),
    "2eu7_na": dict(file="2eu7_02562", iarea=0, icov=0, wm=30000.0, golden="2eu7_na_v_mw30000"),
    "two_v": dict(file="1aki49_atom_01806", iarea=1, icov=0, wm=5000.0, golden="two_v_mw5000"),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_transport_parity_and_golden(bh, oracle, solver, case):
    c = CASES[case]
    path = os.path.join(GOLDEN, c["file"])
    ref = _oracle_run(oracle, path, iarea=c["iarea"], cg=1.0, ivisco=1, icov=c["icov"], wm=c["wm"])
    from best_hydro_amd import meshes as M
    solver.set_mesh(M.read_best(path))
    g = golden_expected(c["golden"])
    for prec in (64, 32):
        print(f"case {case} precision {prec}", flush=True)
        r = solver.solve_mesh(iarea=c["iarea"], cg=1.0, ivisco=1, icov=c["icov"], precision=prec, wm=c["wm"])
        st = solver.stats()
        assert st["pivoted"] == 0
        assert rel(r["dtrans"], ref["dtrans"]) < TOL
        assert rel(r["dr"], ref["dr"]) < TOL
        assert rel(r["chiprime"], ref["chiprime"]) < TOL
        assert np.max(np.abs(r["dtr"] - ref["dtr"])) < TOL * np.max(np.abs(ref["dtr"]))
        assert np.max(np.abs(r["urc"] - ref["urc"])) < 1e-8 * max(1.0, np.max(np.abs(ref["urc"])))
        if c["icov"]:
            assert np.max(np.abs(r["dvisc"] - ref["dvisc"])) < 1e-8
        got = dict(dtrans=r["dtrans"], dr=r["dr"], chi=r["chi"], eta=r["eta"])
        if c["icov"]:
            got.update(dx=r["dvisc"][0], dy=r["dvisc"][1], dz=r["dvisc"][2])
        for key, val in g.items():
            if key != "flags":
                assert rel(got[key], val[0]) <= val[1], (key, got[key], val[0])
        if case == "two_v":
            assert r["ncomp"] == 2 and r["ndef"] == 2


def test_solutions_match_lapack(bh, oracle, solver):
    """The 11 solution vectors (DIFCO's 6 flows, VISCO's 5) against LAPACK on the oracle's GR."""
    path = os.path.join(GOLDEN, "2eu7_02562")
    ref = _oracle_run(oracle, path, ivisco=1)
    from best_hydro_amd import meshes as M
    solver.set_mesh(M.read_best(path))
    want = np.concatenate([ref["sol_rigid"], ref["sol_strain"]], axis=1)
    for prec in (64, 32):
        solver.solve_mesh(precision=prec)
        got = solver.solution(11)
        for k in range(11):
            assert np.max(np.abs(got[:, k] - want[:, k])) < 1e-11 * np.max(np.abs(want[:, k]))


def test_sphere_c1_stokes_einstein(bh, oracle, solver):
    from best_hydro_amd import meshes as M
    t = M.sphere_mesh(3, 15.0)
    ref = oracle.best_run(t.v1, t.v2, t.v3, t.truearea, t.truevol, ivisco=1)
    solver.set_mesh(t)
    for prec in (64, 32):
        r = solver.solve_mesh(precision=prec)
        assert rel(r["dtrans"], ref["dtrans"]) < TOL and rel(r["dr"], ref["dr"]) < TOL
        assert rel(r["chi"], ref["chi"]) < TOL
        assert rel(r["dtrans"], ref["tunit"] * 4.0 / 45.0) < 5e-3   # kT / (6 pi eta a)
        assert rel(r["chi"], 2.5) < 1e-2


def test_preaveraged_option(bh, oracle, solver):
    """-p (cg = 0): the pre-averaged tensor and the 4/3, 3/4 factors."""
    path = os.path.join(GOLDEN, "1aki49_atom_01806")
    ref = _oracle_run(oracle, path, iarea=1, cg=0.0, ivisco=1, icov=0)
    from best_hydro_amd import meshes as M
    solver.set_mesh(M.read_best(path))
    r = solver.solve_mesh(cg=0.0, precision=32)
    assert rel(r["dtrans"], ref["dtrans"]) < TOL and rel(r["chiprime"], ref["chiprime"]) < TOL


def test_hmatrix_abi_sequence(bh, oracle):
    """HMPREP's calls (best/besolve.f90:164-233) and BESOLV's (:95-134) through the nine bhm_*
    functions: the solutions equal LAPACK's on the oracle's deflated GR."""
    hm = bh.hm
    path = os.path.join(GOLDEN, "1aki49_atom_01806")
    v1, v2, v3, ta, tv = oracle.read_best(path)
    m = oracle.prep(v1, v2, v3, 1, ta)
    n = m["nar"]
    icomp, ncomp, nopen = oracle.compon(m)
    assert hm.available() == 1
    assert hm.create(n, m["v1"], m["v2"], m["v3"], m["area"], m["afac"], 1.0, 0) == 0
    x = hm.incenters(n)
    xo, sb = oracle.self_blocks(m)
    assert np.array_equal(x.view(np.int64), xo.view(np.int64))
    head = hm.entries(0, 9, 0, 9)
    assert np.max(np.abs(head - oracle.entries(m, xo, sb, 0, 9, 0, 9))) < 1e-12 * np.max(np.abs(head))
    diagm = hm.diag_mean()
    xmu, gdef = oracle.deflation_weights(m, diagm, icomp, ncomp)
    for single in (0, 1):
        assert hm.build(1e-4, 0, 0.1, ncomp, gdef, icomp, xmu, single, 11) == 0
        b = oracle.difco_rhs(n, xo)
        rc, sol = hm.solve(b)
        assert rc == 0
        x_o, gr = oracle.tgetgr(m)
        oracle.deflate(m, gr, icomp, ncomp)
        lu, piv, info = oracle.lapack_getrf(gr)
        want = oracle.lapack_getrs(lu, piv, oracle.difco_rhs(n, xo))
        assert np.max(np.abs(sol - want)) < 1e-11 * np.max(np.abs(want))
    lines = hm.log_lines()
    assert any("dense GPU LU" in l for l in lines)
    hm.destroy()


def test_large_mesh_precision_agreement(bh, solver):
    """The 5,120-triangle bumpy ellipsoid: FP32 LU + FP64 refinement against FP64 LU (both on the
    GPU) -- the size-independent check at the benchmark's size."""
    from best_hydro_amd import meshes as M
    solver.set_mesh(M.bumpy_ellipsoid_mesh())
    r64 = solver.solve_mesh(precision=64)
    s64 = solver.solution(11)
    r32 = solver.solve_mesh(precision=32)
    s32 = solver.solution(11)
    assert rel(r32["dtrans"], r64["dtrans"]) < TOL
    assert rel(r32["dr"], r64["dr"]) < TOL
    assert rel(r32["chiprime"], r64["chiprime"]) < TOL
    assert np.max(np.abs(s32 - s64)) < 1e-10 * np.max(np.abs(s64))
    assert solver.stats()["ir_iterations"] <= 8


def test_row_exchanges_on_a_delaunay_mesh(bh, oracle, solver):
    """A random-point Delaunay mesh whose GR needs 6 row exchanges under partial pivoting (LAPACK's
    IPIV is not the identity): the FP64 LU takes the pivoted path and the FP32 LU accepts the
    multipliers (up to 16) that refinement corrects; both within 1e-10 of the oracle."""
    from best_hydro_amd import meshes as M
    t = M.series_meshes(sizes=(2500,))[0]
    ref = oracle.best_run(t.v1, t.v2, t.v3, t.truearea, t.truevol, iarea=1, cg=1.0, ivisco=1, icov=0)
    solver.set_mesh(t)
    r64 = solver.solve_mesh(precision=64)
    assert solver.stats()["pivoted"] == 1
    r32 = solver.solve_mesh(precision=32)
    assert solver.stats()["pivoted"] == 0
    for r in (r64, r32):
        for key in ("dtrans", "dr", "chiprime"):
            assert rel(r[key], ref[key]) < TOL, key


@pytest.mark.parametrize("case", ["2eu7_v_mw30000", "2eu7_vc_mw30000", "2eu7_na_v_mw30000", "two_v_mw5000"])
def test_be_file_of_the_gpu_solve(bh, oracle, solver, case):
    """The .be file of the GPU's solve (FP32 LU + FP64 refinement, the device algebra) for each
    case of the reference's expected_best_*.txt: the printed values carry the fixture's digits,
    and every byte equals the .be the oracle's FP64 run gives."""
    from conftest import golden_expected
    from test_befile import check_against_fixture
    from best_hydro_amd import meshes as M
    f = golden_expected(case)["flags"]
    t = M.read_best(os.path.join(GOLDEN, f["file"]))
    print(f"{case}: GPU solve and oracle run", flush=True)
    solver.set_mesh(t)
    r = solver.solve_mesh(iarea=f["iarea"], ivisco=1, icov=f["icov"], precision=32)
    text = bh.format_be(r, nar=t.nar, truearea=t.truearea, truevol=t.truevol, iarea=f["iarea"], mw=1, wm=f["wm"],
                        ivisco=1, icov=f["icov"], cg=1.0)
    check_against_fixture(oracle, text.decode(), case)
    ref = oracle.best_run(t.v1, t.v2, t.v3, t.truearea, t.truevol, iarea=f["iarea"], cg=1.0, ivisco=1,
                          icov=f["icov"], wm=f["wm"])
    want = oracle.format_be("4.0 SFSU/ARAGON", f["iarea"], t.truearea, ref["afac"], t.truevol, 1, f["wm"], ref["nar"],
                            ref["centroid"], ref["urc"], ref["udr"], ref["dtt"], ref["dtr"], ref["drr"], 1, f["icov"],
                            ref["dvisc"], ref["chiprime"], 1.0, ref["totarea"])
    diff = [(a, b) for a, b in zip(text.split(b"\n"), want.split(b"\n")) if a != b]
    assert not diff, diff[:3]
