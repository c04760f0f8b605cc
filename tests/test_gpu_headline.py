"""Parity at the sizes the benchmark runs, through the C ABI: the 5,120-triangle bumpy ellipsoid
(both LU precisions) and all eight meshes of the 1/N series in the captured graph the benchmark
times, each against the oracle's dense FP64 path (LAPACK); the FP32 path's fallbacks (row
exchanges after a large multiplier, after a refinement that does not converge, and the series'
re-solve), each forced by an option and checked against the oracle; the transport algebra
against what the mathematics fixes (tests/transport_checks.py); the device algebra (k_transport)
against the same algebra on the host, bit for bit, and in the captured graph with the centre of
viscosity against the oracle.

Bars: D_t, D_r, CHI*V and the JACOBI eigenvalues within 1e-10 relative; the coupling tensor Dtr
within 1e-10 of its largest entry (of 1e-3 sqrt(|Dtt| |Drr|) where it vanishes, as on a sphere); the centre of resistance within 1e-10 of the body's size; the
deflated GR within 1e-12 of each 3x3 block's largest entry of the oracle's (the prep's running
sums are the reference's bit for bit: nearly aligned pairs amplify the last bits of the shifted
coordinates); the 11 solution vectors with normwise backward errors below 1e-14 against the
GPU's matrix and within 1e-9 of LAPACK's (relative to each column's largest entry; two
backward-stable solves differ by about cond(GR) u at this size).
"""
import numpy as np
import pytest
import transport_checks as TC
from conftest import rel, no_torch_loaded

pytestmark = pytest.mark.gpu

TOL = 1e-10


def _compare(r, ref, size):
    errs = {k: rel(r[k], ref[k]) for k in ("dtrans", "dr", "chiprime")}
    # the coupling tensor relative to its largest entry, or to 1e-3 of its natural scale
    # sqrt(|Dtt| |Drr|) for bodies where it vanishes (the sphere: discretization noise only)
    scale = np.sqrt(np.max(np.abs(ref["dtt"])) * np.max(np.abs(ref["drr"])))
    errs["dtr"] = float(np.max(np.abs(r["dtr"] - ref["dtr"])) / max(np.max(np.abs(ref["dtr"])), 1e-3 * scale))
    errs["urc"] = float(np.max(np.abs(r["urc"] - ref["urc"])) / size)
    for k in ("dtt_eig", "drr_eig"):
        a, b = np.sort(r[k]), np.sort(ref[k])
        errs[k] = float(np.max(np.abs(a - b)) / np.max(np.abs(b)))
    return errs


def _size(t):
    return float(max(np.max(np.abs(t.v1)), np.max(np.abs(t.v2)), np.max(np.abs(t.v3))))


@pytest.fixture(scope="module")
def c2(bh, oracle):
    from best_hydro_amd import meshes as M
    t = M.bumpy_ellipsoid_mesh()
    print(f"bumpy ellipsoid: {t.nar} triangles; oracle run", flush=True)
    ref = oracle.best_run(t.v1, t.v2, t.v3, t.truearea, t.truevol, iarea=1, cg=1.0, ivisco=1, icov=0)
    # the oracle's deflated GR and the 11 right-hand sides, for the backward errors
    m = oracle.prep(t.v1, t.v2, t.v3, 1, t.truearea)
    x, gr = oracle.tgetgr(m)
    oracle.deflate(m, gr, np.ones(m["nar"], dtype=np.int32), 1)
    b = np.concatenate([oracle.difco_rhs(m["nar"], x), oracle.strain_rhs(m["nar"], x)], axis=1)
    return t, ref, gr, b


def _backward_errors(gr, x, b):
    """Normwise backward error of each column: |b - GR x| / (|GR| |x| + |b|), infinity norms."""
    r = b - gr @ x
    na = float(np.max(np.sum(np.abs(gr), axis=1)))
    return np.max(np.abs(r), axis=0) / (na * np.max(np.abs(x), axis=0) + np.max(np.abs(b), axis=0))


@pytest.mark.parametrize("prec", [32, 64])
def test_c2_against_oracle(bh, oracle, c2, prec):
    t, ref, gr, rhs = c2
    s = bh.Solver(0)
    s.set_mesh(t)
    print(f"C2 precision {prec}", flush=True)
    r = s.solve_mesh(ivisco=1, precision=prec)
    st = s.stats()
    got = s.solution(11)
    _, gr_gpu = s.tgetgr()  # the GPU's own GR (no deflation), after the solve
    s.close()
    errs = _compare(r, ref, _size(t))
    print("  errors against the oracle: " + ", ".join(f"{k} {v:.1e}" for k, v in errs.items()), flush=True)
    for k, v in errs.items():
        assert v < TOL, k
    assert st["pivoted"] == 0 and st["ir_above_tol"] == 0
    # The 11 solutions: backward errors against the oracle's matrix at the FP64 level (as small as
    # LAPACK's own), and the difference from LAPACK's solutions, which the conditioning of GR
    # bounds (two backward-stable FP64 solves differ by about cond(GR) u: 3e-10 here)
    want = np.concatenate([ref["sol_rigid"], ref["sol_strain"]], axis=1)
    gr_gpu = np.asfortranarray(gr_gpu)
    def _block(a, b):  # the largest difference relative to its 3x3 block's largest entry
        n3 = b.shape[0]
        bb = np.abs(b).reshape(n3 // 3, 3, n3 // 3, 3).max(axis=(1, 3))
        return float(np.max(np.abs(a - b) / np.repeat(np.repeat(bb, 3, axis=0), 3, axis=1)))
    mm = oracle.prep(t.v1, t.v2, t.v3, 1, t.truearea)
    oracle.deflate(mm, gr_gpu, np.ones(mm["nar"], dtype=np.int32), 1)
    d_entries = _block(gr_gpu, gr)
    eta_own = _backward_errors(gr_gpu, got, rhs)
    eta, eta_l = _backward_errors(gr, got, rhs), _backward_errors(gr, want, rhs)
    fwd = np.max(np.abs(got - want), axis=0) / np.max(np.abs(want), axis=0)
    print(f"  backward error against the GPU's GR {eta_own.max():.1e}, against the oracle's {eta.max():.1e} "
          f"(LAPACK {eta_l.max():.1e}); GR entries within {d_entries:.1e} of their block; difference from LAPACK "
          f"{fwd.max():.1e}", flush=True)
    assert np.all(eta_own < 1e-14)
    assert d_entries < 1e-12
    assert np.all(fwd < 1e-9)
    # the algebra on the GPU's own tensors
    mob = TC.mobility_errors(r)
    assert mob["dtr"] < 1e-12 and mob["drr"] < 1e-12 and mob["dtt"] < mob["dtt_bound"]
    e_cr, b_cr, e_cd, b_cd = TC.centre_errors(r)
    assert e_cr < b_cr and e_cd < b_cd
    for k, v in TC.eigen_errors(r).items():
        assert v < 1e-12, k


@pytest.fixture(scope="module")
def series(bh):
    """The series exactly as the benchmark runs it: every mesh resident, autotuned, one launch of
    the captured graph (a warm launch first)."""
    no_torch_loaded()
    from best_hydro_amd import meshes as M
    tris = M.series_meshes()
    print("series sizes", [t.nar for t in tris], flush=True)
    ser = bh.Series(tris)
    tune = ser.autotune(ivisco=1)
    print("tuning", tune, flush=True)
    ser.solve(ivisco=1)
    res, ms = ser.solve(ivisco=1)
    print(f"graph launch {ms:.1f} ms", flush=True)
    stats = [ser.stats(k) for k in range(len(tris))]
    ser.close()
    return tris, res, stats, tune


@pytest.mark.parametrize("k", range(8))
def test_series_mesh_against_oracle(bh, oracle, series, k):
    tris, res, stats, tune = series
    t = tris[k]
    print(f"series mesh {k}: {t.nar} triangles, tuning {tune[k]}, oracle run", flush=True)
    ref = oracle.best_run(t.v1, t.v2, t.v3, t.truearea, t.truevol, iarea=1, cg=1.0, ivisco=1, icov=0)
    errs = _compare(res[k], ref, _size(t))
    print("  errors against the oracle: " + ", ".join(f"{k2} {v:.1e}" for k2, v in errs.items()), flush=True)
    for key, v in errs.items():
        assert v < TOL, key
    assert stats[k]["ir_above_tol"] == 0


@pytest.fixture(scope="module")
def delaunay(bh, oracle):
    """The 2,500-triangle random-Delaunay mesh: partial pivoting exchanges rows there, and the
    exchange-free FP32 factorization meets multipliers up to about 1.6."""
    from best_hydro_amd import meshes as M
    t = M.series_meshes(sizes=(2500,))[0]
    ref = oracle.best_run(t.v1, t.v2, t.v3, t.truearea, t.truevol, iarea=1, cg=1.0, ivisco=1, icov=0)
    return t, ref


def _check_oracle(r, ref, t):
    for key, v in _compare(r, ref, _size(t)).items():
        assert v < TOL, key


def test_fp32_growth_fallback_to_row_exchanges(bh, delaunay):
    """A growth threshold of 1 makes the FP32 factorization meet a multiplier above it: the matrix
    is refactored with row exchanges (lu_pivoted<float>) and refined against the FP64 matrix."""
    t, ref = delaunay
    s = bh.Solver(0)
    s.set_mesh(t)
    s.set_option("fp32_growth", 1.0)
    r = s.solve_mesh(ivisco=1, precision=32)
    st = s.stats()
    log = s.log()
    s.close()
    print("stats", st, flush=True)
    assert st["pivoted"] == 1
    assert any("multiplier exceeded" in l for l in log)
    _check_oracle(r, ref, t)


def test_refinement_fallback_to_row_exchanges(bh, delaunay):
    """One refinement step allowed with the exchange-free factors: the refinement cannot have
    converged, so the FP32 matrix is refactored with row exchanges (refactor_pivoted32) and the
    refinement runs again with those factors."""
    t, ref = delaunay
    s = bh.Solver(0)
    s.set_mesh(t)
    s.set_option("ir_steps_before_pivoting", 1)
    r = s.solve_mesh(ivisco=1, precision=32)
    st = s.stats()
    log = s.log()
    s.close()
    print("stats", st, flush=True)
    assert st["pivoted"] == 1
    assert any("refinement did not converge with the exchange-free factors" in l for l in log)
    _check_oracle(r, ref, t)


def test_series_resolves_a_mesh_past_the_growth_threshold(bh, oracle, delaunay):
    """In the series, a mesh whose captured factorization met a multiplier above the threshold is
    solved again eagerly with row exchanges; the other mesh keeps its graph solve."""
    no_torch_loaded()
    from best_hydro_amd import meshes as M
    t, ref = delaunay
    sphere = M.sphere_mesh(3, 15.0)
    ser = bh.Series([t, sphere])
    ser.set_option("fp32_growth", 1.0)
    for k in range(2):
        ser.set_tuning(k, strip=64, gemm_bk=0, ir_iters=3)
    res, _ = ser.solve(ivisco=1)
    piv = [ser.stats(k)["pivoted"] for k in range(2)]
    ser.close()
    assert piv == [1, 0]
    _check_oracle(res[0], ref, t)
    ref_s = oracle.best_run(sphere.v1, sphere.v2, sphere.v3, sphere.truearea, sphere.truevol, ivisco=1)
    _check_oracle(res[1], ref_s, sphere)


def test_series_rejects_steps_beyond_the_limit(bh):
    from best_hydro_amd import meshes as M
    ser = bh.Series([M.sphere_mesh(2, 15.0)])
    with pytest.raises(bh.HipError):
        ser.set_tuning(0, strip=64, gemm_bk=0, ir_iters=30)
    ser.close()


def test_transport_algebra_on_gpu(bh, oracle):
    """On the two-surface example mesh (FP32 path): scaling every coordinate by 2 (no area
    correction) scales every output exactly; the cyclic rotation x -> y -> z -> x rotates the
    tensors; the algebra's invariants hold on the GPU's tensors."""
    import os
    from conftest import GOLDEN
    from best_hydro_amd import meshes as M
    t = M.read_best(os.path.join(GOLDEN, "1aki49_atom_01806"))
    s = bh.Solver(0)
    s.set_mesh(t)
    r1 = s.solve_mesh(iarea=0, ivisco=1, precision=32)
    t2 = M.Triangulation(t.v1 * 2.0, t.v2 * 2.0, t.v3 * 2.0, 0.0, 0.0, name="scaled")
    s.set_mesh(t2)
    r2 = s.solve_mesh(iarea=0, ivisco=1, precision=32)
    sc = TC.scaling_errors(r1, r2, 2.0)
    print("scaling by 2:", {k: f"{v:.1e}" for k, v in sc.items()}, flush=True)
    for k, v in sc.items():
        assert v <= 1e-13, k
    s.set_mesh(t)
    r = s.solve_mesh(ivisco=1, precision=32)
    tc = M.Triangulation(TC.cyclic(t.v1), TC.cyclic(t.v2), TC.cyclic(t.v3), t.truearea, t.truevol, name="rotated")
    s.set_mesh(tc)
    rr = s.solve_mesh(ivisco=1, precision=32)
    s.close()
    cov = TC.covariance_errors(r, rr)
    print("cyclic rotation:", {k: f"{v:.1e}" for k, v in cov.items()}, flush=True)
    for k, v in cov.items():
        assert v < 1e-8, k
    mob = TC.mobility_errors(r)
    assert mob["dtr"] < 1e-12 and mob["drr"] < 1e-12 and mob["dtt"] < mob["dtt_bound"]
    e_cr, b_cr, e_cd, b_cd = TC.centre_errors(r)
    assert e_cr < b_cr and e_cd < b_cd


_KEYS = ("dtt", "dtr", "drr", "urc", "udr", "ktt", "krt", "krr", "drr_eig", "dtt_eig", "dr", "dtrans",
         "chiprime", "chi", "dvisc")


@pytest.mark.parametrize("prec", [32, 64])
def test_device_algebra_matches_host(bh, prec):
    """The reductions summed and BEST's algebra run on the device (k_transport), and the same on
    the host (host_algebra 1) from the same solve: the same bits, with and without the move to
    the centre of viscosity (the second reduction reads the centre from device memory)."""
    import os
    from conftest import GOLDEN
    from best_hydro_amd import meshes as M
    t = M.read_best(os.path.join(GOLDEN, "1aki49_atom_01806"))
    s = bh.Solver(0)
    s.set_mesh(t)
    for icov in (0, 1):
        print(f"precision {prec}, icov {icov}", flush=True)
        rd = s.solve_mesh(ivisco=1, icov=icov, precision=prec)
        s.set_option("host_algebra", 1)
        rh = s.solve_mesh(ivisco=1, icov=icov, precision=prec)
        s.set_option("host_algebra", 0)
        for k in _KEYS:
            assert np.array_equal(np.asarray(rd[k]), np.asarray(rh[k])), (icov, k)
        assert (np.any(np.asarray(rd["dvisc"]) != 0.0)) == bool(icov)
    s.close()


def test_series_graph_centre_of_viscosity(bh, oracle):
    """icov 1 in the captured graph (the centre computed and the second reduction run on the
    device, no host step between): against the oracle's run with the centre of viscosity."""
    no_torch_loaded()
    import os
    from conftest import GOLDEN
    from best_hydro_amd import meshes as M
    t = M.read_best(os.path.join(GOLDEN, "1aki49_atom_01806"))
    sphere = M.sphere_mesh(3, 15.0)
    ser = bh.Series([t, sphere])
    for k in range(2):
        ser.set_tuning(k, strip=64, gemm_bk=0, ir_iters=4)
    ser.solve(ivisco=1, icov=1)
    res, _ = ser.solve(ivisco=1, icov=1)
    ser.close()
    for r, m in zip(res, (t, sphere)):
        ref = oracle.best_run(m.v1, m.v2, m.v3, m.truearea, m.truevol, iarea=1, cg=1.0, ivisco=1, icov=1)
        _check_oracle(r, ref, m)
        d = float(np.max(np.abs(np.asarray(r["dvisc"]) - np.asarray(ref["dvisc"]))))
        print(f"{m.nar} triangles: centre of viscosity within {d:.1e} A of the oracle's", flush=True)
        assert d < 1e-10 * _size(m)
