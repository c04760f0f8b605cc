"""The series of meshes for the 1/N extrapolation (the bumpy ellipsoid at eight uneven N): one
captured HIP graph solves every mesh; each mesh's results equal its own single solve's and the
oracle's; the 1/N fit equals numpy's."""
import numpy as np
import pytest
from conftest import rel, no_torch_loaded


def test_extrapolation_matches_numpy(bh):
    rng = np.random.default_rng(3)
    n = np.array([2500, 3144, 3786, 4430, 5072, 5714, 6358, 7000])
    y = 1.7e-7 + 3.2e-4 / n + rng.normal(scale=1e-12, size=n.size)
    a, b = bh.extrapolate_1n(n, y)
    slope, icpt = np.polyfit(1.0 / n, y, 1)
    assert rel(a, icpt) < 1e-9 and rel(b, slope) < 1e-9
    # exact on exact data
    a, b = bh.extrapolate_1n(n, 2.0 + 5.0 / n)
    assert abs(a - 2.0) < 1e-12 and abs(b - 5.0) < 1e-9


@pytest.mark.gpu
def test_series_graph_matches_single_solves_and_oracle(bh, oracle):
    no_torch_loaded()
    from best_hydro_amd import meshes as M
    sizes = (2500, 3143, 3786)
    tris = M.series_meshes(sizes=sizes)
    print("series sizes", [len(t.v1) for t in tris], flush=True)
    ser = bh.Series(tris)
    tune = ser.autotune(ivisco=1)
    print("tuning", tune, flush=True)
    res, ms = ser.solve(ivisco=1)
    res2, ms2 = ser.solve(ivisco=1)
    print(f"graph {ms:.2f} ms, again {ms2:.2f} ms", flush=True)
    s = bh.Solver(0)
    for k, t in enumerate(tris):
        s.set_mesh(t)
        one = s.solve_mesh(ivisco=1, precision=32)
        for key in ("dtrans", "dr", "chiprime"):
            assert rel(res[k][key], one[key]) < 1e-12, (k, key)
            assert res2[k][key] == res[k][key]
        assert ser.stats(k)["pivoted"] == 0
    s.close()
    t = tris[0]
    ref = oracle.best_run(t.v1, t.v2, t.v3, t.truearea, t.truevol, iarea=1, cg=1.0, ivisco=1, icov=0)
    for key in ("dtrans", "dr", "chiprime"):
        assert rel(res[0][key], ref[key]) < 1e-10
    ser.close()


@pytest.mark.gpu
def test_series_fixed_tuning_and_icov(bh):
    """Hand-set tiles (no autotune) give the same answers; -vc needs the eager second pass."""
    no_torch_loaded()
    from best_hydro_amd import meshes as M
    tris = M.series_meshes(sizes=(2500, 3143))
    ser = bh.Series(tris)
    for k in range(2):
        ser.set_tuning(k, strip=32, gemm_bk=16, ir_iters=3)
    a, _ = ser.solve(ivisco=1, icov=1)
    for k in range(2):
        ser.set_tuning(k, strip=128, gemm_bk=64, ir_iters=5)
    b, _ = ser.solve(ivisco=1, icov=1)
    for k in range(2):
        for key in ("dtrans", "dr", "chiprime"):
            assert rel(a[k][key], b[k][key]) < 1e-12
        assert np.max(np.abs(a[k]["dvisc"] - b[k]["dvisc"])) < 1e-9
    ser.close()
