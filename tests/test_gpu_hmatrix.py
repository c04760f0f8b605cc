"""The H-matrix path on the GPU (cross-approximated blocks, the block-tree product, GMRES) against
the dense FP64 oracle and against the reference's own H-matrix library (oracle/_ref).

Bars: the product (GR + deflation) x within 10 eps of the dense product; GMRES to its relative
residual (preconditioned by the hierarchical factorization at eps_pre, or block Jacobi); D_t, D_r and [eta] within 1e-5 (eps = 1e-6) of the dense answers; the reference library
at the same eps agrees with the dense answers to the same order, so the two H-matrix answers agree
to twice that."""
import os
import re

import numpy as np
import pytest
from conftest import GOLDEN, rel

pytestmark = pytest.mark.gpu

EPS = 1e-6


@pytest.fixture(scope="module")
def solver(bh):
    s = bh.Solver(0)
    yield s
    s.close()


def _prepared(oracle, fname):
    v1, v2, v3, ta, tv = oracle.read_best(os.path.join(GOLDEN, fname))
    return oracle.prep(v1, v2, v3, 1, ta), (v1, v2, v3, ta, tv)


def test_product_matches_dense(bh, oracle, solver):
    from best_hydro_amd import meshes as M
    path = os.path.join(GOLDEN, "2eu7_02562")
    tri = M.read_best(path)
    m, (v1, v2, v3, ta, tv) = _prepared(oracle, "2eu7_02562")
    n = m["nar"]
    solver.set_mesh(tri)
    solver.set_method("hmatrix", eps=EPS, tol=1e-10)
    solver.solve_mesh(precision=64)
    st = solver.hm_stats()
    print("hm stats", st, flush=True)
    assert st["n_lowrank"] > 0 and st["kmax"] > 0
    x_o, gr = oracle.tgetgr(m)
    icomp = np.ones(n, dtype=np.int32)
    oracle.deflate(m, gr, icomp, 1)
    rng = np.random.default_rng(5)
    x = rng.standard_normal((3 * n, 4))
    y = solver.hm_apply(x)
    want = gr @ x
    err = np.linalg.norm(y - want) / np.linalg.norm(want)
    print(f"product relative error {err:.2e}", flush=True)
    assert err < 10 * EPS
    # deterministic: the same product twice, bit for bit
    assert np.array_equal(solver.hm_apply(x), y)
    solver.set_method("dense")


def test_recompression_ranks_match_svd(bh, oracle, solver):
    """The recompression of the cross approximations (the reference's truncate: QR of both
    factors, SVD of Ru Rv^T, singular values above eps * s_max kept; hmatrix/src/mat.cpp:129-160):
    every low-rank block's rank against the numerical rank at eps of the exact block (numpy's SVD
    of the oracle's GR entries), the ranks never above the cross approximation's, and the product
    within 10 eps of the dense one with and without the recompression."""
    from best_hydro_amd import meshes as M
    path = os.path.join(GOLDEN, "2eu7_02562")
    m, _ = _prepared(oracle, "2eu7_02562")
    n = m["nar"]
    x_o, gr = oracle.tgetgr(m)
    grd = gr.copy(order="F")
    oracle.deflate(m, grd, np.ones(n, dtype=np.int32), 1)
    rng = np.random.default_rng(11)
    x = rng.standard_normal((3 * n, 3))
    want = grd @ x
    solver.set_mesh(M.read_best(path))
    solver.set_method("hmatrix", eps=EPS, tol=1e-10)
    res = {}
    for trunc in (0, 1):
        solver.set_hm_option("trunc", trunc)
        solver.solve_mesh(precision=64)
        st = solver.hm_stats()
        blocks, perm = solver.hm_blocks()
        err = np.linalg.norm(solver.hm_apply(x) - want) / np.linalg.norm(want)
        res[trunc] = (st, blocks, perm, err)
        print(f"trunc={trunc}: {st['n_lowrank']} low-rank blocks, rank max {st['kmax']} mean {st['kmean']:.2f} "
              f"(cross approximation {st['kmean_aca']:.2f}), {st['bytes_lowrank'] / 1e6:.1f} MB low-rank, "
              f"{st['n_trunc_failed']} kept, {st['t_trunc']:.3f} s; product error {err:.2e}", flush=True)
        assert err < 10 * EPS
    (st0, b0, p0, _), (st1, b1, p1, _) = res[0], res[1]
    assert np.array_equal(p0, p1) and np.array_equal(b0[:, [0, 2, 3, 4, 5]], b1[:, [0, 2, 3, 4, 5]])
    lr = b1[:, 0] == 1
    assert np.all(b1[lr, 1] <= b0[lr, 1])
    assert st1["bytes_lowrank"] < 0.9 * st0["bytes_lowrank"]
    assert st1["n_trunc_failed"] <= 0.01 * st1["n_lowrank"]
    # unknown 3 p + a in tree order is BEST's 3 perm[p] + a
    idx = (3 * p1.astype(np.int64)[:, None] + np.arange(3)).ravel()
    diffs = []
    for t, k, r0, nr, c0, nc in b1[lr]:
        a = gr[np.ix_(idx[r0:r0 + nr], idx[c0:c0 + nc])]
        s = np.linalg.svd(a, compute_uv=False)
        diffs.append(int(k) - int(np.sum(s > EPS * s[0])))
    diffs = np.array(diffs)
    print(f"rank - numerical rank at eps over {len(diffs)} blocks: min {diffs.min()} max {diffs.max()}, "
          f"{np.mean(diffs == 0):.3f} equal, {np.mean(np.abs(diffs) <= 1):.3f} within one", flush=True)
    assert np.all(np.abs(diffs) <= 2)
    assert np.mean(np.abs(diffs) <= 1) >= 0.95
    solver.set_method("dense")


@pytest.mark.parametrize("single", [False, True])
def test_transport_matches_dense_oracle(bh, oracle, solver, single):
    from best_hydro_amd import meshes as M
    path = os.path.join(GOLDEN, "2eu7_02562")
    v1, v2, v3, ta, tv = oracle.read_best(path)
    ref = oracle.best_run(v1, v2, v3, ta, tv, iarea=1, cg=1.0, ivisco=1, icov=0)
    solver.set_mesh(M.read_best(path))
    solver.set_method("hmatrix", eps=EPS, single=single, tol=1e-10)
    r = solver.solve_mesh(ivisco=1)
    st = solver.hm_stats()
    print(f"single={single}: iterations {st['iters_min']}..{st['iters_max']}, relres {st['relres']:.1e}, "
          f"D_t {rel(r['dtrans'], ref['dtrans']):.1e} D_r {rel(r['dr'], ref['dr']):.1e} "
          f"eta {rel(r['chiprime'], ref['chiprime']):.1e}", flush=True)
    assert st["relres"] <= 1e-10
    tol = 1e-5 if not single else 3e-5
    assert rel(r["dtrans"], ref["dtrans"]) < tol
    assert rel(r["dr"], ref["dr"]) < tol
    assert rel(r["chiprime"], ref["chiprime"]) < tol
    solver.set_method("dense")


def test_two_surfaces(bh, oracle, solver):
    """NDEF = 2: the per-surface deflation in the H-matrix product."""
    from best_hydro_amd import meshes as M
    path = os.path.join(GOLDEN, "1aki49_atom_01806")
    v1, v2, v3, ta, tv = oracle.read_best(path)
    ref = oracle.best_run(v1, v2, v3, ta, tv, iarea=1, cg=1.0, ivisco=1, icov=0)
    solver.set_mesh(M.read_best(path))
    solver.set_method("hmatrix", eps=EPS, tol=1e-10)
    r = solver.solve_mesh(ivisco=1)
    assert r["ndef"] == 2
    assert rel(r["dtrans"], ref["dtrans"]) < 1e-5 and rel(r["chiprime"], ref["chiprime"]) < 1e-5
    solver.set_method("dense")


def test_hmatrix_abi_against_reference_library(bh, oracle, monkeypatch):
    """bhm_build/bhm_solve on the H-matrix path (BEST_HIP_SOLVER=hmatrix) and the reference's
    library at the same eps: both close to LAPACK on the dense GR, and to each other."""
    R = oracle.reference_lib()
    m, _ = _prepared(oracle, "2eu7_02562")
    n = m["nar"]
    icomp, ncomp, nopen = oracle.compon(m)
    x_o, gr = oracle.tgetgr(m)
    oracle.deflate(m, gr, icomp, ncomp)
    lu, piv, info = oracle.lapack_getrf(gr)
    xo, sb = oracle.self_blocks(m)
    b = oracle.difco_rhs(n, xo)
    want = oracle.lapack_getrs(lu, piv, b.copy())
    monkeypatch.setenv("BEST_HIP_SOLVER", "hmatrix")
    monkeypatch.setenv("BEST_HMATRIX_GMRES_TOL", "1e-10")
    hm = bh.hm
    assert hm.create(n, m["v1"], m["v2"], m["v3"], m["area"], m["afac"], 1.0, 0) == 0
    diagm = hm.diag_mean()
    xmu, gdef = oracle.deflation_weights(m, diagm, icomp, ncomp)
    assert hm.build(EPS, 0, 0.1, ncomp, gdef, icomp, xmu, 0, 6) == 0
    rc, sol = hm.solve(b.copy())
    lines = hm.log_lines()
    print("\n".join(lines), flush=True)
    assert rc == 0
    assert any("GMRES" in l for l in lines)
    e_gpu = np.max(np.abs(sol - want)) / np.max(np.abs(want))
    print(f"GPU H-matrix vs LAPACK: {e_gpu:.2e}", flush=True)
    assert e_gpu < 1e-4
    hm.destroy()
    if R is None:
        return
    assert R.bhm_create(n, m["v1"], m["v2"], m["v3"], m["area"], m["afac"], 1.0, 8) == 0
    assert R.bhm_build(EPS, 0, 0.1, ncomp, gdef, icomp, xmu, 0, 6) == 0
    bref = b.copy(order="F").reshape(-1, order="F")
    assert R.bhm_solve(6, bref) == 0
    sol_ref = bref.reshape(b.shape, order="F")
    e_ref = np.max(np.abs(sol_ref - want)) / np.max(np.abs(want))
    e_pair = np.max(np.abs(sol - sol_ref)) / np.max(np.abs(want))
    print(f"reference H-matrix vs LAPACK: {e_ref:.2e}; GPU vs reference: {e_pair:.2e}", flush=True)
    assert e_pair < max(2 * e_ref, 1e-4)


def test_large_mesh_against_gpu_dense(bh, solver):
    """20,480 triangles (61,440 unknowns): the H-matrix answer against the dense GPU solve (FP32
    LU + FP64 refinement, itself within 1e-10 of FP64) -- the size-independent check of the path
    between the oracle's sizes and the 81,920-triangle case."""
    from best_hydro_amd import meshes as M
    tri = M.hmatrix_sphere_mesh(level=5)
    solver.set_mesh(tri)
    solver.set_method("dense")
    d = solver.solve_mesh(ivisco=1, precision=32)
    solver.set_method("hmatrix", eps=EPS, tol=1e-10)
    h = solver.solve_mesh(ivisco=1)
    st = solver.hm_stats()
    print(f"20480: iterations {st['iters_min']}..{st['iters_max']}, D_t {rel(h['dtrans'], d['dtrans']):.1e} "
          f"D_r {rel(h['dr'], d['dr']):.1e} eta {rel(h['chiprime'], d['chiprime']):.1e}", flush=True)
    assert st["relres"] <= 1e-10
    for key in ("dtrans", "dr", "chiprime"):
        assert rel(h[key], d[key]) < 1e-6, key
    solver.set_method("dense")


def test_c5_matches_reference_library(bh, solver):
    """The 81,920-triangle case (245,760 unknowns, eps 1e-6, GMRES to 1e-10) against the
    reference's own H-matrix library run on the CPU with the oracle's prep, deflation weights and
    DIFCO reductions around it (tests/golden/c5_reference_hmatrix.json, written by
    tests/golden/make_c5_reference.py): the dense oracle cannot run at this size.  Both are
    eps = 1e-6 approximations of the same GR solved to 1e-10, so D_t, D_r and the tensors agree
    far below eps (bars 1e-8 on D_t and D_r, 1e-7 of the largest entry on the tensors)."""
    import json
    from best_hydro_amd import meshes as M
    with open(os.path.join(GOLDEN, "c5_reference_hmatrix.json")) as f:
        ref = json.load(f)
    tri = M.hmatrix_sphere_mesh()
    assert tri.nar == ref["triangles"]
    solver.set_mesh(tri)
    solver.set_method("hmatrix", eps=ref["eps"], tol=ref["gmres_tol"])
    r = solver.solve_mesh(ivisco=0)
    st = solver.hm_stats()
    errs = {"dtrans": rel(r["dtrans"], ref["dtrans"]), "dr": rel(r["dr"], ref["dr"])}
    for key in ("dtt", "drr", "dtr"):
        want = np.array(ref[key])
        errs[key] = float(np.max(np.abs(np.asarray(r[key]) - want)) / np.max(np.abs(want)))
    print(f"C5: {st['n_dense']} dense + {st['n_lowrank'] + st['n_small'] + st['n_capped']} admissible blocks, "
          f"{(st['bytes_dense'] + st['bytes_lowrank']) / 1e9:.2f} GB, GMRES {st['iters_min']}..{st['iters_max']} "
          f"iterations; against the reference library: "
          + ", ".join(f"{k} {v:.1e}" for k, v in errs.items()), flush=True)
    # the block tree is the reference's: its log line "77584 dense and 130698 low-rank blocks"
    line = next(l for l in ref["log"] if l.startswith("H-matrix in"))
    nd, nlr = (int(w) for w in re.search(r"(\d+) dense and (\d+) low-rank", line).groups())
    assert st["n_dense"] == nd and st["n_lowrank"] + st["n_small"] + st["n_capped"] == nlr
    assert errs["dtrans"] < 1e-8 and errs["dr"] < 1e-8
    for key in ("dtt", "drr", "dtr"):
        assert errs[key] < 1e-7, key
    # the estimate made before the build bounds what the build took
    print(f"memory: estimate {st['mem_estimate'] / 1e9:.1f} GB, the build took {st['mem_used'] / 1e9:.1f} GB", flush=True)
    assert st["mem_used"] <= st["mem_estimate"]
    solver.set_method("dense")


def test_hierarchical_preconditioner(bh, solver):
    """GMRES preconditioned by the hierarchical factorization of the H-matrix at eps_pre = 0.1
    (BEST's EPSPRE) on the 2,562-triangle example: fewer iterations than block Jacobi on the diagonal
    leaves, and the same answers (within 1e-8 of the dense GPU solve; the H-matrix at eps 1e-6,
    GMRES to 1e-10)."""
    from best_hydro_amd import meshes as M
    solver.set_mesh(M.read_best(os.path.join(GOLDEN, "2eu7_02562")))
    d = solver.solve_mesh(ivisco=1, precision=64)
    iters = {}
    for pc in (0, 1):
        solver.set_method("hmatrix", eps=EPS, tol=1e-10)
        solver.set_hm_option("precond", pc)
        h = solver.solve_mesh(ivisco=1)
        st = solver.hm_stats()
        iters[pc] = st["iters_max"]
        print(f"precond {pc}: {st['iters_min']}..{st['iters_max']} iterations, eps_pre {st['eps_pre']}, pivot ratio "
              f"{st['piv_ratio']:.1e}, levels {st['pre_levels']}, widest basis {st['pre_rank_max']}", flush=True)
        assert st["relres"] <= 1e-10
        for key in ("dtrans", "dr", "chiprime"):
            assert rel(h[key], d[key]) < 1e-8, (pc, key)
        if pc:
            assert st["eps_pre"] == 0.1 and st["piv_ratio"] > 0 and st["pre_levels"] == 7
    assert iters[1] < 0.7 * iters[0]
    solver.set_hm_option("precond", 0)
    solver.set_method("dense")


def test_preconditioner_rebuilt_finer(bh, solver):
    """GMRES that does not converge within its step limit rebuilds its preconditioner ten times
    finer and runs again (the reference's finer(), hmatrix/src/api.cpp:554-578): eps_pre 0.3 with
    25 steps allowed."""
    from best_hydro_amd import meshes as M
    solver.set_mesh(M.read_best(os.path.join(GOLDEN, "2eu7_02562")))
    solver.set_method("hmatrix", eps=EPS, tol=1e-10)
    solver.set_hm_option("precond", 1)
    solver.set_hm_option("eps_pre", 0.3)
    solver.set_hm_option("maxsteps", 25)
    h = solver.solve_mesh(ivisco=0)
    st = solver.hm_stats()
    log = solver.log()
    print("\n".join(l for l in log if "precond" in l or "GMRES" in l), flush=True)
    assert any("rebuilding it at eps 3e-02" in l for l in log)
    assert abs(st["eps_pre"] - 0.03) < 1e-12 and st["relres"] <= 1e-10
    solver.set_hm_option("eps_pre", 0.1)
    solver.set_hm_option("maxsteps", 500)
    solver.set_hm_option("precond", 0)
    solver.set_method("dense")


def test_block_jacobi_falls_back_to_the_factorization(bh, solver):
    """Block Jacobi (the default) with 30 GMRES steps allowed does not converge on the 2,562-triangle
    example (it needs about 50); the solve then builds the hierarchical factorization at eps_pre and
    converges with it."""
    from best_hydro_amd import meshes as M
    solver.set_mesh(M.read_best(os.path.join(GOLDEN, "2eu7_02562")))
    solver.set_method("hmatrix", eps=EPS, tol=1e-10)
    solver.set_hm_option("maxsteps", 30)
    h = solver.solve_mesh(ivisco=0)
    st = solver.hm_stats()
    log = solver.log()
    solver.set_hm_option("maxsteps", 500)
    solver.set_method("dense")
    assert any("block Jacobi was not enough" in l for l in log)
    assert st["eps_pre"] == 0.1 and st["relres"] <= 1e-10


def test_direct_hlu_against_reference_library(bh, oracle, monkeypatch):
    """-hlu through the C ABI: bhm_build with direct = 1 factors the H-matrix hierarchically at eps
    and bhm_solve solves with those factors, refined against the H-matrix (no Krylov method), on the
    2,562-triangle example at eps 1e-3.  The reference's own direct mode at the same eps (its H-LU,
    hmatrix/src/api.cpp:461-495, oracle/_ref) is 6e-3 from LAPACK there; ours must be at least as
    close."""
    R = oracle.reference_lib()
    m, _ = _prepared(oracle, "2eu7_02562")
    n = m["nar"]
    icomp, ncomp, nopen = oracle.compon(m)
    x_o, gr = oracle.tgetgr(m)
    oracle.deflate(m, gr, icomp, ncomp)
    lu, piv, info = oracle.lapack_getrf(gr)
    xo, sb = oracle.self_blocks(m)
    b = oracle.difco_rhs(n, xo)
    want = oracle.lapack_getrs(lu, piv, b.copy())
    eps = 1e-3
    monkeypatch.setenv("BEST_HIP_SOLVER", "hmatrix")
    hm = bh.hm
    assert hm.create(n, m["v1"], m["v2"], m["v3"], m["area"], m["afac"], 1.0, 0) == 0
    diagm = hm.diag_mean()
    xmu, gdef = oracle.deflation_weights(m, diagm, icomp, ncomp)
    assert hm.build(eps, 1, 0.1, ncomp, gdef, icomp, xmu, 0, 6) == 0
    rc, sol = hm.solve(b.copy())
    lines = hm.log_lines()
    print("\n".join(lines), flush=True)
    assert rc == 0
    assert any("-hlu" in l for l in lines) and not any("by GMRES" in l for l in lines)
    e_gpu = np.max(np.abs(sol - want)) / np.max(np.abs(want))
    hm.destroy()
    e_ref = 6.2e-3  # the reference library's direct mode here (measured below when it is built)
    if R is not None:
        assert R.bhm_create(n, m["v1"], m["v2"], m["v3"], m["area"], m["afac"], 1.0, 8) == 0
        assert R.bhm_build(eps, 1, 0.1, ncomp, gdef, icomp, xmu, 0, 6) == 0
        bref = b.copy(order="F").reshape(-1, order="F")
        assert R.bhm_solve(6, bref) == 0
        e_ref = np.max(np.abs(bref.reshape(b.shape, order="F") - want)) / np.max(np.abs(want))
    print(f"-hlu at eps {eps}: GPU {e_gpu:.2e} from LAPACK, the reference library {e_ref:.2e}", flush=True)
    assert e_gpu <= e_ref


def test_memory_estimate_refuses_before_the_build(bh, oracle, solver, monkeypatch):
    """The memory contract of hmatrix/best_hmatrix.h:48-60 on device memory: with a floor above the
    device's memory the build is refused before anything is allocated, with a message; with
    BEST_HMATRIX_MEMCHECK=0 it goes ahead (handle API and bhm_build)."""
    from best_hydro_amd import meshes as M
    solver.set_mesh(M.read_best(os.path.join(GOLDEN, "1aki49_atom_01806")))
    solver.set_method("hmatrix", eps=1e-4, tol=1e-8)
    solver.set_hm_option("mem_floor_gb", 1e6)
    with pytest.raises(bh.HipError, match="not enough device memory for the H-matrix solver"):
        solver.solve_mesh(ivisco=0)
    solver.set_hm_option("memcheck", 0)
    solver.solve_mesh(ivisco=0)
    solver.set_hm_option("memcheck", 1)
    solver.set_hm_option("mem_floor_gb", -1)
    solver.set_method("dense")
    m, _ = _prepared(oracle, "1aki49_atom_01806")
    n = m["nar"]
    icomp, ncomp, nopen = oracle.compon(m)
    monkeypatch.setenv("BEST_HIP_SOLVER", "hmatrix")
    monkeypatch.setenv("BEST_HMATRIX_MEMFLOOR", "1000000")
    hm = bh.hm
    assert hm.create(n, m["v1"], m["v2"], m["v3"], m["area"], m["afac"], 1.0, 0) == 0
    xmu, gdef = oracle.deflation_weights(m, hm.diag_mean(), icomp, ncomp)
    assert hm.build(1e-4, 0, 0.1, ncomp, gdef, icomp, xmu, 0, 6) != 0
    lines = hm.log_lines()
    assert any("not enough device memory" in l and "BEST_HMATRIX_MEMCHECK=0" in l for l in lines)
    monkeypatch.setenv("BEST_HMATRIX_MEMCHECK", "0")
    assert hm.build(1e-4, 0, 0.1, ncomp, gdef, icomp, xmu, 0, 6) == 0
    hm.destroy()
