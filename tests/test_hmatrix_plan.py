"""The H-matrix path's cluster and block trees against the reference library (oracle/_ref, the
reference's hmatrix sources compiled as they are): the same depth, leaf clusters and leaf sizes
(hmatrix/src/cluster.cpp:91-112), and the same numbers of inadmissible and admissible leaf
blocks (hmatrix/src/hmatrix.cpp:27-54) as its build logs.  Host code only (no GPU)."""
import os
import re

import numpy as np
import pytest
from conftest import GOLDEN


def _ref_tree(oracle, R, m):
    n = m["nar"]
    assert R.bhm_create(n, m["v1"], m["v2"], m["v3"], m["area"], m["afac"], 1.0, 4) == 0
    x = np.zeros(3 * n)
    R.bhm_incenters(x)
    icomp, ncomp, nopen = oracle.compon(m)
    diagm = R.bhm_diag_mean()
    xmu, gdef = oracle.deflation_weights(m, diagm, icomp, ncomp)
    assert R.bhm_build(1e-4, 0, 0.1, ncomp, gdef, icomp, xmu, 0, 6) == 0
    lines = oracle.ref_log(R)
    tree = next(l for l in lines if l.startswith("cluster tree:"))
    d, nl, lo, hi = map(int, re.search(r"depth (\d+), (\d+) leaf clusters of (\d+) to (\d+)", tree).groups())
    blocks = next(l for l in lines if "dense and" in l and "low-rank blocks" in l)
    nd, nlr = map(int, re.search(r"(\d+) dense and (\d+) low-rank blocks", blocks).groups())
    return x, dict(depth=d, nleaves=nl, leaf_min=lo, leaf_max=hi, n_dense=nd, n_lowrank=nlr)


@pytest.mark.parametrize("fname", ["2eu7_02562", "1aki49_atom_01806"])
def test_trees_match_the_reference(bh, oracle, fname):
    R = oracle.reference_lib()
    if R is None:
        pytest.skip("oracle/_ref not built")
    v1, v2, v3, ta, tv = oracle.read_best(os.path.join(GOLDEN, fname))
    m = oracle.prep(v1, v2, v3, 1, ta)
    n = m["nar"]
    x, ref = _ref_tree(oracle, R, m)

    def nx3(a):
        return np.asarray(a).reshape(3, n).T

    plan = bh.hm_plan(nx3(m["v1"]), nx3(m["v2"]), nx3(m["v3"]), nx3(x), leaf=32, eta=2.0)
    print(fname, plan, ref)
    for k in ("depth", "nleaves", "leaf_min", "leaf_max", "n_dense"):
        assert plan[k] == ref[k], k
    # the reference stores small admissible blocks low-rank too
    assert plan["n_lowrank"] + plan["n_small"] == ref["n_lowrank"]


def test_tree_levels_and_leaf_sizes(bh):
    """Every leaf at one depth, sizes differing by at most one (level-consistent tree)."""
    from best_hydro_amd import meshes as M
    t = M.sphere_mesh(4, 15.0)
    x = (t.v1 + t.v2 + t.v3) / 3.0
    plan = bh.hm_plan(t.v1, t.v2, t.v3, x, leaf=32)
    n = len(t.v1)
    assert plan["nleaves"] == 2 ** plan["depth"]
    assert plan["leaf_max"] - plan["leaf_min"] <= 1
    assert plan["leaf_max"] <= 32 and (n + 2 ** (plan["depth"] - 1) - 1) // 2 ** (plan["depth"] - 1) > 32
