"""COMPON on the device (csrc/compon.hip, run by bhg_set_mesh): ICOMP, NCOMP and NOPEN against the
oracle's COMPON (bo_compon, best/compon.f90:19-115) on the vertices after the centroid shift, as
best/best.f:323 calls it, and against a dictionary-based restatement where the oracle's sort
cannot be used (NaN coordinates).  Meshes: the reference's examples, the two-surface mesh, the
81,920-triangle sphere, an open mesh, forty separate spheres in shuffled triangle order (the
numbering follows the surfaces' first triangles), corners at -0 and +0, a NaN vertex.  Also the
mesh file read by the library (bhg_set_mesh_file) solves to the same bits as the same mesh given
as arrays."""
import os

import numpy as np
import pytest
from conftest import GOLDEN, no_torch_loaded

pytestmark = pytest.mark.gpu


def _device(bh, t):
    s = bh.Solver(0)
    s.set_mesh(t)
    ic, nc, no = s.components()
    s.close()
    return ic, nc, no


def _oracle(oracle, t):
    m = oracle.prep(t.v1, t.v2, t.v3, 0, 0.0)
    return oracle.compon(m)


def _restated(t):
    """best/compon.f90 by a dictionary of exact coordinates (-0 equal to 0, NaN equal to nothing)
    and a sequential union-find, on the shifted vertices the oracle's prep gives."""
    import oracle as O
    m = O.prep(t.v1, t.v2, t.v3, 0, 0.0)
    n = m["nar"]
    vs = [O._uncol(m[k], n) for k in ("v1", "v2", "v3")]
    ids, table = np.zeros((n, 3), dtype=np.int64), {}
    for t_ in range(n):
        for k in range(3):
            p = tuple(0.0 if c == 0.0 else c for c in vs[k][t_])
            ids[t_, k] = len(table) + 10 ** 9 * (t_ * 3 + k + 1) if any(c != c for c in p) else table.setdefault(p, len(table))
    parent = list(range(n))
    def find(i):
        while parent[i] != i:
            parent[i] = parent[parent[i]]
            i = parent[i]
        return i
    first = {}
    for t_ in range(n):
        for k in range(3):
            v = ids[t_, k]
            if v in first:
                a, b = find(t_), find(first[v])
                if a != b:
                    parent[max(a, b)] = min(a, b)
            else:
                first[v] = t_
    lab, ic = {}, np.zeros(n, dtype=np.int32)
    for t_ in range(n):
        r = find(t_)
        lab.setdefault(r, len(lab) + 1)
        ic[t_] = lab[r]
    edges = {}
    for t_ in range(n):
        for c in range(3):
            a, b = ids[t_, c], ids[t_, (c + 1) % 3]
            e = (min(a, b), max(a, b))
            edges[e] = edges.get(e, 0) + 1
    return ic, len(lab), sum(1 for v in edges.values() if v != 2)


def _check(bh, oracle, t, want=None):
    ic, nc, no = _device(bh, t)
    ref = want if want is not None else _oracle(oracle, t)
    print(f"{t.name}: {t.nar} triangles, NCOMP {nc}, NOPEN {no}", flush=True)
    assert nc == ref[1] and no == ref[2]
    assert np.array_equal(ic, ref[0])
    return nc, no


@pytest.mark.parametrize("fname", ["2eu7_02562", "1aki49_atom_01806"])
def test_examples(bh, oracle, fname):
    from best_hydro_amd import meshes as M
    nc, no = _check(bh, oracle, M.read_best(os.path.join(GOLDEN, fname)))
    assert no == 0 and nc == (2 if fname.startswith("1aki") else 1)


def test_two_surface_and_large(bh, oracle):
    from best_hydro_amd import meshes as M
    assert _check(bh, oracle, M.two_surface_mesh()) == (2, 0)
    assert _check(bh, oracle, M.hmatrix_sphere_mesh()) == (1, 0)


def test_open_mesh(bh, oracle):
    from best_hydro_amd import meshes as M
    t = M.sphere_mesh(3, 15.0)
    keep = np.ones(t.nar, dtype=bool)
    keep[[5, 400, 401]] = False
    o = M.Triangulation(t.v1[keep], t.v2[keep], t.v3[keep], t.truearea, t.truevol, name="open")
    nc, no = _check(bh, oracle, o)
    assert no > 0


def test_forty_spheres_shuffled(bh, oracle):
    from best_hydro_amd import meshes as M
    parts = []
    for k in range(40):
        s = M.sphere_mesh(1, 2.0)
        off = np.array([7.0 * (k % 8), 7.0 * (k // 8), 0.5 * k])
        parts.append((s.v1 + off, s.v2 + off, s.v3 + off))
    v = [np.concatenate([p[j] for p in parts]) for j in range(3)]
    perm = np.random.default_rng(11).permutation(len(v[0]))
    t = M.Triangulation(v[0][perm], v[1][perm], v[2][perm], 1.0, 1.0, name="forty")
    assert _check(bh, oracle, t) == (40, 0)


def test_signed_zero_and_nan(bh, oracle):
    from best_hydro_amd import meshes as M
    t = M.sphere_mesh(2, 1.0)
    # the icosphere has exact zero coordinates; write some of them as -0 in some triangles
    v1, v2, v3 = (a.copy() for a in (t.v1, t.v2, t.v3))
    for a in (v1, v3):
        z = a[::3] == 0.0
        a[::3][z] = -0.0
    assert np.signbit(v1[v1 == 0.0]).any() and not np.signbit(v2[v2 == 0.0]).all()
    z = M.Triangulation(v1, v2, v3, t.truearea, t.truevol, name="signed zeros")
    assert _check(bh, oracle, z) == (1, 0)
    v1 = v1.copy()
    v1[7, 1] = np.nan
    n = M.Triangulation(v1, v2, v3, t.truearea, t.truevol, name="nan vertex")
    _check(bh, oracle, n, want=_restated(n))


def test_mesh_file_solves_to_the_same_bits(bh, tmp_path):
    no_torch_loaded()
    from best_hydro_amd import meshes as M
    path = os.path.join(GOLDEN, "2eu7_02562")
    a = bh.Solver(0)
    a.set_mesh_file(path)
    ra = a.solve_mesh(ivisco=1, precision=32)
    a.close()
    b = bh.Solver(0)
    b.set_mesh(M.read_best(path))
    rb = b.solve_mesh(ivisco=1, precision=32)
    b.close()
    for k in ("dtrans", "dr", "chiprime", "dtt", "drr", "dtr"):
        assert np.array_equal(np.asarray(ra[k]), np.asarray(rb[k])), k
    ser = bh.Series([M.sphere_mesh(2, 15.0)])
    ser.set_mesh_file(0, path)
    ser.set_tuning(0, strip=64, gemm_bk=0, ir_iters=4)
    res, _ = ser.solve(ivisco=1)
    ser.close()
    assert abs(res[0]["dtrans"] - ra["dtrans"]) <= 1e-12 * abs(ra["dtrans"])
