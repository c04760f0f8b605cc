"""bhg_besolv_noalloc (the BESOLV twin whose work vectors live in the library's static device work
space, for callers that cannot allocate) against bhg_besolv on the same factors, through the C ABI:
two calls in a row, the second with a smaller input (a smaller mesh, fewer right-hand sides), both
LU precisions, and the size limit rejected with an error.

Bar: the same kernels on the same data in a fixed order, so the solutions agree to 1e-15 of each
column's largest entry (bit for bit in practice; the test reports it)."""
import ctypes as C
import os

import numpy as np
import pytest
from conftest import GOLDEN

pytestmark = pytest.mark.gpu


def _bind(L):
    L.bhg_besolv_noalloc.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_long, C.c_void_p, C.c_long]
    L.bhg_besolv_noalloc.restype = C.c_int


def _hip():
    """The HIP runtime libbesthip.so links (test plumbing for device buffers; torch carries a HIP
    runtime of its own, which must not be mixed with the library's in one process)."""
    H = C.CDLL("libamdhip64.so.7")
    H.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    H.hipFree.argtypes = [C.c_void_p]
    H.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    return H


def _noalloc(bh, s, b):
    """b (3n x nrhs, host) -> solutions by bhg_besolv_noalloc, with device buffers of the test's own."""
    L = bh.lib()
    _bind(L)
    H = _hip()
    n3, nrhs = b.shape
    hb = np.ascontiguousarray(np.asfortranarray(b).reshape(-1, order="F"))
    hx = np.full(n3 * nrhs, np.nan)
    db, dx = C.c_void_p(), C.c_void_p()
    assert H.hipMalloc(C.byref(db), hb.nbytes) == 0 and H.hipMalloc(C.byref(dx), hx.nbytes) == 0
    try:
        assert H.hipMemcpy(db, hb.ctypes.data, hb.nbytes, 1) == 0   # host to device
        assert H.hipMemcpy(dx, hx.ctypes.data, hx.nbytes, 1) == 0   # NaN: every output entry must be written
        rc = L.bhg_besolv_noalloc(s.h, nrhs, db, n3, dx, n3)
        if rc != 0:
            raise bh.HipError(L.bhg_last_error(s.h).decode())
        assert H.hipMemcpy(hx.ctypes.data, dx, hx.nbytes, 2) == 0   # device to host
    finally:
        H.hipFree(db)
        H.hipFree(dx)
    return hx.reshape(nrhs, n3).T


def _rhs(oracle, t, nrhs):
    m = oracle.prep(t.v1, t.v2, t.v3, 1, t.truearea)
    x, _ = oracle.self_blocks(m)
    b = np.concatenate([oracle.difco_rhs(m["nar"], x), oracle.strain_rhs(m["nar"], x)], axis=1)
    return np.asfortranarray(b[:, :nrhs])


def _agree(a, b):
    for k in range(a.shape[1]):
        assert np.max(np.abs(a[:, k] - b[:, k])) <= 1e-15 * np.max(np.abs(b[:, k])), k


@pytest.mark.parametrize("prec", [32, 64])
def test_noalloc_matches_besolv_twice(bh, oracle, prec):
    from best_hydro_amd import meshes as M
    big = M.read_best(os.path.join(GOLDEN, "2eu7_02562"))
    small = M.sphere_mesh(3, 15.0)
    for t, nrhs in ((big, 11), (small, 6)):
        print(f"precision {prec}: {t.nar} triangles, {nrhs} right-hand sides", flush=True)
        s = bh.Solver(0)
        s.set_mesh(t)
        s.solve_mesh(ivisco=1, precision=prec)  # factors resident
        b = _rhs(oracle, t, nrhs)
        want = s.besolv(b)
        got = _noalloc(bh, s, b)
        print(f"  bit-identical: {np.array_equal(got, want)}", flush=True)
        _agree(got, want)
        s.close()


def test_noalloc_rejects_beyond_its_work_space(bh):
    from best_hydro_amd import meshes as M
    t = M.series_meshes(sizes=(6358,))[0]   # 3 x 6,358 unknowns: padded order 19,456 > 16,384
    s = bh.Solver(0)
    s.set_mesh(t)
    s.solve_mesh(ivisco=0, precision=32)
    b = np.zeros((3 * t.nar, 1))
    b[0, 0] = 1.0
    with pytest.raises(bh.HipError, match="exceeds"):
        _noalloc(bh, s, b)
    s.close()
