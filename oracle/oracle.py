"""The oracle: a CPU FP64 restatement of BEST's dense boundary-element path.

TEST INFRASTRUCTURE ONLY.  Tests, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` import this module as the checker; the product (``best-hydro_amd``) never does.

The arithmetic lives in ``best_oracle.c`` (built into ``liboracle.so`` by ``oracle/Makefile``),
which restates the reference routine by routine; each wrapper below names the reference lines.
The dense factorization is LAPACK's ``dgetrf``/``dgetrs`` -- the same third-party routines the
reference calls at ``best/best.f:444`` and ``best/besolve.f90:133`` -- through SciPy's bundled
LAPACK, with the C textbook LU (``bo_getrf``) as an independent cross-check for small systems.

``reference_lib()`` loads ``oracle/_ref/libbhm_ref.so``: the reference's own hmatrix library built
from its sources by ``make -C oracle ref``.  Its ``bhm_entries``/``bhm_incenters``/``bhm_diag_mean``
pin the oracle's matrix entries bit for bit, and its ``bhm_build``/``bhm_solve`` are the H-matrix
CPU reference.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_REF = None

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


def build(ref: bool | None = None) -> None:
    """Compiles liboracle.so (and oracle/_ref when the reference sources are present)."""
    targets = ["all"]
    if ref is None:
        ref = os.path.isdir("/root/reference/hmatrix/src")
    if ref:
        targets.append("ref")
    subprocess.run(["make", "-s", "-C", HERE, *targets, "-j8"], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            build(ref=False)
        L = C.CDLL(path)
        i, d, pd, pi_ = C.c_int, C.c_double, _dp, _ip
        L.bo_prep.argtypes = [i, pd, pd, pd, i, d, pd, pd, pd, pd]
        L.bo_compon.argtypes = [i, pd, pd, pd, pi_, C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.bo_defgam.argtypes = [i, i, pi_, pd, d, pd]
        L.bo_self.argtypes = [i, i, d, d, pd, pd, pd, pd, pd]
        L.bo_pair_xc.argtypes = [i, i, pd, d, pd, pd, pd, d, pd]
        L.bo_selfall.argtypes = [i, d, d, pd, pd, pd, pd, pd]
        L.bo_entries.argtypes = [i, d, d, pd, pd, pd, pd, pd, pd, i, i, i, i, pd, i]
        L.bo_tgetgr.argtypes = [i, d, d, pd, pd, pd, pd, pd, pd, C.c_long]
        L.bo_deflate.argtypes = [i, pd, pd, pd, pd, i, pi_, C.c_void_p, C.c_long, pd, pd, pd]
        L.bo_getrf.argtypes = [i, pd, C.c_long, pi_]
        L.bo_getrf.restype = C.c_int
        L.bo_getrs.argtypes = [i, i, pd, C.c_long, pi_, pd, C.c_long]
        L.bo_matin3.argtypes = [pd, pd]
        L.bo_matin3.restype = C.c_int
        L.bo_jacobi.argtypes = [pd, pd]
        L.bo_difco_rhs.argtypes = [i, pd, pd]
        L.bo_difco_post.argtypes = [i, pd, pd, d, pd, pd, pd, pd, pd, pd, pd]
        L.bo_strain.argtypes = [i, pd, pd]
        L.bo_visco_post.argtypes = [i, pd, pd, i, d, pd, pd, pd]
        L.bo_visco_post.restype = C.c_double
        L.bo_units.argtypes = [pd, pd, pd]
        L.bo_gauss_table.argtypes = [pd, pd]
        _LIB = L
    return _LIB


def reference_lib():
    """The reference's hmatrix library built from its own sources, or None if not built."""
    global _REF
    if _REF is None:
        path = os.path.join(HERE, "_ref", "libbhm_ref.so")
        if not os.path.exists(path):
            return None
        R = C.CDLL(path)
        i, d = C.c_int, C.c_double
        R.bhm_available.restype = i
        R.bhm_create.argtypes = [i, _dp, _dp, _dp, _dp, d, d, i]
        R.bhm_create.restype = i
        R.bhm_incenters.argtypes = [_dp]
        R.bhm_entries.argtypes = [i, i, i, i, _dp, i]
        R.bhm_diag_mean.restype = d
        R.bhm_build.argtypes = [d, i, d, i, _dp, _ip, _dp, i, i]
        R.bhm_build.restype = i
        R.bhm_solve.argtypes = [i, _dp]
        R.bhm_solve.restype = i
        R.bhm_log_line.argtypes = [C.c_char_p, i]
        R.bhm_log_line.restype = i
        _REF = R
    return _REF


def ref_log(R) -> list[str]:
    out, buf = [], C.create_string_buffer(512)
    while True:
        n = R.bhm_log_line(buf, 512)
        if n < 0:
            return out
        out.append(buf.raw[:n].decode(errors="replace"))


# ------------------------------------------------------------------ mesh files

def read_best(path: str):
    """BEST's input file (best/best.f:189-220): NAR[, TRUEAREA, TRUEVOL], then 9 numbers per
    triangle.  Returns (v1, v2, v3 as NAR x 3 Fortran-order arrays, truearea, truevol)."""
    with open(path) as f:
        head = f.readline().split()
        body = np.array(f.read().split(), dtype=np.float64)
    nar = int(head[0])
    truearea = float(head[1]) if len(head) > 1 else 0.0
    truevol = float(head[2]) if len(head) > 2 else 0.0
    t = body[: 9 * nar].reshape(nar, 9)
    return (np.asfortranarray(t[:, 0:3]), np.asfortranarray(t[:, 3:6]), np.asfortranarray(t[:, 6:9]),
            truearea, truevol)


def _col(v):
    """NAR x 3 array -> the flat column-major buffer BEST passes."""
    return np.ascontiguousarray(np.asarray(v, dtype=np.float64).T).ravel()


def _uncol(buf, nar):
    return buf.reshape(3, nar).T.copy()


# ------------------------------------------------------------------ wrappers

def prep(v1, v2, v3, iarea=1, truearea=0.0):
    """best/best.f:261-315.  Returns dict(v1, v2, v3 (shifted, flat column-major), area, afac,
    totarea, centroid)."""
    nar = len(v1)
    b1, b2, b3 = _col(v1), _col(v2), _col(v3)
    area = np.zeros(nar)
    afac, tot, cen = np.zeros(1), np.zeros(1), np.zeros(3)
    lib().bo_prep(nar, b1, b2, b3, int(iarea), float(truearea), area, afac, tot, cen)
    return dict(nar=nar, v1=b1, v2=b2, v3=b3, area=area, afac=float(afac[0]), totarea=float(tot[0]),
                centroid=cen)


def compon(m):
    """best/compon.f90:19-115 -> (icomp 1-based int32[N], ncomp, nopen)."""
    nar = m["nar"]
    icomp = np.zeros(nar, dtype=np.int32)
    nc, no = C.c_int(), C.c_int()
    lib().bo_compon(nar, m["v1"], m["v2"], m["v3"], icomp, C.byref(nc), C.byref(no))
    return icomp, nc.value, no.value


def self_blocks(m, cg=1.0):
    """best/tgetgr.f:35-145 -> (x flat column-major N x 3, self blocks N x 3 x 3)."""
    nar = m["nar"]
    x = np.zeros(3 * nar)
    sb = np.zeros(9 * nar)
    lib().bo_selfall(nar, m["afac"], float(cg), m["v1"], m["v2"], m["v3"], x, sb)
    return x, sb.reshape(nar, 3, 3)


def entries(m, x, selfb, i0, mrows, j0, ncols, cg=1.0):
    """GR(i0:i0+m, j0:j0+nc) before deflation (bhm_entries semantics), Fortran-order array."""
    a = np.zeros(mrows * ncols)
    lib().bo_entries(m["nar"], m["afac"], float(cg), m["v1"], m["v2"], m["v3"], m["area"], x,
                     np.ascontiguousarray(selfb).ravel(), int(i0), int(mrows), int(j0), int(ncols), a, int(mrows))
    return a.reshape(ncols, mrows).T


def tgetgr(m, cg=1.0):
    """best/tgetgr.f:1-173 -> (x flat column-major, GR as an n x n Fortran-order array)."""
    nar = m["nar"]
    n3 = 3 * nar
    x = np.zeros(3 * nar)
    gr = np.zeros(n3 * n3)
    lib().bo_tgetgr(nar, m["afac"], float(cg), m["v1"], m["v2"], m["v3"], x, m["area"], gr, n3)
    return x, gr.reshape(n3, n3).T  # column-major buffer viewed as Fortran-order


def deflate(m, gr, icomp, ndef):
    """best/best.f:384-426, in place on gr (Fortran-order n x n).  Returns (xmu, diagm, gdefk)."""
    nar = m["nar"]
    xmu = np.zeros(3 * nar)
    dm = np.zeros(1)
    gd = np.zeros(max(ndef, 1))
    assert gr.flags.f_contiguous
    ptr = gr.ctypes.data_as(C.c_void_p)
    lib().bo_deflate(nar, m["v1"], m["v2"], m["v3"], m["area"], int(ndef), icomp, ptr, 3 * nar, xmu, dm, gd)
    return xmu, float(dm[0]), gd


def deflation_weights(m, diagm, icomp, ndef):
    """xmu and the weights without a GR (the H-matrix path: best/besolve.f90:190-220)."""
    nar = m["nar"]
    xmu = np.zeros(3 * nar)
    dm = np.array([diagm])
    gd = np.zeros(max(ndef, 1))
    lib().bo_deflate(nar, m["v1"], m["v2"], m["v3"], m["area"], int(ndef), icomp, None, 3 * nar, xmu, dm, gd)
    return xmu, gd


def getrf_textbook(a):
    """Unblocked GEPP (dgetf2 semantics) on a copy; returns (lu Fortran-order, ipiv 1-based, info)."""
    lu = np.asfortranarray(a, dtype=np.float64).copy()
    n = lu.shape[0]
    ipiv = np.zeros(n, dtype=np.int32)
    flat = lu.reshape(-1, order="F")
    info = lib().bo_getrf(n, flat, n, ipiv)
    return flat.reshape(n, n, order="F"), ipiv, info


def getrs_textbook(lu, ipiv, b):
    n = lu.shape[0]
    bb = np.asfortranarray(b, dtype=np.float64).copy()
    nrhs = 1 if bb.ndim == 1 else bb.shape[1]
    flat = bb.reshape(-1, order="F")
    lib().bo_getrs(n, nrhs, np.asfortranarray(lu).reshape(-1, order="F"), n, ipiv, flat, n)
    return flat.reshape(bb.shape, order="F")


def lapack_getrf(gr):
    """LAPACK dgetrf, in place on the Fortran-order array (the call at best/best.f:444)."""
    from scipy.linalg import lapack
    lu, piv, info = lapack.dgetrf(gr, overwrite_a=True)
    return lu, piv, info


def lapack_getrs(lu, piv, b):
    """LAPACK dgetrs('N') (best/besolve.f90:133)."""
    from scipy.linalg import lapack
    x, info = lapack.dgetrs(lu, piv, b)
    return x


def matin3(a):
    a = np.asfortranarray(a, dtype=np.float64).reshape(-1, order="F").copy()
    y = np.zeros(9)
    ierr = lib().bo_matin3(a, y)
    return y.reshape(3, 3, order="F"), ierr


def jacobi(a):
    a = np.asfortranarray(a, dtype=np.float64).reshape(-1, order="F").copy()
    d = np.zeros(3)
    lib().bo_jacobi(a, d)
    return d


def gauss_table():
    y, w = np.zeros(21), np.zeros(21)
    lib().bo_gauss_table(y, w)
    return y, w


def difco_rhs(nar, x):
    b = np.zeros(3 * nar * 6)
    lib().bo_difco_rhs(nar, x, b)
    return b.reshape(6, 3 * nar).T.copy(order="F")


def strain_rhs(nar, x):
    v = np.zeros(3 * nar * 5)
    lib().bo_strain(nar, x, v)
    return v.reshape(5, 3 * nar).T.copy(order="F")


def difco_post(nar, x, area, cg, b6):
    """best/difco.f:51-401 -> dict of 3x3 tensors (column-major semantics kept as numpy arrays)."""
    bflat = np.asfortranarray(b6).reshape(-1, order="F").copy()
    dtt, dtr, drr = np.zeros(9), np.zeros(9), np.zeros(9)
    urc, udr, fric = np.zeros(3), np.zeros(3), np.zeros(36)
    lib().bo_difco_post(nar, x, area, float(cg), bflat, dtt, dtr, drr, urc, udr, fric)
    f = lambda a: a.reshape(3, 3, order="F")
    return dict(dtt=f(dtt), dtr=f(dtr), drr=f(drr), urc=urc, udr=udr, ktt=f(fric[0:9]), ktr=f(fric[9:18]),
                krt=f(fric[18:27]), krr=f(fric[27:36]))


def visco_post(nar, x, area, icov, cg, b6, v5):
    """best/visco.f:25-211 -> (chiprime, (D1, D2, D3))."""
    bflat = np.asfortranarray(b6).reshape(-1, order="F").copy()
    vflat = np.asfortranarray(v5).reshape(-1, order="F").copy()
    d = np.zeros(3)
    chip = lib().bo_visco_post(nar, x, area, int(icov), float(cg), bflat, vflat, d)
    return chip, d


def units(drr, dtt):
    """best/output.f -> dict(dr, dtrans, drr_eig, dtt_eig, tunit, runit)."""
    out = np.zeros(10)
    lib().bo_units(np.asfortranarray(drr).reshape(-1, order="F").copy(),
                   np.asfortranarray(dtt).reshape(-1, order="F").copy(), out)
    return dict(dr=out[0], dtrans=out[1], drr_eig=out[2:5].copy(), dtt_eig=out[5:8].copy(), tunit=out[8],
                runit=out[9])


# ------------------------------------------------------------------ the whole dense run

def best_run(v1, v2, v3, truearea=0.0, truevol=0.0, iarea=1, cg=1.0, ivisco=1, icov=0, wm=None,
             lu="lapack", keep=False):
    """BEST90's dense path (best/best.f:261-498) on the oracle: prep, COMPON, TGETGR, deflation,
    LU, DIFCO, VISCO, OUTPUT.  Returns a dict of the printed values (dtrans, dr, chi, eta, ...) and
    the tensors.  lu='lapack' uses dgetrf/dgetrs (as the reference does); 'textbook' the C GEPP."""
    m = prep(v1, v2, v3, iarea, truearea)
    nar = m["nar"]
    icomp, ncomp, nopen = compon(m)
    ndef = 1
    if nopen > 0:
        icomp[:] = 1
    elif ncomp > 1:
        ndef = ncomp
    x, gr = tgetgr(m, cg)
    gr_raw_head = np.array(gr[:9, :9]) if keep else None
    xmu, diagm, gdefk = deflate(m, gr, icomp, ndef)
    if lu == "lapack":
        fac, piv, info = lapack_getrf(gr)
        solve = lambda b: lapack_getrs(fac, piv, b)
    else:
        fac, piv, info = getrf_textbook(gr)
        solve = lambda b: getrs_textbook(fac, piv, b)
    if info != 0:
        raise RuntimeError(f"dgetrf info={info}")
    b6 = solve(difco_rhs(nar, x))
    res = difco_post(nar, x, m["area"], cg, b6)
    res.update(units(res["drr"], res["dtt"]))
    res.update(nar=nar, afac=m["afac"], totarea=m["totarea"], centroid=m["centroid"], ncomp=ncomp, nopen=nopen,
               ndef=ndef, diagm=diagm, gdefk=gdefk, x=x)
    if ivisco:
        v5 = solve(strain_rhs(nar, x))
        chip, d = visco_post(nar, x, m["area"], icov, cg, b6, v5)
        res["chiprime"] = chip
        res["dvisc"] = d
        if iarea == 1:
            res["chi"] = chip / truevol
        if wm is not None:
            res["eta"] = chip * 0.602214 / wm
        res["sol_strain"] = v5
    res["sol_rigid"] = b6
    if keep:
        res["gr_head"] = gr_raw_head
    return res
