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
CPU reference.  ``output_ref()`` loads ``oracle/_ref/libbest_output_ref.so``: the reference's
``OUTPUT`` and ``JACOBI`` (``best/output.f``, ``best/jacobi.f``) compiled by the same recipe with
amdflang; they pin ``format_be`` byte for byte and ``jacobi`` bit for bit.

``format_be`` (pure Python, below) restates ``OUTPUT``'s records and the Fortran edit descriptors
they use (Ew.d, Fw.d, Iw and list-directed character output as the Fortran runtime here writes
them).
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


_OUT = None


def output_ref():
    """The reference's OUTPUT and JACOBI compiled from best/output.f and best/jacobi.f, or None."""
    global _OUT
    if _OUT is None:
        path = os.path.join(HERE, "_ref", "libbest_output_ref.so")
        if not os.path.exists(path):
            return None
        _OUT = C.CDLL(path)
    return _OUT


def ref_output_be(path, ver, iarea, truearea, afac, truevol, mw, wm, nar, centroid, urc, udr, dtt, dtr, drr,
                  ivisco, icov, dvisc, chiprime, cg, totarea) -> bytes:
    """The .be file the reference's OUTPUT writes (best/output.f:1-176) for these values; path is a
    scratch file name (OUTPUT opens it with status 'old')."""
    R = output_ref()
    if R is None:
        raise RuntimeError("oracle/_ref/libbest_output_ref.so is not built")
    d = lambda x: C.byref(C.c_double(float(x)))
    i = lambda x: C.byref(C.c_int(int(x)))
    keep = [np.asfortranarray(np.asarray(a, dtype=np.float64)).copy(order="F") for a in (urc, udr, dtt, dtr, drr)]
    fn = path.encode().ljust(80)
    if len(fn) > 80:
        raise ValueError("path longer than 80 characters")
    open(path, "wb").close()
    ptr = [a.ctypes.data_as(C.c_void_p) for a in keep]
    R.output_(fn, ver.encode().ljust(16)[:16], i(iarea), d(truearea), d(afac), d(truevol), i(mw), d(wm), i(nar),
              d(centroid[0]), d(centroid[1]), d(centroid[2]), *ptr, i(ivisco), i(icov), d(dvisc[0]), d(dvisc[1]),
              d(dvisc[2]), d(chiprime), d(cg), d(totarea), C.c_long(80), C.c_long(16))
    with open(path, "rb") as f:
        return f.read()


def ref_jacobi(a):
    """The reference's JACOBI (best/jacobi.f) on a 3x3 symmetric matrix: the eigenvalues."""
    R = output_ref()
    a = np.asfortranarray(np.asarray(a, dtype=np.float64)).copy(order="F")
    dd, v, nrot = np.zeros(3), np.zeros((3, 3), order="F"), C.c_int(0)
    n = C.c_int(3)
    R.jacobi_(a.ctypes.data_as(C.c_void_p), C.byref(n), C.byref(n), dd.ctypes.data_as(C.c_void_p),
              v.ctypes.data_as(C.c_void_p), C.byref(nrot))
    return dd


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


# ------------------------------------------------------------------ the .be file (best/output.f)

def _fortran_e(x, w, d):
    """Ew.d output editing: [-]0.x1..xd followed by E+zz (|exponent| <= 99) or +zzz (<= 999),
    right-justified in w; the leading zero dropped when w has no room for it, asterisks when the
    field does not fit; Inf and NaN right-justified."""
    x = float(x)
    if x != x:
        t = "NaN"
        return t.rjust(w) if len(t) <= w else "*" * w
    if x in (float("inf"), float("-inf")):
        t = ("-" if x < 0 else "") + "Inf"
        return t.rjust(w) if len(t) <= w else "*" * w
    neg = x < 0 or (x == 0.0 and str(x).startswith("-"))
    if x == 0.0:
        digits, e = "0" * d, 0
    else:
        m, ee = ("%.*e" % (d - 1, abs(x))).split("e")
        digits, e = m.replace(".", ""), int(ee) + 1
    if abs(e) <= 99:
        ex = "E" + ("-" if e < 0 else "+") + "%02d" % abs(e)
    elif abs(e) <= 999:
        ex = ("-" if e < 0 else "+") + "%03d" % abs(e)
    else:
        return "*" * w
    t = ("-" if neg else "") + "0." + digits + ex
    if len(t) > w:
        t = ("-" if neg else "") + "." + digits + ex
    return t.rjust(w) if len(t) <= w else "*" * w


def _fortran_f(x, w, d):
    """Fw.d output editing: [-]i.ffff rounded to d decimals, right-justified in w; the zero before
    the point dropped when w has no room for it, asterisks when the field does not fit."""
    x = float(x)
    t = "%.*f" % (d, abs(x))
    neg = x < 0 or (x == 0.0 and str(x).startswith("-"))
    t = ("-" if neg else "") + t
    if len(t) > w and t.lstrip("-").startswith("0."):
        t = ("-" if neg else "") + t.lstrip("-")[1:]
    return t.rjust(w) if len(t) <= w else "*" * w


def _fortran_i(n, w):
    t = str(int(n))
    return t.rjust(w) if len(t) <= w else "*" * w


def format_be(ver, iarea, truearea, afac, truevol, mw, wm, nar, centroid, urc, udr, dtt, dtr, drr, ivisco, icov,
              dvisc, chiprime, cg, totarea) -> bytes:
    """The records OUTPUT writes to the .be file (best/output.f:36-168), in order, with the
    constants and the arithmetic of best/output.f:16-20 (tunit, runit), :99-106 and :121-128
    (eigenvalues by JACOBI, 1/3 trace, anisotropy, unit conversions), :151-164 (CHI, ETA).
    Tensors are 3x3 arrays (row j printed as dtt(j,1..3)); ver is padded or cut to 16 characters."""
    E = _fortran_e
    ld = lambda *items: " " + "".join(items)   # list-directed: a blank, then the characters
    fmt340 = lambda a, b, c: "".join("   " + E(v, 12, 5) for v in (a, b, c))
    temperature, eta, pi = 293.15, 0.01002, 3.141592653
    tunit = (1.38066E-16) * temperature * 1.0E8 / (8.0 * eta * pi)
    runit = tunit * 1.0e16
    dtt, dtr, drr = (np.asarray(a, dtype=np.float64).reshape(3, 3) for a in (dtt, dtr, drr))
    L = [ld("BEST Ver. ", (ver + " " * 16)[:16]), "", " BE - MACROMOLECULE TRANSPORT TENSORS", ""]
    L.append("    " + " Area (A^2)            " + E(totarea, 12, 5))
    if iarea == 1:
        L.append("    " + " True Area (A^2)       " + E(truearea, 12, 5))
        L.append("    " + " Areafactor                 " + _fortran_f(afac, 7, 4))
        L.append("    " + " True Volume (A^3)     " + E(truevol, 12, 5))
    if mw == 1:
        L.append("    " + " Molecular Weight (g/mol)  " + _fortran_f(wm, 8, 1))
    L.append("    " + " Temperature (K)           " + _fortran_f(temperature, 8, 3))
    L.append("    " + " Solvent Viscosity (Poise) " + _fortran_f(eta, 8, 5))
    L.append("    " + " Number of Surface Patches    " + _fortran_i(nar, 5))
    L.append(ld(" "))
    if cg == 0:
        L.append(ld(" PRE-AVERAGING USED FOR THIS CALCULATION"))
    L.append(ld(" "))
    L.append(ld(" Centroid Coordinates (A)"))
    L.append(fmt340(*centroid))
    L.append(ld(" Vector between Center of Resistance/Centroid "))
    L.append(fmt340(*urc))
    L.append(ld(" Vector between Centers of Diffusion/Resistance"))
    L.append(fmt340(*udr))
    L += [ld(" "), ld(" Diffusion Tensors in Center of Mobility"), ld(" "), ld(" Dtt Tensor (1/A)")]
    L += [fmt340(*dtt[j]) for j in range(3)]
    L += [ld(" "), ld(" Symmetric Dtr Tensor (1/A^2)")]
    L += [fmt340(*dtr[j]) for j in range(3)]
    L += [ld(" "), ld("  Drr TENSOR (1/A^3)")]
    L += [fmt340(*drr[j]) for j in range(3)]

    def eig_block(t):
        dd = jacobi(t)
        trace = (dd[0] + dd[1] + dd[2]) / 3.
        dl = np.sqrt(abs(dd[0] * dd[0] + dd[1] * dd[1] + dd[2] * dd[2] - dd[0] * dd[1] - dd[0] * dd[2] - dd[1] * dd[2]))
        return dd, trace, dl

    dd, trace, dl = eig_block(drr)
    L += [ld(" "), ld("EIGENVALUES OF Drr TENSOR (1/A^3)"), fmt340(*dd), "  1/3 Trace = " + E(trace, 12, 5),
          "  Anisotropy=" + E(dl, 12, 5), ld(" "), ld("EIGENVALUES OF Drr TENSOR (1/s)"),
          fmt340(runit * dd[0], runit * dd[1], runit * dd[2]),
          "  1/3 Trace DRR(EXACT) =" + E(runit * trace, 15, 5) + " 1/s", ""]
    dd, trace, dl = eig_block(dtt)
    L += [ld("EIGENVALUES OF Dtt TENSOR (1/A)"), fmt340(*dd), "  1/3 Trace = " + E(trace, 12, 5),
          "  Anisotropy=" + E(dl, 12, 5), ld(" "), ld("EIGENVALUES OF Dtt TENSOR (cm^2/s)"),
          fmt340(tunit * dd[0], tunit * dd[1], tunit * dd[2]),
          "  1/3 Trace DTT(EXACT) =" + E(tunit * trace, 15, 5) + " cm^2/s", ""]
    if ivisco:
        what = "INTRINSIC VISCOSITY" if mw == 1 else "VISCOSITY FACTOR"
        if icov == 1:
            L.append(ld(what + " AT THE CENTER OF VISCOSITY"))
            L.append(ld(" Vector between Center of Viscosity/Centroid"))
            L.append("   " + "  Dx = " + E(dvisc[0], 12, 5) + "  Dy = " + E(dvisc[1], 12, 5) + "  Dz = " +
                     E(dvisc[2], 12, 5))
        else:
            L.append(ld(what + " AT THE CENTROID"))
        if iarea == 1:
            with np.errstate(all="ignore"):
                chi = float(np.float64(chiprime) / np.float64(truevol))
            L += ["", "     CHI               " + E(chi, 16, 6)]
            L += ["     CHI*VOLUME (A^3)  " + E(chiprime, 16, 6), ""]
        else:
            L += ["", "     CHI*VOLUME (A^3)  " + E(chiprime, 16, 6), ""]
        if mw == 1:
            with np.errstate(all="ignore"):
                eta_v = float(np.float64(chiprime) * 0.602214 / np.float64(wm))
            L += ["     ETA (cm^3/g)      " + E(eta_v, 16, 6), ""]
    return ("\n".join(L) + "\n").encode()
