"""best-hydro_amd: BEST's boundary-element Stokes solve on one MI355X GPU.

The compute lives in ``libbesthip.so`` (HIP kernels and the C ABI declared in
``include/best_hip.h``), built in-tree by ``build()``.  This module is the host-side mirror of the
reference's call surface for that path:

* ``hm``: the nine ``bhm_*`` functions of ``hmatrix/best_hmatrix.h`` exactly as
  ``best/besolve.f90:48-89`` binds them (one solver per process);
* ``Solver``: a handle with the dense GPU path -- ``tgetgr`` (TGETGR), ``factor`` (the deflation
  and DGETRF of best/best.f:384-468), ``besolv`` (BESOLV) and ``solve_mesh`` (best/best.f:261-498
  without file I/O);
* ``meshes``: BEST's input format and the triangulations of the workloads.

There is no CPU fallback: if the library is missing every call raises.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "libbesthip.so")
_LIB = None

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


def build(jobs: int = 8) -> str:
    """Compiles libbesthip.so for gfx950 (hipcc), in-tree."""
    subprocess.run(["make", "-s", "-C", os.path.join(HERE, "csrc"), f"-j{jobs}"], check=True)
    return LIBPATH


class BhResult(C.Structure):
    _fields_ = [("dtt", C.c_double * 9), ("dtr", C.c_double * 9), ("drr", C.c_double * 9), ("urc", C.c_double * 3),
                ("udr", C.c_double * 3), ("ktt", C.c_double * 9), ("krt", C.c_double * 9), ("krr", C.c_double * 9),
                ("drr_eig", C.c_double * 3), ("dtt_eig", C.c_double * 3), ("dr", C.c_double), ("dtrans", C.c_double),
                ("chiprime", C.c_double), ("chi", C.c_double), ("dvisc", C.c_double * 3), ("afac", C.c_double),
                ("totarea", C.c_double), ("centroid", C.c_double * 3), ("diagm", C.c_double), ("ncomp", C.c_int),
                ("nopen", C.c_int), ("ndef", C.c_int)]

    def as_dict(self):
        out = {}
        for name, _ in self._fields_:
            v = getattr(self, name)
            if isinstance(v, C.Array):
                a = np.frombuffer(v, dtype=np.float64).copy()  # through the buffer protocol, no list
                out[name] = a.reshape(3, 3, order="F") if len(a) == 9 else a
            else:
                out[name] = v
        return out


class BhBeParams(C.Structure):
    _fields_ = [("ver", C.c_char_p), ("iarea", C.c_int), ("truearea", C.c_double), ("truevol", C.c_double),
                ("mw", C.c_int), ("wm", C.c_double), ("nar", C.c_int), ("ivisco", C.c_int), ("icov", C.c_int),
                ("cg", C.c_double)]


class BhStats(C.Structure):
    _fields_ = [("t_prep", C.c_double), ("t_assemble", C.c_double), ("t_factor", C.c_double), ("t_solve", C.c_double),
                ("t_reduce", C.c_double), ("t_total", C.c_double), ("ir_iterations", C.c_int),
                ("ir_correction", C.c_double), ("pivoted", C.c_int), ("npad", C.c_int),
                ("t_adjoint_solve", C.c_double), ("t_adjoint_pairs", C.c_double), ("t_adjoint_tri", C.c_double),
                ("ir_above_tol", C.c_int)]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class BhHmStats(C.Structure):
    _fields_ = [("depth", C.c_int), ("nleaves", C.c_int), ("leaf_min", C.c_int), ("leaf_max", C.c_int),
                ("n_dense", C.c_long), ("n_lowrank", C.c_long), ("n_small", C.c_long), ("n_capped", C.c_long),
                ("bytes_dense", C.c_double), ("bytes_lowrank", C.c_double), ("kmax", C.c_int), ("kmean", C.c_double),
                ("kernel_blocks", C.c_double), ("t_tree", C.c_double), ("t_fill", C.c_double), ("t_pre", C.c_double),
                ("iters_min", C.c_int), ("iters_max", C.c_int), ("steps", C.c_int), ("relres", C.c_double),
                ("t_solve", C.c_double), ("n_trunc_failed", C.c_long), ("kmean_aca", C.c_double),
                ("t_trunc", C.c_double), ("eps_pre", C.c_double), ("piv_ratio", C.c_double), ("pre_levels", C.c_int),
                ("pre_rank_max", C.c_int), ("pre_bytes", C.c_double), ("mem_estimate", C.c_double),
                ("mem_free", C.c_double), ("mem_floor", C.c_double), ("mem_used", C.c_double)]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


def lib():
    """The loaded libbesthip.so (raises if it has not been built)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIBPATH):
        raise RuntimeError(f"{LIBPATH} is missing: build it with best_hydro_amd.build() (hipcc, gfx950)")
    L = C.CDLL(LIBPATH)
    i, d, vp, lg = C.c_int, C.c_double, C.c_void_p, C.c_long
    L.bhm_available.restype = i
    L.bhm_create.argtypes = [i, _dp, _dp, _dp, _dp, d, d, i]
    L.bhm_create.restype = i
    L.bhm_incenters.argtypes = [_dp]
    L.bhm_entries.argtypes = [i, i, i, i, _dp, i]
    L.bhm_diag_mean.restype = d
    L.bhm_build.argtypes = [d, i, d, i, _dp, _ip, _dp, i, i]
    L.bhm_build.restype = i
    L.bhm_solve.argtypes = [i, _dp]
    L.bhm_solve.restype = i
    L.bhm_log_line.argtypes = [C.c_char_p, i]
    L.bhm_log_line.restype = i
    L.bhg_create.argtypes = [i]
    L.bhg_create.restype = vp
    L.bhg_destroy.argtypes = [vp]
    L.bhg_last_error.argtypes = [vp]
    L.bhg_last_error.restype = C.c_char_p
    L.bhg_log_line.argtypes = [vp, C.c_char_p, i]
    L.bhg_log_line.restype = i
    L.bhg_compon.argtypes = [i, _dp, _dp, _dp, _ip, C.POINTER(i), C.POINTER(i)]
    L.bhg_compon.restype = i
    L.bhg_set_mesh.argtypes = [vp, i, _dp, _dp, _dp, d, d, vp]
    L.bhg_set_mesh.restype = i
    L.bhg_set_mesh_file.argtypes = [vp, C.c_char_p, i]
    L.bhg_set_mesh_file.restype = i
    L.bhg_get_components.argtypes = [vp, C.POINTER(i), vp, C.POINTER(i), C.POINTER(i)]
    L.bhg_get_components.restype = i
    L.bhg_series_set_mesh_file.argtypes = [vp, i, C.c_char_p, i]
    L.bhg_series_set_mesh_file.restype = i
    L.bhg_read_best.argtypes = [C.c_char_p, i, i, C.POINTER(i), C.POINTER(d), C.POINTER(d), vp, vp, vp, C.c_char_p, i]
    L.bhg_read_best.restype = i
    L.bhg_write_best.argtypes = [C.c_char_p, i, d, d, _dp, _dp, _dp, C.c_char_p, i]
    L.bhg_write_best.restype = i
    L.bhg_solve_mesh.argtypes = [vp, i, d, i, i, i, C.POINTER(BhResult)]
    L.bhg_solve_mesh.restype = i
    L.bhg_get_solution.argtypes = [vp, i, _dp, lg]
    L.bhg_get_solution.restype = i
    L.bhg_get_stats.argtypes = [vp, C.POINTER(BhStats)]
    L.bhg_get_stats.restype = i
    L.bhg_set_profiling.argtypes = [vp, i]
    L.bhg_set_profiling.restype = i
    L.bhg_get_profile.argtypes = [vp, _dp, _dp, np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")]
    L.bhg_get_profile.restype = i
    L.bhg_set_refinement.argtypes = [vp, d, i]
    L.bhg_set_refinement.restype = i
    L.bhg_load_prepared.argtypes = [vp, i, _dp, _dp, _dp, _dp, d, d]
    L.bhg_load_prepared.restype = i
    L.bhg_incenters.argtypes = [vp, _dp]
    L.bhg_incenters.restype = i
    L.bhg_entries.argtypes = [vp, i, i, i, i, _dp, lg]
    L.bhg_entries.restype = i
    L.bhg_diag_mean.argtypes = [vp]
    L.bhg_diag_mean.restype = d
    L.bhg_tgetgr.argtypes = [vp, _dp, _dp, lg]
    L.bhg_tgetgr.restype = i
    L.bhg_factor_deflated.argtypes = [vp, i, vp, vp, vp, i]
    L.bhg_factor_deflated.restype = i
    L.bhg_besolv.argtypes = [vp, i, _dp, lg]
    L.bhg_besolv.restype = i
    L.bhg_set_option.argtypes = [vp, C.c_char_p, d]
    L.bhg_set_option.restype = i
    L.bhg_series_set_option.argtypes = [vp, C.c_char_p, d]
    L.bhg_series_set_option.restype = i
    L.bhg_set_method.argtypes = [vp, i, d, i, d]
    L.bhg_set_method.restype = i
    L.bhg_set_hm_option.argtypes = [vp, C.c_char_p, d]
    L.bhg_set_hm_option.restype = i
    L.bhg_get_hm_stats.argtypes = [vp, C.POINTER(BhHmStats)]
    L.bhg_get_hm_stats.restype = i
    L.bhg_hm_plan.argtypes = [i, _dp, _dp, _dp, _dp, i, d, np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")]
    L.bhg_hm_plan.restype = i
    L.bhg_series_create.argtypes = [i, i]
    L.bhg_series_create.restype = vp
    L.bhg_series_destroy.argtypes = [vp]
    L.bhg_series_last_error.argtypes = [vp]
    L.bhg_series_last_error.restype = C.c_char_p
    L.bhg_series_set_mesh.argtypes = [vp, i, i, _dp, _dp, _dp, d, d]
    L.bhg_series_set_mesh.restype = i
    L.bhg_series_autotune.argtypes = [vp, i, d, i]
    L.bhg_series_autotune.restype = i
    L.bhg_series_set_tuning.argtypes = [vp, i, i, i, i]
    L.bhg_series_set_tuning.restype = i
    L.bhg_series_get_tuning.argtypes = [vp, i, _ip]
    L.bhg_series_get_tuning.restype = i
    L.bhg_series_solve.argtypes = [vp, i, d, i, i, C.POINTER(BhResult), C.POINTER(d)]
    L.bhg_series_solve.restype = i
    L.bhg_series_get_stats.argtypes = [vp, i, C.POINTER(BhStats)]
    L.bhg_series_get_stats.restype = i
    L.bhg_extrapolate_1n.argtypes = [i, _ip, _dp, C.POINTER(d), C.POINTER(d)]
    L.bhg_extrapolate_1n.restype = i
    L.bhg_hm_time_product.argtypes = [vp, i, i, C.POINTER(d), C.POINTER(d)]
    L.bhg_hm_time_product.restype = i
    L.bhg_hm_apply.argtypes = [vp, i, _dp, lg, _dp, lg]
    L.bhg_hm_apply.restype = i
    L.bhg_hm_blocks.argtypes = [vp, C.c_void_p, lg, C.c_void_p, C.POINTER(lg)]
    L.bhg_hm_blocks.restype = i
    L.bhg_gradient.argtypes = [vp, i, d, i, C.POINTER(BhResult), _dp]
    L.bhg_gradient.restype = i
    L.bhg_format_be.argtypes = [C.POINTER(BhResult), C.POINTER(BhBeParams), C.c_char_p, lg, C.c_char_p, i]
    L.bhg_format_be.restype = lg
    L.bhg_write_be.argtypes = [C.c_char_p, C.POINTER(BhResult), C.POINTER(BhBeParams), C.c_char_p, i]
    L.bhg_write_be.restype = i
    _LIB = L
    return L


def _col(v):
    """NAR x 3 -> BEST's flat column-major buffer."""
    return np.ascontiguousarray(np.asarray(v, dtype=np.float64).T).ravel()


class HipError(RuntimeError):
    pass


class Solver:
    """One GPU solver handle (``bhg_*`` in include/best_hip.h)."""

    def __init__(self, device: int = 0):
        self.L = lib()
        self.h = self.L.bhg_create(int(device))
        if not self.h:
            raise HipError(f"no HIP device {device}")
        self.n = 0

    def close(self):
        if self.h:
            self.L.bhg_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc, what):
        if rc != 0:
            raise HipError(f"{what}: {self.L.bhg_last_error(self.h).decode(errors='replace')}")

    def log(self):
        out, buf = [], C.create_string_buffer(512)
        while True:
            k = self.L.bhg_log_line(self.h, buf, 512)
            if k < 0:
                return out
            out.append(buf.raw[:k].decode(errors="replace"))

    # -- the whole mesh solve
    def set_mesh(self, tri, icomp=None):
        """A triangulation as BEST reads it (meshes.Triangulation or an object with v1, v2, v3,
        truearea, truevol)."""
        self.n = len(tri.v1)
        ic = None
        if icomp is not None:
            self._icomp = np.ascontiguousarray(icomp, dtype=np.int32)
            ic = self._icomp.ctypes.data_as(C.c_void_p)
        self._check(self.L.bhg_set_mesh(self.h, self.n, _col(tri.v1), _col(tri.v2), _col(tri.v3), float(tri.truearea),
                                        float(tri.truevol), ic), "set_mesh")

    def set_mesh_file(self, path, iarea=1):
        """BEST's mesh file read by the library (bhg_set_mesh_file) and made resident, COMPON on
        the device."""
        self._check(self.L.bhg_set_mesh_file(self.h, os.fsencode(path), int(iarea)), "set_mesh_file")
        nar = C.c_int()
        self._check(self.L.bhg_get_components(self.h, C.byref(nar), None, None, None), "get_components")
        self.n = nar.value

    def components(self):
        """(ICOMP, NCOMP, NOPEN) of the resident mesh, from COMPON (bhg_get_components)."""
        ic = np.zeros(self.n, dtype=np.int32)
        nc, no = C.c_int(), C.c_int()
        self._check(self.L.bhg_get_components(self.h, None, ic.ctypes.data_as(C.c_void_p), C.byref(nc), C.byref(no)),
                    "get_components")
        return ic, nc.value, no.value

    def solve_mesh(self, iarea=1, cg=1.0, ivisco=1, icov=0, precision=32, wm=None):
        """best/best.f:261-498 on the resident mesh; returns the .be values and tensors."""
        res = BhResult()
        self._check(self.L.bhg_solve_mesh(self.h, int(iarea), float(cg), int(ivisco), int(icov), int(precision),
                                          C.byref(res)), "solve_mesh")
        out = res.as_dict()
        if wm is not None and ivisco:
            out["eta"] = out["chiprime"] * 0.602214 / wm
        return out

    def gradient(self, iarea=1, cg=1.0, precision=64, wm=None):
        """One solve of the resident mesh (strain flows on, icov 0) and the adjoint gradient.

        Returns (values, grad): values as solve_mesh's; grad[o, i, v, c] = d(output o) /
        d(coordinate c of vertex v of triangle i), outputs D_t (dtrans), D_r (dr), CHI V
        (chiprime), and with wm also [eta] (0.602214 CHI V / wm) as a fourth row."""
        res = BhResult()
        g = np.zeros(27 * self.n)
        self._check(self.L.bhg_gradient(self.h, int(iarea), float(cg), int(precision), C.byref(res), g), "gradient")
        out = res.as_dict()
        grad = g.reshape(3, 3, 3, self.n).transpose(0, 3, 1, 2).copy()
        if wm is not None:
            out["eta"] = out["chiprime"] * 0.602214 / wm
            grad = np.concatenate([grad, grad[2:3] * (0.602214 / wm)])
        return out, grad

    def solution(self, nrhs=11):
        out = np.zeros(3 * self.n * nrhs)
        self._check(self.L.bhg_get_solution(self.h, nrhs, out, 3 * self.n), "solution")
        return out.reshape(nrhs, 3 * self.n).T

    def stats(self):
        s = BhStats()
        self.L.bhg_get_stats(self.h, C.byref(s))
        return s.as_dict()

    # FP32 GEMM launches by tile: gemm32 on 64 x 128 and 128 x 64 (panels, U row panels),
    # gemm32_64 on 64 x 64 (narrow update parts), gemm32_128 on 128 x 128 (wide trailing updates)
    PROFILE_CLASSES = ("pairs", "gemm64", "gemm32", "diag", "trsm", "resid", "other", "gemm32_128", "gemm32_64")

    def set_profiling(self, on=True):
        self.L.bhg_set_profiling(self.h, int(bool(on)))

    def profile(self):
        n = len(self.PROFILE_CLASSES)
        ms, fl = np.zeros(n), np.zeros(n)
        cnt = np.zeros(n, dtype=np.int64)
        self.L.bhg_get_profile(self.h, ms, fl, cnt)
        return {k: dict(ms=float(ms[i]), flops=float(fl[i]), launches=int(cnt[i]))
                for i, k in enumerate(self.PROFILE_CLASSES)}

    def set_refinement(self, tol, maxit=30):
        self.L.bhg_set_refinement(self.h, float(tol), int(maxit))

    def set_option(self, name, value):
        """A setting of the dense path (bhg_set_option): 'fp32_growth', 'ir_steps_before_pivoting'."""
        self._check(self.L.bhg_set_option(self.h, name.encode(), float(value)), f"set_option {name}")

    # -- the H-matrix path
    def set_method(self, method="dense", eps=1e-6, single=False, tol=0.0):
        """'dense' (LU) or 'hmatrix' (cross-approximated blocks + GMRES) for the next solve_mesh."""
        m = {"dense": 0, "hmatrix": 1}[method]
        self._check(self.L.bhg_set_method(self.h, m, float(eps), int(bool(single)), float(tol)), "set_method")

    def set_hm_option(self, name, value):
        """One of the H-matrix settings that are not BEST options (include/best_hip.h)."""
        if self.L.bhg_set_hm_option(self.h, name.encode(), float(value)) != 0:
            raise HipError(f"unknown H-matrix option {name!r}")

    def hm_stats(self):
        s = BhHmStats()
        self.L.bhg_get_hm_stats(self.h, C.byref(s))
        return s.as_dict()

    def hm_time_product(self, nrhs=6, reps=10):
        """(mean ms, algorithmic bytes) of the H-matrix product on device-resident vectors."""
        ms, by = C.c_double(), C.c_double()
        self._check(self.L.bhg_hm_time_product(self.h, int(nrhs), int(reps), C.byref(ms), C.byref(by)), "hm_time_product")
        return ms.value, by.value

    def hm_apply(self, x):
        """(GR + deflation) x with the built H-matrix."""
        x = np.asfortranarray(x, dtype=np.float64)
        nrhs = 1 if x.ndim == 1 else x.shape[1]
        y = np.zeros(3 * self.n * nrhs)
        self._check(self.L.bhg_hm_apply(self.h, nrhs, x.reshape(-1, order="F").copy(), 3 * self.n, y, 3 * self.n),
                    "hm_apply")
        return y.reshape(x.shape, order="F")

    def hm_blocks(self):
        """The leaf blocks of the built H-matrix as an (nblocks, 6) int64 array (1 if low-rank, rank,
        first row, rows, first column, columns; unknowns in tree order) and the tree order of the
        triangles (BEST's triangle, from 0, at each tree position)."""
        cnt = C.c_long()
        self._check(self.L.bhg_hm_blocks(self.h, None, 0, None, C.byref(cnt)), "hm_blocks")
        out = np.zeros((cnt.value, 6), dtype=np.int64)
        perm = np.zeros(self.n, dtype=np.int32)
        self._check(self.L.bhg_hm_blocks(self.h, out.ctypes.data, cnt.value, perm.ctypes.data, C.byref(cnt)),
                    "hm_blocks")
        return out, perm

    # -- the dense path piece by piece (the besolve module's twins)
    def load_prepared(self, n, v1, v2, v3, area, afac, cg=1.0):
        """The mesh as bhm_create receives it: flat column-major shifted vertices, corrected areas."""
        self.n = int(n)
        self._check(self.L.bhg_load_prepared(self.h, self.n, np.ascontiguousarray(v1, dtype=np.float64),
                                             np.ascontiguousarray(v2, dtype=np.float64),
                                             np.ascontiguousarray(v3, dtype=np.float64),
                                             np.ascontiguousarray(area, dtype=np.float64), float(afac), float(cg)),
                    "load_prepared")

    def incenters(self):
        x = np.zeros(3 * self.n)
        self._check(self.L.bhg_incenters(self.h, x), "incenters")
        return x

    def entries(self, i0, m, j0, nc):
        a = np.zeros(m * nc)
        self._check(self.L.bhg_entries(self.h, int(i0), int(m), int(j0), int(nc), a, int(m)), "entries")
        return a.reshape(nc, m).T

    def diag_mean(self):
        return self.L.bhg_diag_mean(self.h)

    def tgetgr(self):
        n3 = 3 * self.n
        x = np.zeros(3 * self.n)
        gr = np.zeros(n3 * n3)
        self._check(self.L.bhg_tgetgr(self.h, x, gr, n3), "tgetgr")
        return x, gr.reshape(n3, n3).T

    def factor(self, precision=32, ndef=1, gdef=None, comp=None, xmu=None):
        keep = []

        def ptr(a, dt):
            if a is None:
                return None
            a = np.ascontiguousarray(a, dtype=dt)
            keep.append(a)
            return a.ctypes.data_as(C.c_void_p)

        self._check(self.L.bhg_factor_deflated(self.h, int(ndef), ptr(gdef, np.float64), ptr(comp, np.int32),
                                               ptr(xmu, np.float64), int(precision)), "factor")

    def besolv(self, b):
        b = np.asfortranarray(b, dtype=np.float64).copy()
        nrhs = 1 if b.ndim == 1 else b.shape[1]
        flat = b.reshape(-1, order="F")
        self._check(self.L.bhg_besolv(self.h, nrhs, flat, 3 * self.n), "besolv")
        return flat.reshape(b.shape, order="F")


class Series:
    """A series of meshes of one body (``bhg_series_*``): every mesh resident, all of them solved
    by one launch of a captured HIP graph (FP32 LU + FP64 refinement), tiles picked by the
    autotuner; ``extrapolate`` fits each result linearly in 1/N."""

    def __init__(self, meshes, device: int = 0):
        self.L = lib()
        self.meshes = list(meshes)
        self.h = self.L.bhg_series_create(int(device), len(self.meshes))
        if not self.h:
            raise HipError(f"no HIP device {device}")
        for k, t in enumerate(self.meshes):
            self._check(self.L.bhg_series_set_mesh(self.h, k, len(t.v1), _col(t.v1), _col(t.v2), _col(t.v3),
                                                   float(t.truearea), float(t.truevol)), "set_mesh")

    def close(self):
        if self.h:
            self.L.bhg_series_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc, what):
        if rc != 0:
            raise HipError(f"{what}: {self.L.bhg_series_last_error(self.h).decode(errors='replace')}")

    def set_mesh_file(self, k, path, iarea=1):
        """Mesh k from BEST's mesh file (bhg_series_set_mesh_file)."""
        self._check(self.L.bhg_series_set_mesh_file(self.h, int(k), os.fsencode(path), int(iarea)), "set_mesh_file")

    def set_option(self, name, value):
        """bhg_set_option for every mesh (bhg_series_set_option)."""
        self._check(self.L.bhg_series_set_option(self.h, name.encode(), float(value)), f"set_option {name}")

    def autotune(self, iarea=1, cg=1.0, ivisco=1):
        self._check(self.L.bhg_series_autotune(self.h, int(iarea), float(cg), int(ivisco)), "autotune")
        return self.tuning()

    def set_tuning(self, k, strip=64, gemm_bk=0, ir_iters=3):
        self._check(self.L.bhg_series_set_tuning(self.h, int(k), int(strip), int(gemm_bk), int(ir_iters)), "set_tuning")

    def tuning(self):
        out = []
        for k in range(len(self.meshes)):
            t = np.zeros(4, dtype=np.int32)
            self.L.bhg_series_get_tuning(self.h, k, t)
            out.append(dict(strip=int(t[0]), gemm_bk=int(t[1]), ir_iters=int(t[2]), parallel=int(t[3])))
        return out

    def solve(self, iarea=1, cg=1.0, ivisco=1, icov=0, wm=None):
        """Results per mesh (dicts as Solver.solve_mesh) and the device time of the launch (ms)."""
        res = (BhResult * len(self.meshes))()
        ms = C.c_double()
        self._check(self.L.bhg_series_solve(self.h, int(iarea), float(cg), int(ivisco), int(icov), res, C.byref(ms)),
                    "solve")
        out = []
        for r in res:
            d = r.as_dict()
            if wm is not None and ivisco:
                d["eta"] = d["chiprime"] * 0.602214 / wm
            out.append(d)
        return out, ms.value

    def stats(self, k):
        s = BhStats()
        self.L.bhg_series_get_stats(self.h, int(k), C.byref(s))
        return s.as_dict()


def extrapolate_1n(n, y):
    """(intercept, slope) of the least-squares line y = intercept + slope / N (bhg_extrapolate_1n)."""
    n = np.ascontiguousarray(n, dtype=np.int32)
    y = np.ascontiguousarray(y, dtype=np.float64)
    a, b = C.c_double(), C.c_double()
    if lib().bhg_extrapolate_1n(len(n), n, y, C.byref(a), C.byref(b)) != 0:
        raise HipError("extrapolate_1n: fewer than two meshes or equal N")
    return a.value, b.value


def read_best_file(path, iarea=1):
    """The mesh file as the library reads it (bhg_read_best: best/best.f:189-220) ->
    meshes.Triangulation."""
    from . import meshes
    L = lib()
    nar, ta, tv = C.c_int(), C.c_double(), C.c_double()
    err = C.create_string_buffer(512)
    if L.bhg_read_best(os.fsencode(path), int(iarea), 0, C.byref(nar), C.byref(ta), C.byref(tv), None, None, None,
                       err, 512) != 0:
        raise HipError(err.value.decode(errors="replace"))
    n = nar.value
    v = [np.zeros(3 * n) for _ in range(3)]
    if L.bhg_read_best(os.fsencode(path), int(iarea), n, C.byref(nar), C.byref(ta), C.byref(tv),
                       *[a.ctypes.data_as(C.c_void_p) for a in v], err, 512) != 0:
        raise HipError(err.value.decode(errors="replace"))
    f = lambda a: a.reshape(3, n).T.copy()
    return meshes.Triangulation(f(v[0]), f(v[1]), f(v[2]), ta.value, tv.value, os.path.basename(path))


def write_best_file(path, tri):
    """The mesh as rcoal writes it (bhg_write_best: list-directed header, then 9e17.8 lines)."""
    err = C.create_string_buffer(512)
    if lib().bhg_write_best(os.fsencode(path), len(tri.v1), float(tri.truearea), float(tri.truevol), _col(tri.v1),
                            _col(tri.v2), _col(tri.v3), err, 512) != 0:
        raise HipError(err.value.decode(errors="replace"))


BEST_VERSION = "4.0 SFSU/ARAGON"  # best/best.f:180


def _be_args(result, nar, truearea, truevol, iarea, mw, wm, ivisco, icov, cg, ver):
    r = BhResult()
    for name, _ in BhResult._fields_:
        if name not in result:
            continue
        v = result[name]
        if isinstance(getattr(r, name), C.Array):
            a = np.asarray(v, dtype=np.float64)
            a = a.reshape(-1, order="F") if a.ndim == 2 else a.reshape(-1)
            getattr(r, name)[:] = [float(t) for t in a]
        else:
            setattr(r, name, type(getattr(r, name))(v))
    p = BhBeParams(ver.encode(), int(iarea), float(truearea), float(truevol), int(mw), float(wm), int(nar),
                   int(ivisco), int(icov), float(cg))
    return r, p


def format_be(result, nar, truearea=0.0, truevol=0.0, iarea=1, mw=0, wm=0.0, ivisco=1, icov=0, cg=1.0,
              ver=BEST_VERSION) -> bytes:
    """The .be file BEST's OUTPUT (best/output.f) writes for a solve's result (a ``solve_mesh``
    dict), byte for byte (bhg_format_be).  Raises HipError when the tensors or the viscosity
    values are not finite, as BEST stops (best/best.f:483-496)."""
    r, p = _be_args(result, nar, truearea, truevol, iarea, mw, wm, ivisco, icov, cg, ver)
    err = C.create_string_buffer(256)
    n = lib().bhg_format_be(C.byref(r), C.byref(p), None, 0, err, 256)
    if n < 0:
        raise HipError(err.value.decode())
    buf = C.create_string_buffer(n + 1)
    lib().bhg_format_be(C.byref(r), C.byref(p), buf, n + 1, err, 256)
    return buf.raw[:n]


def write_be(path, result, nar, truearea=0.0, truevol=0.0, iarea=1, mw=0, wm=0.0, ivisco=1, icov=0, cg=1.0,
             ver=BEST_VERSION) -> None:
    """format_be's text written to path (bhg_write_be)."""
    r, p = _be_args(result, nar, truearea, truevol, iarea, mw, wm, ivisco, icov, cg, ver)
    err = C.create_string_buffer(256)
    if lib().bhg_write_be(os.fsencode(path), C.byref(r), C.byref(p), err, 256) != 0:
        raise HipError(err.value.decode())


def hm_plan(v1, v2, v3, x, leaf=32, eta=2.0):
    """Cluster and block trees on the host: depth, leaf clusters, smallest and largest leaf, dense
    leaf blocks, admissible blocks (low-rank), small admissible blocks.  v1, v2, v3, x: NAR x 3."""
    out = np.zeros(7, dtype=np.int64)
    n = len(v1)
    rc = lib().bhg_hm_plan(n, _col(v1), _col(v2), _col(v3), _col(x), int(leaf), float(eta), out)
    if rc != 0:
        raise HipError("hm_plan failed")
    keys = ("depth", "nleaves", "leaf_min", "leaf_max", "n_dense", "n_lowrank", "n_small")
    return {k: int(v) for k, v in zip(keys, out)}


class _HM:
    """The nine functions of hmatrix/best_hmatrix.h, as best/besolve.f90 calls them."""

    @staticmethod
    def available():
        return lib().bhm_available()

    @staticmethod
    def create(n, v1, v2, v3, area, afac, cg, threads=0):
        return lib().bhm_create(int(n), v1, v2, v3, area, float(afac), float(cg), int(threads))

    @staticmethod
    def incenters(n):
        x = np.zeros(3 * n)
        lib().bhm_incenters(x)
        return x

    @staticmethod
    def entries(i0, m, j0, nc):
        a = np.zeros(m * nc)
        lib().bhm_entries(int(i0), int(m), int(j0), int(nc), a, int(m))
        return a.reshape(nc, m).T

    @staticmethod
    def diag_mean():
        return lib().bhm_diag_mean()

    @staticmethod
    def build(eps, direct, eps_pre, ndef, gdef, comp, xmu, single, nrhs):
        return lib().bhm_build(float(eps), int(direct), float(eps_pre), int(ndef),
                               np.ascontiguousarray(gdef, dtype=np.float64), np.ascontiguousarray(comp, dtype=np.int32),
                               np.ascontiguousarray(xmu, dtype=np.float64), int(single), int(nrhs))

    @staticmethod
    def solve(b):
        b = np.asfortranarray(b, dtype=np.float64).copy()
        nrhs = 1 if b.ndim == 1 else b.shape[1]
        flat = b.reshape(-1, order="F")
        rc = lib().bhm_solve(nrhs, flat)
        return rc, flat.reshape(b.shape, order="F")

    @staticmethod
    def log_lines():
        out, buf = [], C.create_string_buffer(512)
        while True:
            k = lib().bhm_log_line(buf, 512)
            if k < 0:
                return out
            out.append(buf.raw[:k].decode(errors="replace"))

    @staticmethod
    def destroy():
        lib().bhm_destroy()


hm = _HM()
