"""Mesh solves/s of BEST's boundary-element Stokes solve on one MI355X.

One step = one whole mesh solve of the 5,120-triangle bumpy ellipsoid (semi-axes 30/20/15 A, an
l <= 6 radial bump field of sd 0.03, level-4 icosphere; 15,360 unknowns) with the vertices already
resident in HBM: centroid shift and areas, incenters and self blocks, the 26.2 M pair blocks with
the deflation, LU (FP32 on MFMA, FP64 iterative refinement), the 11 right-hand sides of -v, and
the DIFCO/VISCO reductions to D_t, D_r, the coupling tensor and CHI*V ([eta]).

Prints one JSON line.  Besides the contract fields it carries the FP64-LU path's rate on the same
mesh, the errors of both paths against the oracle (the reference's dense FP64 path restated in C
with LAPACK), the roofline of the dominant kernel class (HIP events around every launch of the
class during the timed steps), each phase of a solve against its own roofline (phase_rooflines:
assembly on the FP64 VALU, the LU on the matrix cores, the solves and residuals on HBM), the
oracle's CPU rate on this host, and the H-matrix path on the
81,920-triangle sphere (hmatrix_c5: solves/s, build and GMRES times, and the HBM roofline of the
H-matrix product), the series of eight meshes for the 1/N extrapolation solved by one captured
graph launch (series_c3), and the adjoint gradient of D_t, D_r and CHI*V with respect to the
vertex coordinates on the 5,120-triangle mesh and the two-surface mesh (adjoint_c4).
"""
import argparse
import importlib.util
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))

METRIC = "mesh solves/s: FP32-MFMA LU+FP64 IR vs FP64 LU at 1e-10 rel err in D,[eta]"
PEAKS = {  # MI355X published dense peaks
    "fp64_valu_tflops": 78.6, "fp32_mfma_tflops": 157.3, "fp64_mfma_tflops": 78.6, "hbm_gbs": 8000.0,
}
PAIR_FLOP = 3457.0  # FP64 flops of one triangle pair's 3x3 integral (special functions counted once)


def _package():
    if "best_hydro_amd" in sys.modules:
        return sys.modules["best_hydro_amd"]
    d = os.path.join(ROOT, "best-hydro_amd")
    spec = importlib.util.spec_from_file_location("best_hydro_amd", os.path.join(d, "__init__.py"),
                                                  submodule_search_locations=[d])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["best_hydro_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def _cores():
    """Threads for the CPU baseline: the CPUs this process may run on, capped by the CPU share
    the host grants (OMP_NUM_THREADS where it is set; the GPU hosts set it to that share)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    share = os.environ.get("OMP_NUM_THREADS", "")
    if share.isdigit() and int(share) > 0:
        n = min(n, int(share))
    return max(1, n)


def _host_cpu():
    """CPU model and core counts from lscpu (falls back to /proc/cpuinfo)."""
    info = {}
    try:
        txt = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in txt.splitlines():
            k, _, v = line.partition(":")
            info[k.strip()] = v.strip()
    except Exception:
        pass
    model = info.get("Model name", "")
    if not model:
        try:
            with open("/proc/cpuinfo") as f:
                model = next((l.split(":", 1)[1].strip() for l in f if l.startswith("model name")), "")
        except OSError:
            model = ""
    out = {"model": model}
    try:
        per = int(info.get("Core(s) per socket", "0")) * int(info.get("Socket(s)", "0"))
        if per:
            out["physical_cores"] = per
        if info.get("CPU(s)"):
            out["logical_cpus"] = int(info["CPU(s)"])
    except ValueError:
        pass
    return out


def cpu_baseline(tri, cores, runs=3):
    """The oracle on this host: whole solves of the same mesh (C assembly with OpenMP, LAPACK
    dgetrf/dgetrs through SciPy's OpenBLAS, C DIFCO/VISCO); the median wall time of `runs`."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    O.lib()
    times, ref = [], None
    for _ in range(runs):
        t0 = time.perf_counter()
        ref = O.best_run(tri.v1, tri.v2, tri.v3, tri.truearea, tri.truevol, iarea=1, cg=1.0, ivisco=1, icov=0)
        times.append(time.perf_counter() - t0)
    return ref, float(np.median(times)), times


def hmatrix_c5(bh, M, steps):
    """The H-matrix path on the 81,920-triangle bumpy sphere (245,760 unknowns; dense GR would be
    483 GB): cross approximation at eps 1e-6, GMRES on 6 right-hand sides to 1e-10.  One step is
    a whole solve (prep, H-matrix build, GMRES, reductions) with the mesh resident."""
    tri = M.hmatrix_sphere_mesh()
    s = bh.Solver(0)
    s.set_mesh(tri)
    s.set_method("hmatrix", eps=1e-6, tol=1e-10)
    s.solve_mesh(ivisco=0)  # warm
    t0 = time.perf_counter()
    for _ in range(steps):
        r = s.solve_mesh(ivisco=0)
    el = (time.perf_counter() - t0) / steps
    st = s.hm_stats()
    ms, nbytes = s.hm_time_product(6, 10)
    gbs = nbytes / (ms * 1e-3) / 1e9
    s.close()
    return {"workload": "40 A sphere x (1 + l<=8 bumps sd 0.05), level-6 icosphere, 81,920 triangles, 245,760 unknowns, "
                        "eps 1e-6, 6 RHS, GMRES to 1e-10",
            "value": round(1.0 / el, 4), "unit": "solves/s", "s_per_solve": round(el, 3), "steps": steps,
            "build_s": round(st["t_tree"] + st["t_fill"] + st["t_pre"], 3), "gmres_s": round(st["t_solve"], 3),
            "gmres_iterations": [st["iters_min"], st["iters_max"]], "relres": st["relres"],
            "h_gb": round((st["bytes_dense"] + st["bytes_lowrank"]) / 1e9, 2), "rank_max": st["kmax"],
            "rank_mean": round(st["kmean"], 1), "rank_mean_before_recompression": round(st["kmean_aca"], 1),
            "recompression_s": round(st["t_trunc"], 3), "blocks_not_recompressed": st["n_trunc_failed"],
            "dtrans": r["dtrans"], "dr": r["dr"],
            # the build against the FP64 VALU roofline of its kernel evaluations: every 3x3 block the
            # cross approximations and the dense blocks evaluate, at the pair integral's 3,457 flop
            # (special functions counted once), over the build's fill time (recompression inside)
            "build_roofline": {"kernel_blocks": st["kernel_blocks"], "flop": st["kernel_blocks"] * PAIR_FLOP,
                               "fill_s": round(st["t_fill"], 3),
                               "achieved": round(st["kernel_blocks"] * PAIR_FLOP / st["t_fill"] / 1e12, 2),
                               "peak": PEAKS["fp64_valu_tflops"], "unit": "TFLOP/s",
                               "frac": round(st["kernel_blocks"] * PAIR_FLOP / st["t_fill"] / 1e12 / PEAKS["fp64_valu_tflops"], 4)},
            "roofline": {"kernel": "H-matrix product (k_hm_vtx + k_hm_rows + sums)", "bound": "hbm",
                         "achieved": round(gbs, 1), "peak": PEAKS["hbm_gbs"], "unit": "GB/s",
                         "frac": round(gbs / PEAKS["hbm_gbs"], 4), "avg_product_ms": round(ms, 4),
                         "bytes_per_product": nbytes}}


def series_c3(bh, M, steps):
    """The 1/N series: the bumpy ellipsoid remeshed at 8 uneven sizes (2,500 ... 7,000 triangles),
    all meshes resident; the autotuner picks the tiles and captures the schedule; one step is one
    launch of the captured graph (every mesh's FP32-LU + FP64-IR solve) plus the per-mesh checks
    and reductions; then the linear fits in 1/N."""
    tris = M.series_meshes()
    sizes = [t.nar for t in tris]
    ser = bh.Series(tris)
    t0 = time.perf_counter()
    tune = ser.autotune(ivisco=1)
    t_tune = time.perf_counter() - t0
    ser.solve(ivisco=1)  # warm
    t0 = time.perf_counter()
    gms = []
    for _ in range(steps):
        res, ms = ser.solve(ivisco=1)
        gms.append(ms)
    el = (time.perf_counter() - t0) / steps
    fits = {}
    for key in ("dtrans", "dr", "chiprime"):
        a, b = bh.extrapolate_1n(sizes, [r[key] for r in res])
        fits[key] = {"intercept": a, "slope": b}
    pivoted = [ser.stats(k)["pivoted"] for k in range(len(tris))]
    ser.close()
    return {"workload": "bumpy ellipsoid remeshed by random-point Delaunay at N = " + ", ".join(map(str, sizes)),
            "value": round(len(tris) / el, 4), "unit": "mesh solves/s", "s_per_series": round(el, 4),
            "graph_ms": round(min(gms), 3), "steps": steps, "autotune_s": round(t_tune, 2),
            "tuning": tune, "pivoted": pivoted, "fit_1n": fits}


def adjoint_c4(bh, M, steps):
    """The adjoint: one solve (FP32 LU + FP64 IR) and the gradient of D_t, D_r and CHI*V with
    respect to every vertex coordinate, on the 5,120-triangle bumpy ellipsoid and on the
    two-surface mesh (that ellipsoid plus an 80-triangle 2 A sphere 12 A away, two deflation
    terms).  The pair term of the gradient (k_adj_pairs) is timed with HIP events."""
    out = {}
    for name, tri in (("bumpy_ellipsoid_5120", M.bumpy_ellipsoid_mesh()), ("two_surface_5200", M.two_surface_mesh())):
        s = bh.Solver(0)
        s.set_mesh(tri)
        s.gradient(precision=32)  # warm
        t0 = time.perf_counter()
        for _ in range(steps):
            s.solve_mesh(precision=32)
        t_solve = (time.perf_counter() - t0) / steps
        t0 = time.perf_counter()
        for _ in range(steps):
            r, g = s.gradient(precision=32)
        t_all = (time.perf_counter() - t0) / steps
        st = s.stats()
        V, idx = M.vertex_index(tri)
        gv = M.vertex_gradient(g, idx, len(V))
        s.close()
        out[name] = {"triangles": tri.nar, "vertices": len(V), "coordinates": 3 * len(V),
                     "solve_ms": round(1e3 * t_solve, 2), "solve_and_gradient_ms": round(1e3 * t_all, 2),
                     "gradient_ms": round(1e3 * (t_all - t_solve), 2),
                     "adjoint_pairs_ms": round(st["t_adjoint_pairs"], 2),
                     "grad_norm": {"dtrans": float(np.linalg.norm(gv[0])), "dr": float(np.linalg.norm(gv[1])),
                                   "chiprime": float(np.linalg.norm(gv[2]))}}
    return out


def counter_pass(s, M, steps):
    """The command the HBM traffic figures come from (rocprofv3 --pmc FETCH_SIZE, then WRITE_SIZE):
    mixed solves of the workload, then three device-to-device copies of 1 GiB, whose counter values
    give the scale of each counter (bytes moved / bytes counted)."""
    import ctypes as C
    for k in range(steps):
        r = s.solve_mesh(iarea=1, cg=1.0, ivisco=1, icov=0, precision=32)
        print(f"counter pass: solve {k}: D_t {r['dtrans']:.10e}", flush=True)
    s.close()
    # the HIP runtime the library already loaded (its device-to-device copy kernel)
    with open("/proc/self/maps") as f:
        path = next(line.split()[-1] for line in f if "libamdhip64" in line)
    hip = C.CDLL(path)
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipFree.argtypes = [C.c_void_p]
    nbytes = 1 << 30
    a, b = C.c_void_p(), C.c_void_p()
    for rc in (hip.hipMalloc(C.byref(a), nbytes), hip.hipMalloc(C.byref(b), nbytes), hip.hipMemset(a, 0, nbytes)):
        if rc != 0:
            raise RuntimeError(f"HIP error {rc} in the calibration copy")
    for _ in range(3):
        if hip.hipMemcpy(b, a, nbytes, 3) != 0:  # hipMemcpyDeviceToDevice
            raise RuntimeError("hipMemcpy failed")
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)
    print("counter pass: three 1 GiB copies done", flush=True)


def valu_pass(s, M):
    """The command the assembly's VALU counters come from (rocprofv3 --pmc SQ_...): one mixed solve
    of the workload (k_pairs: 5,120 x 5,120 pairs) and one gradient (a solve, then k_adj_pairs),
    after a warm-up of each."""
    for k in range(2):
        r = s.solve_mesh(iarea=1, cg=1.0, ivisco=1, icov=0, precision=32)
        print(f"valu pass: solve {k}: D_t {r['dtrans']:.10e}", flush=True)
    for k in range(2):
        r, _ = s.gradient(precision=32)
        print(f"valu pass: gradient {k}: D_t {r['dtrans']:.10e}", flush=True)
    s.close()


def traffic_from_profile(kernel_class):
    """Calibrated HBM bytes per launch of the kernel class, from the committed counter summary
    (profiles/*_traffic.md, a JSON block written from two rocprofv3 --pmc runs), or None."""
    import glob
    import re
    for path in sorted(glob.glob(os.path.join(ROOT, "profiles", "*_traffic.md"))):
        with open(path) as f:
            m = re.search(r"```json\n(.*?)\n```", f.read(), re.S)
        if not m:
            continue
        rec = json.loads(m.group(1))
        if rec.get("kernel_class") == kernel_class:
            return rec["fetch_bytes_per_launch"] + rec["write_bytes_per_launch"], os.path.basename(path)
    return None, None


def valu_from_profile():
    """The VALU counter records of k_pairs and k_adj_pairs from the committed summary
    (profiles/*_valu.md, a JSON block written from rocprofv3 --pmc runs of --valu-pass), or {}."""
    import glob
    import re
    for path in sorted(glob.glob(os.path.join(ROOT, "profiles", "*_valu.md"))):
        with open(path) as f:
            m = re.search(r"```json\n(.*?)\n```", f.read(), re.S)
        if m:
            rec = json.loads(m.group(1))
            for v in rec.values():
                v["source"] = os.path.basename(path)
            return rec
    return {}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true", help="skip the oracle CPU baseline and error report")
    ap.add_argument("--no-hmatrix", action="store_true", help="skip the H-matrix path's 81,920-triangle case")
    ap.add_argument("--no-series", action="store_true", help="skip the 8-mesh series")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint gradient")
    ap.add_argument("--counter-pass", action="store_true",
                    help="for rocprofv3 --pmc runs: --steps mixed solves of the workload, then three 1 GiB device "
                         "copies (the calibration of FETCH_SIZE / WRITE_SIZE); prints no JSON line")
    ap.add_argument("--valu-pass", action="store_true",
                    help="for rocprofv3 --pmc runs: two mixed solves and two gradients of the workload (the VALU "
                         "counters of the assembly and of the adjoint's pair term); prints no JSON line")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {args.gpus} is not supported (the solve does not shard across "
              "GPUs; see DESIGN.md for how it would)", file=sys.stderr)
        sys.exit(2)

    cores = _cores()
    os.environ.setdefault("OMP_NUM_THREADS", str(cores))
    os.environ.setdefault("OPENBLAS_NUM_THREADS", str(cores))
    bh = _package()
    from best_hydro_amd import meshes as M

    tri = M.bumpy_ellipsoid_mesh()
    n = tri.nar
    s = bh.Solver(0)
    s.set_mesh(tri)  # inputs resident in HBM before any timing
    if args.counter_pass:
        counter_pass(s, M, args.steps)
        return
    if args.valu_pass:
        valu_pass(s, M)
        return

    def run(prec, steps, profile=False):
        s.set_profiling(profile)
        t0 = time.perf_counter()
        res = None
        for _ in range(steps):
            res = s.solve_mesh(iarea=1, cg=1.0, ivisco=1, icov=0, precision=prec)
        el = time.perf_counter() - t0
        prof = s.profile() if profile else None
        s.set_profiling(False)
        return res, el, prof, s.stats()

    for _ in range(args.warmup):
        run(32, 1)
        run(64, 1)
    # The timed steps: the metric's path (FP32 LU + FP64 refinement), then the FP64-LU comparator.
    r32, el32, _, st32 = run(32, args.steps)
    r64, el64, _, st64 = run(64, args.steps)
    ms32 = 1e3 * el32 / args.steps
    ms64 = 1e3 * el64 / args.steps
    # The same steps again with HIP events around every launch of the timed kernel classes (on
    # the stream each kernel runs on), for the roofline; the events cost a few percent, so this
    # pass is kept out of the value above.
    _, elp, prof32, _ = run(32, args.steps, profile=True)
    # one profiled step of the FP64-LU comparator: its GEMM kernel's own rate
    _, _, prof64, _ = run(64, 1, profile=True)
    g64 = prof64["gemm64"]
    fp64_gemm = None
    if g64["launches"]:
        fp64_gemm = {"kernel": "k_gemm<double, 16, *, *> (FP64 LU: panels and trailing updates; 2 waves per SIMD)",
                     "achieved": round(g64["flops"] / (g64["ms"] * 1e-3) / 1e12, 3), "peak": PEAKS["fp64_mfma_tflops"],
                     "unit": "TFLOP/s", "launches": g64["launches"], "ms_per_step": round(g64["ms"], 3)}
        fp64_gemm["frac"] = round(fp64_gemm["achieved"] / fp64_gemm["peak"], 4)

    # Dominant kernel class of the mixed path, by measured time.
    flop_classes = {"pairs": ("fp64_valu_tflops", "TFLOP/s"), "gemm32": ("fp32_mfma_tflops", "TFLOP/s"),
                    "gemm32_128": ("fp32_mfma_tflops", "TFLOP/s"), "gemm32_64": ("fp32_mfma_tflops", "TFLOP/s"),
                    "gemm64": ("fp64_mfma_tflops", "TFLOP/s"), "diag": ("fp32_mfma_tflops", "TFLOP/s"),
                    "trsm": ("fp32_mfma_tflops", "TFLOP/s"), "resid": ("fp64_valu_tflops", "TFLOP/s")}
    dom = max((k for k in prof32 if prof32[k]["launches"] > 0 and k in flop_classes), key=lambda k: prof32[k]["ms"])
    p = prof32[dom]
    avg_ms = p["ms"] / p["launches"]
    flops_per_launch = p["flops"] / p["launches"]
    achieved = flops_per_launch / (avg_ms * 1e-3) / 1e12
    peak = PEAKS[flop_classes[dom][0]]
    traffic, traffic_src = traffic_from_profile(dom)
    kernels = {"gemm32_128": "k_gemm<float, 16, 128, 128> (FP32 trailing updates, 128 x 128 tiles)",
               "gemm32": "k_gemm<float, 32, 64, 128> and <float, 32, 128, 64> (FP32 panels and U row panels)",
               "gemm32_64": "k_gemm<float, 32, 64, 64> and k_gemm_multi<float, 32> (FP32 narrow update parts, 64 x 64 tiles)",
               "pairs": "k_pairs<true> (assembly)", "diag": "k_lu_diag<float>"}
    roofline = {"bound": "mfma" if dom.startswith("gemm") else "valu", "kernel_class": dom,
                "kernel": kernels.get(dom, dom),
                "achieved": round(achieved, 3), "peak": peak, "unit": "TFLOP/s", "frac": round(achieved / peak, 4),
                "traffic": traffic, "traffic_source": traffic_src,
                "avg_launch_ms": round(avg_ms, 5), "launches_per_step": p["launches"] / args.steps,
                "flops_per_launch": flops_per_launch,
                "share_of_step": round(p["ms"] / args.steps / ms32, 3),
                "profiled_ms_per_step": round(1e3 * elp / args.steps, 3),
                "classes_ms_per_step": {k: round(v["ms"] / args.steps, 3) for k, v in prof32.items() if v["launches"]}}

    # Each phase of one solve against the roofline that bounds it, from the solver's phase timers
    # (HIP events) and the algorithmic work per unit: assembly 3,457 flops per triangle pair
    # (special functions counted once) on the FP64 VALU; the LU 2/3 npad^3 on the matrix cores; the
    # triangular solves and refinement residuals as the bytes of the factors and of GR they stream.
    npad = st32["npad"]
    lu_flops = 2.0 / 3.0 * npad ** 3
    pair_flops = PAIR_FLOP * n * (n - 1.0)
    super_blk = 1024
    f32_solve_bytes = 4.0 * (npad * npad + 2 * npad * super_blk)   # L, U and the super-block inverses
    f64_solve_bytes = 8.0 * (npad * npad + 2 * npad * super_blk)
    resid_bytes = 8.0 * npad * npad
    nsolve = st32["ir_iterations"]

    def phase(work, ms, peak, unit):
        achieved = work / (ms * 1e-3) / (1e12 if unit == "TFLOP/s" else 1e9)
        return {"achieved": round(achieved, 2), "peak": peak, "unit": unit, "frac": round(achieved / peak, 4),
                "ms": round(ms, 3)}
    valu = valu_from_profile()
    phase_rooflines = {
        "assembly_fp64_valu": dict(phase(pair_flops, st32["t_assemble"], PEAKS["fp64_valu_tflops"], "TFLOP/s"),
                                   counters=valu.get("k_pairs")),
        "lu_fp32_mfma": phase(lu_flops, st32["t_factor"], PEAKS["fp32_mfma_tflops"], "TFLOP/s"),
        "lu_fp64_mfma": phase(lu_flops, st64["t_factor"], PEAKS["fp64_mfma_tflops"], "TFLOP/s"),
        "solves_and_refinement_hbm": phase(nsolve * f32_solve_bytes + (nsolve - 1) * resid_bytes, st32["t_solve"],
                                           PEAKS["hbm_gbs"], "GB/s"),
        "solve_fp64_hbm": phase(f64_solve_bytes, st64["t_solve"], PEAKS["hbm_gbs"], "GB/s"),
        "note": "factor phases include the super-block inverses of the solves; the solve phase of the mixed "
                f"path is {nsolve} FP32 triangular solves and {nsolve - 1} FP64 residual sweeps",
    }

    out = {
        "metric": METRIC, "value": round(1e3 / ms32, 4), "unit": "solves/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms32, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32+f64", "data": "synthetic",
        "config": {"workload": "bumpy ellipsoid 30/20/15 A, l<=6 radial bumps sd 0.03 (seeded), level-4 icosphere",
                   "triangles": n, "unknowns": 3 * n, "rhs": 11, "lu": "FP32 MFMA + FP64 iterative refinement",
                   "parallelism": "single GPU"},
        "fp64_lu": {"value": round(1e3 / ms64, 4), "unit": "solves/s", "ms_per_step": round(ms64, 3),
                    "gemm": fp64_gemm},
        "ratio_mixed_over_fp64": round(ms64 / ms32, 3),
        "phases_ms": {k: round(v, 3) for k, v in st32.items() if k.startswith("t_") and not k.startswith("t_adjoint")},
        "ir_iterations": st32["ir_iterations"],
        "roofline": roofline,
        "phase_rooflines": phase_rooflines,
    }
    if not args.no_cpu:
        ref, el, times = cpu_baseline(tri, cores)
        def err(a, b):
            return abs(a - b) / abs(b)
        errors = {}
        for name, r in (("mixed", r32), ("fp64", r64)):
            errors[name] = {"dtrans": err(r["dtrans"], ref["dtrans"]), "dr": err(r["dr"], ref["dr"]),
                            "chiprime_eta": err(r["chiprime"], ref["chiprime"]),
                            "coupling_dtr": float(np.max(np.abs(r["dtr"] - ref["dtr"])) / np.max(np.abs(ref["dtr"])))}
        out["errors_vs_fp64_oracle"] = errors
        host = _host_cpu()
        out["cpu_baseline"] = {"value": round(1.0 / el, 5), "unit": "solves/s", "cores": cores, "kind": "port",
                               "host_cpu": host, "runs_s": [round(t, 3) for t in times],
                               "sample": f"whole {n}-triangle solves of the same mesh by the oracle (C assembly, "
                                         f"OpenMP {cores} threads; LAPACK dgetrf/dgetrs via SciPy OpenBLAS with "
                                         f"{cores} threads; C DIFCO/VISCO), median of {len(times)}: {el:.2f} s"}
    s.close()
    if not args.no_series:
        out["series_c3"] = series_c3(bh, M, max(1, min(args.steps, 5)))
    if not args.no_adjoint:
        out["adjoint_c4"] = adjoint_c4(bh, M, max(1, min(args.steps, 5)))
        out["adjoint_c4"]["k_adj_pairs_counters"] = valu.get("k_adj_pairs")
    if not args.no_hmatrix:
        out["hmatrix_c5"] = hmatrix_c5(bh, M, max(1, min(args.steps, 3)))
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
