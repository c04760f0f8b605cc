"""Writes c5_reference_hmatrix.json: the reference's own H-matrix library (oracle/_ref, built from
hmatrix/src by oracle/Makefile) solving the 81,920-triangle bumpy sphere (245,760 unknowns) on
the CPU at eps 1e-6, GMRES to 1e-10, the 6 DIFCO right-hand sides; the oracle's C restatement of
prep, COMPON, the deflation weights, the right-hand sides and the DIFCO reductions around it, as
BEST calls the library (best/besolve.f90:178-245, best/difco.f:28-406).  The dense oracle cannot
run at this size (GR would be 483 GB), so this is the CPU reference of the H-matrix path at full
size.

    python tests/golden/make_c5_reference.py [--threads N]

Takes some minutes and about 30 GB of memory; records the timings and the reference's log too.
"""
import argparse
import importlib.util
import json
import os
import platform
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def meshes():
    d = os.path.join(ROOT, "best-hydro_amd")
    spec = importlib.util.spec_from_file_location("bh_meshes", os.path.join(d, "meshes.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--threads", type=int, default=os.cpu_count() or 1)
    ap.add_argument("--eps", type=float, default=1e-6)
    ap.add_argument("--level", type=int, default=6, help="icosphere level (6: the 81,920-triangle case)")
    ap.add_argument("--out", default=os.path.join(HERE, "c5_reference_hmatrix.json"))
    args = ap.parse_args()
    os.environ["BEST_HMATRIX_GMRES_TOL"] = "1e-10"
    import oracle as O
    R = O.reference_lib()
    if R is None:
        raise SystemExit("oracle/_ref/libbhm_ref.so is not built (make -C oracle ref)")
    tri = meshes().hmatrix_sphere_mesh(level=args.level)
    m = O.prep(tri.v1, tri.v2, tri.v3, 1, tri.truearea)
    n = m["nar"]
    icomp, ncomp, nopen = O.compon(m)
    ndef = ncomp if ( nopen == 0 and ncomp > 1 ) else 1
    if nopen > 0:
        icomp[:] = 1
    x, _ = O.self_blocks(m)
    t = {}
    t0 = time.perf_counter()
    assert R.bhm_create(n, m["v1"], m["v2"], m["v3"], m["area"], m["afac"], 1.0, args.threads) == 0
    diagm = R.bhm_diag_mean()
    xmu, gdef = O.deflation_weights(m, diagm, icomp, ndef)
    t["create_s"] = time.perf_counter() - t0
    print(f"created ({t['create_s']:.1f} s); building", flush=True)
    t0 = time.perf_counter()
    rc = R.bhm_build(args.eps, 0, 0.1, ndef, gdef, icomp, xmu, 0, 6)
    t["build_s"] = time.perf_counter() - t0
    log = O.ref_log(R)
    print("\n".join(log), flush=True)
    if rc != 0:
        raise SystemExit(f"bhm_build returned {rc}")
    b = O.difco_rhs(n, x)
    bf = np.asfortranarray(b).reshape(-1, order="F").copy()
    t0 = time.perf_counter()
    rc = R.bhm_solve(6, bf)
    t["solve_s"] = time.perf_counter() - t0
    log2 = O.ref_log(R)
    print("\n".join(log2), flush=True)
    if rc != 0:
        raise SystemExit(f"bhm_solve returned {rc}")
    b6 = bf.reshape(b.shape, order="F")
    res = O.difco_post(n, x, m["area"], 1.0, b6)
    res.update(O.units(res["drr"], res["dtt"]))
    out = {"mesh": f"hmatrix_sphere_mesh(level={args.level}): 40 A sphere x (1 + l<=8 bumps sd 0.05)",
           "triangles": n, "eps": args.eps, "gmres_tol": 1e-10, "threads": args.threads,
           "host": platform.processor() or platform.machine(), "timings": t, "log": log + log2}
    for k, v in res.items():
        if isinstance(v, np.ndarray):
            out[k] = v.tolist()
        elif isinstance(v, (float, int, np.floating, np.integer)):
            out[k] = float(v)
    with open(args.out, "w") as f:
        json.dump(out, f, indent=1)
    print(f"D_t {out['dtrans']!r}  D_r {out['dr']!r}; build {t['build_s']:.1f} s, solve {t['solve_s']:.1f} s", flush=True)


if __name__ == "__main__":
    main()
