"""Solve one BEST mesh on the GPU and print the values BEST writes to its .be file.

    python best-hydro_amd/solve_mesh.py MESHFILE [--method dense|hmatrix] [--precision 32|64]
    python best-hydro_amd/solve_mesh.py --workload sphere|bumpy_ellipsoid|two_surface|hmatrix_sphere ...
    python best-hydro_amd/solve_mesh.py ... --gradient [--gradient-out FILE.npy]

MESHFILE is BEST's input (NAR TRUEAREA TRUEVOL, then 9 coordinates per triangle).  --workload
generates one of the synthetic meshes instead.  -v (the 5 strain flows, [eta]) is on unless
--no-visco; --icov is -vc, --no-area is -na, --preaveraged is -p.  With --method hmatrix, --eps
is BEST's -he and --single its -hs; --tol is the GMRES relative residual.  --gradient also runs
the adjoint (dense path): the gradient of D_t, D_r and CHI*V with respect to the shared vertex
coordinates, saved as an (outputs x vertices x 3) array with --gradient-out.
"""
import argparse
import importlib.util
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def _package():
    if "best_hydro_amd" in sys.modules:
        return sys.modules["best_hydro_amd"]
    spec = importlib.util.spec_from_file_location("best_hydro_amd", os.path.join(HERE, "__init__.py"),
                                                  submodule_search_locations=[HERE])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["best_hydro_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("mesh", nargs="?")
    ap.add_argument("--workload", choices=["sphere", "bumpy_ellipsoid", "two_surface", "hmatrix_sphere"])
    ap.add_argument("--level", type=int, default=None, help="icosphere level of the generated workload")
    ap.add_argument("--method", choices=["dense", "hmatrix"], default="dense")
    ap.add_argument("--precision", type=int, choices=[32, 64], default=32)
    ap.add_argument("--eps", type=float, default=1e-6)
    ap.add_argument("--tol", type=float, default=0.0)
    ap.add_argument("--single", action="store_true")
    ap.add_argument("--no-visco", action="store_true")
    ap.add_argument("--icov", action="store_true")
    ap.add_argument("--no-area", action="store_true")
    ap.add_argument("--preaveraged", action="store_true")
    ap.add_argument("--mw", type=float, default=None, help="molecular weight for [eta]")
    ap.add_argument("--repeat", type=int, default=1)
    ap.add_argument("--gradient", action="store_true")
    ap.add_argument("--gradient-out", default=None)
    a = ap.parse_args(argv)
    bh = _package()
    from best_hydro_amd import meshes as M
    if a.mesh:
        tri = M.read_best(a.mesh)
    elif a.workload == "sphere":
        tri = M.sphere_mesh(a.level or 3, 15.0)
    elif a.workload == "bumpy_ellipsoid":
        tri = M.bumpy_ellipsoid_mesh(a.level or 4)
    elif a.workload == "two_surface":
        tri = M.two_surface_mesh(a.level or 4)
    elif a.workload == "hmatrix_sphere":
        tri = M.hmatrix_sphere_mesh(level=a.level or 6)
    else:
        ap.error("a mesh file or --workload")
    s = bh.Solver(0)
    s.set_mesh(tri)
    if a.method == "hmatrix":
        s.set_method("hmatrix", eps=a.eps, single=a.single, tol=a.tol)
    if a.gradient and (a.method != "dense" or a.no_visco or a.icov):
        ap.error("--gradient needs the dense method, the strain flows and no --icov")
    grad = None
    for _ in range(a.repeat):
        t0 = time.perf_counter()
        if a.gradient:
            r, grad = s.gradient(iarea=0 if a.no_area else 1, cg=0.0 if a.preaveraged else 1.0, precision=a.precision,
                                 wm=a.mw)
        else:
            r = s.solve_mesh(iarea=0 if a.no_area else 1, cg=0.0 if a.preaveraged else 1.0,
                             ivisco=0 if a.no_visco else 1, icov=1 if a.icov else 0, precision=a.precision, wm=a.mw)
        el = time.perf_counter() - t0
    print(f"triangles {len(tri.v1)}  surfaces {r['ncomp']}  method {a.method}  {el * 1e3:.1f} ms")
    print(f"DTRANS (cm^2/s)  {r['dtrans']:.16e}")
    print(f"DR (1/s)         {r['dr']:.16e}")
    if not a.no_visco:
        print(f"CHI              {r['chi']:.16e}")
        print(f"CHI*V (A^3)      {r['chiprime']:.16e}")
        if "eta" in r:
            print(f"[eta] (cm^3/g)   {r['eta']:.16e}")
    if grad is not None:
        V, idx = M.vertex_index(tri)
        gv = M.vertex_gradient(grad, idx, len(V))
        for k, name in enumerate(("D_t", "D_r", "CHI*V", "[eta]")[: len(gv)]):
            print(f"gradient of {name}: |g| = {np.linalg.norm(gv[k]):.6e} per A over {len(V)} vertices")
        if a.gradient_out:
            np.save(a.gradient_out, gv)
    if a.method == "hmatrix":
        st = s.hm_stats()
        print(f"H-matrix: {(st['bytes_dense'] + st['bytes_lowrank']) / 1e9:.2f} GB, rank max {st['kmax']} mean "
              f"{st['kmean']:.1f}; GMRES {st['iters_min']}..{st['iters_max']} iterations, relative residual "
              f"{st['relres']:.1e}")
    s.close()


if __name__ == "__main__":
    main()
