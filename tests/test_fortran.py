"""The Fortran host over iso_c_binding (best-hydro_amd/fortran): the best_hip module binds the
handle API of include/best_hip.h, and bh_solve solves one BEST mesh file with it and writes BEST's
.be file (bhg_write_be)."""
import os
import subprocess

import pytest
from conftest import GOLDEN, ROOT, rel

EXE = os.path.join(ROOT, "best-hydro_amd", "fortran", "bh_solve")

needs_exe = pytest.mark.skipif(not os.path.exists(EXE), reason="bh_solve not built (make -C best-hydro_amd/fortran)")


@needs_exe
def test_usage_without_arguments():
    p = subprocess.run([EXE], capture_output=True, text=True, timeout=60)
    assert p.returncode == 2
    assert "usage: bh_solve" in p.stdout


def _parse(out):
    vals = {}
    for line in out.splitlines():
        for key in ("DTRANS", "DR (1/s)", "CHI*V", "CHI "):
            if line.startswith(key):
                vals[key.split()[0]] = float(line.split()[-1])
    return vals


@needs_exe
@pytest.mark.gpu
@pytest.mark.parametrize("prec", [32, 64])
def test_fortran_solve_matches_oracle(oracle, prec):
    path = os.path.join(GOLDEN, "2eu7_02562")
    print(f"bh_solve {path} {prec}", flush=True)
    p = subprocess.run([EXE, path, str(prec), "1", "1", "0"], capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stdout + p.stderr
    got = _parse(p.stdout)
    v1, v2, v3, ta, tv = oracle.read_best(path)
    ref = oracle.best_run(v1, v2, v3, ta, tv, iarea=1, cg=1.0, ivisco=1, icov=0)
    assert rel(got["DTRANS"], ref["dtrans"]) < 1e-10
    assert rel(got["DR"], ref["dr"]) < 1e-10
    assert rel(got["CHI*V"], ref["chiprime"]) < 1e-10
    assert rel(got["CHI"], ref["chi"]) < 1e-10


@needs_exe
@pytest.mark.gpu
@pytest.mark.parametrize("case", ["2eu7_v_mw30000", "2eu7_vc_mw30000", "2eu7_na_v_mw30000"])
def test_fortran_writes_the_be_file(bh, oracle, tmp_path, case):
    """bh_solve -v -mw 30000 (and -vc, -na) on the reference's example mesh: the .be file is named
    as BEST names it, its printed values carry the reference's expected_best_*.txt digits, and its
    bytes equal the library's writer on the same solve driven from Python."""
    import shutil
    from conftest import golden_expected
    from test_befile import check_against_fixture
    f = golden_expected(case)["flags"]
    mesh = tmp_path / f["file"]
    shutil.copy(os.path.join(GOLDEN, f["file"]), mesh)
    args = [EXE, str(mesh), "32", str(f["iarea"]), "1", str(f["icov"]), str(f["wm"])]
    print(" ".join(args), flush=True)
    p = subprocess.run(args, capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stdout + p.stderr
    suffix = "v" + ("c" if f["icov"] else "") + "m" + ("n" if f["iarea"] == 0 else "")
    be = tmp_path / (f["file"] + suffix + ".be")
    assert be.exists(), p.stdout
    text = be.read_bytes()
    check_against_fixture(oracle, text.decode(), case)
    from best_hydro_amd import meshes as M
    t = M.read_best(str(mesh))
    s = bh.Solver(0)
    s.set_mesh(t)
    r = s.solve_mesh(iarea=f["iarea"], ivisco=1, icov=f["icov"], precision=32)
    s.close()
    want = bh.format_be(r, nar=t.nar, truearea=t.truearea, truevol=t.truevol, iarea=f["iarea"], mw=1, wm=f["wm"],
                        ivisco=1, icov=f["icov"], cg=1.0)
    assert text == want
