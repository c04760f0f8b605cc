"""BEST's mesh file through the library (bhg_read_best, bhg_write_best; host code, no GPU).

The reader against the READ statements of best/best.f:189-220 compiled by amdflang in a program of
their own (tests/fortran/read_mesh.f90), bit for bit, on the reference's two example meshes and on
files that exercise list-directed input: values over several lines, values past the ninth on a
line (skipped: each READ starts on a new line), commas, D and Q exponents, signed exponents with no
letter, repeat counts, CRLF line ends, the -na header; and against meshes.parse_best.  The writer
against the reference's own rcoal-format example (examples/best/2eu7_02562, byte for byte) and
round trips through the reader.
"""
import os
import subprocess

import numpy as np
import pytest
from conftest import GOLDEN, ROOT

FLANG = "/opt/rocm/bin/amdflang"


@pytest.fixture(scope="module")
def fortran_reader(tmp_path_factory):
    if not os.path.exists(FLANG):
        pytest.skip("amdflang not available")
    d = tmp_path_factory.mktemp("fr")
    exe = str(d / "read_mesh")
    subprocess.run([FLANG, "-O0", "-o", exe, os.path.join(ROOT, "tests", "fortran", "read_mesh.f90")], check=True,
                   cwd=str(d))
    return exe


def _fortran_read(exe, path, iarea):
    out = subprocess.run([exe, path, str(iarea)], capture_output=True, text=True, timeout=120, check=True).stdout.split()
    nar = int(out[0])
    bits = np.array([int(h, 16) for h in out[1:]], dtype=np.uint64).view(np.float64)
    ta, tv = bits[0], bits[1]
    v = bits[2:].reshape(nar, 3, 3)  # triangle, vertex, coordinate
    return nar, ta, tv, v[:, 0, :], v[:, 1, :], v[:, 2, :]


def _same(bh, exe, path, iarea):
    t = bh.read_best_file(path, iarea)
    nar, ta, tv, v1, v2, v3 = _fortran_read(exe, path, iarea)
    assert t.nar == nar
    if iarea == 1:
        assert t.truearea.hex() == ta.hex() and t.truevol.hex() == tv.hex()
    for a, b in ((t.v1, v1), (t.v2, v2), (t.v3, v3)):
        assert np.array_equal(a.view(np.uint64), b.view(np.uint64))
    return t


@pytest.mark.parametrize("fname", ["2eu7_02562", "1aki49_atom_01806"])
def test_reader_matches_fortran_on_the_examples(bh, fortran_reader, fname):
    from best_hydro_amd import meshes as M
    path = os.path.join(GOLDEN, fname)
    t = _same(bh, fortran_reader, path, 1)
    p = M.read_best(path)
    assert np.array_equal(t.v1, p.v1) and np.array_equal(t.v2, p.v2) and np.array_equal(t.v3, p.v3)
    _same(bh, fortran_reader, path, 0)


LIST_DIRECTED = {
    "spanning": "3 100.5 200.25\n1 2 3\n4 5 6\n7 8 9\n1 2 3 4 5 6\n7 8 9\n1 2 3 4 5 6 7 8 9\n",
    "extra_values_skipped": "2 1.0 2.0 junk\n1 2 3 4 5 6 7 8 9 99 98\n-1 -2 -3 -4 -5 -6 -7 -8 -9\n",
    "commas_and_exponents": "2,1.5D+2,2.5d-1\n0.1E1,2.0D0,3Q0 ,4.,.5,6.25e-3,7.0+2,8.0-1,9\n1,2,3,4,5,6,7,8,9\n",
    "repeat_counts": "2 1 1\n3*0.5 3*1.25 2*2 3\n9*7\n",
    "crlf": "2 10 20\r\n1 2 3 4 5 6 7 8 9\r\n9 8 7 6 5 4 3 2 1\r\n",
    "blank_lines_inside": "2 10 20\n\n1 2 3 4\n\n5 6 7 8 9\n9 8 7 6 5 4 3 2 1\n",
    "correct_rounding": "1 0.1 0.30000000000000004\n0.1 0.2 0.3 1e-320 2.2250738585072011e-308 9.999999999999999e22 "
                        "1.7976931348623157e308 4.9406564584124654e-324 0.1000000000000000055511151231257827\n",
}


@pytest.mark.parametrize("case", sorted(LIST_DIRECTED))
@pytest.mark.parametrize("iarea", [1, 0])
def test_reader_matches_fortran_list_directed(bh, fortran_reader, tmp_path, case, iarea):
    path = str(tmp_path / (case + ".mesh"))
    with open(path, "w", newline="") as f:
        f.write(LIST_DIRECTED[case])
    _same(bh, fortran_reader, path, iarea)


def test_reader_errors(bh, tmp_path):
    short = tmp_path / "short.mesh"
    short.write_text("3 1 2\n1 2 3 4 5 6 7 8 9\n1 2 3\n")
    with pytest.raises(bh.HipError, match="end of file"):
        bh.read_best_file(str(short))
    bad = tmp_path / "bad.mesh"
    bad.write_text("1 1 2\n1 2 x 4 5 6 7 8 9\n")
    with pytest.raises(bh.HipError, match='line 2: cannot read "x"'):
        bh.read_best_file(str(bad))
    with pytest.raises(bh.HipError, match="cannot open"):
        bh.read_best_file(str(tmp_path / "missing.mesh"))
    zero = tmp_path / "zero.mesh"
    zero.write_text("0 1 2\n")
    with pytest.raises(bh.HipError, match="not a number of triangles"):
        bh.read_best_file(str(zero))


def test_writer_reproduces_the_reference_example(bh, tmp_path):
    """examples/best/2eu7_02562 (written by rcoal: list-directed header, 9e17.8 lines) read and
    written again: the same bytes."""
    src = os.path.join(GOLDEN, "2eu7_02562")
    t = bh.read_best_file(src)
    out = str(tmp_path / "2eu7.mesh")
    bh.write_best_file(out, t)
    with open(out, "rb") as f, open(src, "rb") as g:
        assert f.read() == g.read()


def test_writer_edit_descriptors_and_round_trip(bh, oracle, tmp_path):
    """9e17.8 on every kind of value (the oracle's Ew.d restatement, itself pinned against the
    reference's compiled OUTPUT), and the header's fields; what is written reads back to the
    values rounded to 8 digits."""
    from best_hydro_amd import meshes as M
    rng = np.random.default_rng(5)
    n = 64
    vals = rng.standard_normal((n, 9)) * 10 ** rng.uniform(-5, 5, (n, 9))
    vals[0, :] = [0.0, -0.0, 1e-120, -3.3e150, 9.999999995e99, 1e100, 5e-100, 0.5, -99999.5]
    t = M.Triangulation(vals[:, 0:3], vals[:, 3:6], vals[:, 6:9], 0.0123, 12345678901234567.0, name="edge")
    out = str(tmp_path / "edge.mesh")
    bh.write_best_file(out, t)
    lines = open(out).read().split("\n")
    assert lines[0][:13] == " " + "%12d" % n
    for i in range(n):
        want = "".join(oracle._fortran_e(v, 17, 8) for v in vals[i])
        assert lines[1 + i] == want, i
    def fortran_value(text):  # a three-digit exponent is written without its letter: 0.1+101
        text = text.strip()
        if "E" not in text:
            k = max(text.rfind("+"), text.rfind("-"))
            text = text[:k] + "E" + text[k:]
        return float(text)
    back = bh.read_best_file(out)
    for a, b in ((back.v1, t.v1), (back.v2, t.v2), (back.v3, t.v3)):
        want = np.array([[fortran_value(oracle._fortran_e(v, 17, 8)) for v in row] for row in b])
        assert np.array_equal(a, want)
    assert back.truearea == float("%.16g" % 0.0123) and back.truevol == float("%.16g" % 12345678901234567.0)
