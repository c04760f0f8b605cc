"""The .be file (BEST's OUTPUT, best/output.f): the library's writer (bhg_format_be, bhg_write_be)
and the oracle's restatement (oracle.format_be), checked against the reference's own OUTPUT and
JACOBI compiled from best/output.f and best/jacobi.f (oracle/_ref/libbest_output_ref.so), byte for
byte and bit for bit, on seeded values that reach every record and every edit-descriptor case
(zeros, signed zeros, roundings across a decade, three-digit exponents, overflowing fields, Inf
and NaN); and against the printed values of the reference's expected_best_*.txt fixtures.

No GPU: the writer is host code.
"""
import os
import re

import numpy as np
import pytest
from conftest import GOLDEN, golden_expected

SPECIALS = [0.0, -0.0, 9.999995e-5, 9.999949e-5, -9.999995e-5, 1e-120, -3.3e150, 1e99, 9.99999e99, 9.999996e99,
            1e100, 5e-100, 1.5e-99, 0.5, 99999.5, 12345678.9, 5e-324, 1.7976931348623157e308]


def _ref(oracle):
    if oracle.output_ref() is None:
        pytest.skip("oracle/_ref/libbest_output_ref.so not built (needs the reference sources: make -C oracle ref)")


def _sym(rng, scale):
    a = rng.standard_normal((3, 3)) * scale
    return (a + a.T) / 2


def _cases(n, seed=7):
    """Seeded OUTPUT arguments; every flag combination, values from 1e-12 to 1e12 of either sign
    with the SPECIALS mixed in, and tensors with near-equal, zero and exact-diagonal cases."""
    rng = np.random.default_rng(seed)

    def val(s=1.0):
        if rng.random() < 0.3:
            return float(rng.choice(SPECIALS))
        return float(rng.standard_normal() * 10 ** rng.uniform(-12, 12)) * s

    out = []
    for t in range(n):
        dtt = _sym(rng, 1e-2) + np.eye(3) * 0.02
        drr = _sym(rng, 1e-5) + np.eye(3) * 1e-4
        if t % 13 == 0:
            dtt = np.zeros((3, 3))
        if t % 17 == 0:
            drr = np.diag([3e-4, 1e-4, 2e-4])
        if t % 19 == 0:
            drr = np.full((3, 3), 1e-5) + np.eye(3) * 1e-4  # two equal eigenvalues
        chip = val()
        if t % 11 == 0:
            chip = [float("nan"), float("inf"), -float("inf")][(t // 11) % 3]
        out.append(dict(
            ver=["4.0 SFSU/ARAGON", "x" * 20, ""][t % 3], iarea=t % 2, truearea=abs(val()),
            afac=[1.0123, 123.45, 0.00004, 99.99996, 9.99996, -0.00004][t % 6],
            truevol=[35000.0, 0.0, -0.0, 1e-300][t % 4] if t % 7 == 0 else abs(val()),
            mw=(t // 2) % 2, wm=[30000.0, 1234567.89, 999999.96, 0.04][t % 4], nar=[2562, 99999, 100000, 7][t % 4],
            centroid=[val() for _ in range(3)], urc=[val() for _ in range(3)], udr=[val() for _ in range(3)],
            dtt=dtt, dtr=_sym(rng, 1e-5), drr=drr, ivisco=(t // 4) % 2, icov=(t // 8) % 2,
            dvisc=[val() for _ in range(3)], chiprime=chip, cg=[1.0, 0.0][(t // 3) % 2], totarea=abs(val())))
    return out


def test_oracle_format_matches_reference_output(oracle, tmp_path):
    """oracle.format_be against the reference's OUTPUT on the same arguments: the same bytes."""
    _ref(oracle)
    scratch = str(tmp_path / "o.be")
    if len(scratch) > 80:
        scratch = "/tmp/bh_output_ref.be"
    for k, c in enumerate(_cases(400)):
        a = oracle.format_be(**c)
        b = oracle.ref_output_be(scratch, **c)
        assert a == b, (k, [(x, y) for x, y in zip(a.split(b"\n"), b.split(b"\n")) if x != y][:3])


def test_oracle_jacobi_matches_reference_jacobi(oracle):
    """The oracle's JACOBI restatement against the reference's compiled JACOBI: the same bits, on
    random, diagonal, repeated-eigenvalue and tiny or huge symmetric matrices."""
    _ref(oracle)
    rng = np.random.default_rng(3)
    for t in range(3000):
        a = _sym(rng, 10 ** rng.uniform(-8, 8))
        if t % 7 == 0:
            a = np.diag(rng.standard_normal(3))
        if t % 11 == 0:
            a = np.full((3, 3), rng.standard_normal()) + np.eye(3) * rng.standard_normal()
        if t % 13 == 0:
            a = a + np.eye(3) * 1e3 * np.max(np.abs(a))
        d_o = oracle.jacobi(a)
        d_r = oracle.ref_jacobi(a)
        assert np.array_equal(d_o, d_r), (t, d_o, d_r)


def _as_result(c):
    """A bhg_result dict with the values OUTPUT prints."""
    return dict(dtt=c["dtt"], dtr=c["dtr"], drr=c["drr"], urc=c["urc"], udr=c["udr"], dvisc=c["dvisc"],
                chiprime=c["chiprime"], afac=c["afac"], totarea=c["totarea"], centroid=c["centroid"])


def _lib_format(bh, c):
    return bh.format_be(_as_result(c), nar=c["nar"], truearea=c["truearea"], truevol=c["truevol"], iarea=c["iarea"],
                        mw=c["mw"], wm=c["wm"], ivisco=c["ivisco"], icov=c["icov"], cg=c["cg"], ver=c["ver"])


def test_library_format_matches_oracle(bh, oracle, tmp_path):
    """bhg_format_be and bhg_write_be against oracle.format_be: the same bytes; values that BEST
    refuses before OUTPUT (non-finite tensors or viscosity values) refused with its message."""
    refused = 0
    for k, c in enumerate(_cases(400, seed=11)):
        want = oracle.format_be(**c)
        finite = all(np.all(np.isfinite(np.asarray(c[n], dtype=float))) for n in ("dtt", "dtr", "drr", "urc", "udr"))
        vfin = np.isfinite(c["chiprime"]) and np.all(np.isfinite(c["dvisc"]))
        if not finite or (c["ivisco"] and not vfin):
            refused += 1
            with pytest.raises(bh.HipError, match="are not finite .NaN or Inf.; nothing was written to the .be file"):
                _lib_format(bh, c)
            continue
        got = _lib_format(bh, c)
        assert got == want, (k, [(x, y) for x, y in zip(got.split(b"\n"), want.split(b"\n")) if x != y][:3])
        if k % 25 == 0:
            p = tmp_path / f"m{k}.be"
            bh.write_be(str(p), _as_result(c), nar=c["nar"], truearea=c["truearea"], truevol=c["truevol"],
                        iarea=c["iarea"], mw=c["mw"], wm=c["wm"], ivisco=c["ivisco"], icov=c["icov"], cg=c["cg"],
                        ver=c["ver"])
            assert p.read_bytes() == want
    assert refused > 0


def test_nonfinite_tensor_refused(bh):
    c = _cases(1)[0]
    c["dtt"] = np.array(c["dtt"]).copy()
    c["dtt"][1, 2] = np.nan
    with pytest.raises(bh.HipError, match="the diffusion tensors are not finite"):
        _lib_format(bh, c)


# The labels check_best.pl (test/check_best.pl) looks for, the record's edit descriptor digits.
LABELS = {"dtrans": ("1/3 Trace DTT(EXACT)", 5), "dr": ("1/3 Trace DRR(EXACT)", 5), "chi": ("CHI", 6),
          "eta": ("ETA (cm^3/g)", 6), "dx": ("Dx =", 5), "dy": ("Dy =", 5), "dz": ("Dz =", 5)}


def be_value_text(text, label):
    """The number after label, as test/check_best.pl reads it."""
    m = re.search(re.escape(label) + r"\s*=?\s*([-+]?[0-9.]+(?:[EeDd][-+]?[0-9]+)?)", text)
    return m.group(1) if m else None


def check_against_fixture(oracle, text, case):
    """Every printed value the fixture holds appears in the .be with the fixture's digits (the
    reference's .be prints them with the same edit descriptors)."""
    g = golden_expected(case)
    for key, val in g.items():
        if key == "flags":
            continue
        want = val[0]
        label, d = LABELS[key]
        got = be_value_text(text, label)
        assert got is not None, (case, label)
        assert got == oracle._fortran_e(want, 20, d).strip(), (case, key, got, want)


@pytest.mark.parametrize("case", ["2eu7_v_mw30000", "2eu7_vc_mw30000", "2eu7_na_v_mw30000", "two_v_mw5000"])
def test_oracle_be_prints_the_reference_digits(oracle, bh, case):
    """The oracle's dense run of each fixture case written as a .be file by the library's writer
    (and the oracle's): the printed values equal the reference's expected_best_*.txt digits."""
    g = golden_expected(case)
    f = g["flags"]
    v1, v2, v3, ta, tv = oracle.read_best(os.path.join(GOLDEN, f["file"]))
    r = oracle.best_run(v1, v2, v3, ta, tv, iarea=f["iarea"], cg=1.0, ivisco=1, icov=f["icov"], wm=f["wm"])
    res = dict(dtt=r["dtt"], dtr=r["dtr"], drr=r["drr"], urc=r["urc"], udr=r["udr"], dvisc=r["dvisc"],
               chiprime=r["chiprime"], afac=r["afac"], totarea=r["totarea"], centroid=r["centroid"])
    text = bh.format_be(res, nar=r["nar"], truearea=ta, truevol=tv, iarea=f["iarea"], mw=1, wm=f["wm"], ivisco=1,
                        icov=f["icov"], cg=1.0)
    o = oracle.format_be("4.0 SFSU/ARAGON", f["iarea"], ta, r["afac"], tv, 1, f["wm"], r["nar"], r["centroid"], r["urc"], r["udr"], r["dtt"],
                         r["dtr"], r["drr"], 1, f["icov"], r["dvisc"], r["chiprime"], 1.0, r["totarea"])
    assert text == o
    check_against_fixture(oracle, text.decode(), case)
