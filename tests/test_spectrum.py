"""Spectral output on the CPU: tda.dat layouts, g_spec's host side and the g_spec oracle.

The oracle restatement of g_spec (oracle/gspec.py) is pinned against the
outputs of the reference program under tests/golden/gspec and, where the
reference tree is present, against the program itself on fresh inputs.  The
package's host side (keyword parsing, number reader, sort, grid, rots.dat)
is compared with the oracle; the broadening kernel runs in the GPU tests.
"""

import glob
import lzma
import os
import subprocess
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(HERE, "golden"))

import gspec as OG                                   # noqa: E402  (oracle)
from stda_amd import spectrum as S                   # noqa: E402
from stda_amd.constants import molecular_weight      # noqa: E402
from stda_amd.molden import read_molden              # noqa: E402

GOLD = sorted(glob.glob(os.path.join(HERE, "golden", "gspec", "*", "tda.dat")))
WATER = os.path.join(HERE, "golden", "water")
REF = os.path.join(ROOT, "oracle", "_ref", "g_spec")


def golden(path):
    d = os.path.dirname(path)
    with open(path) as f:
        tda = f.read()
    with open(os.path.join(d, "rots.dat")) as f:
        rots = f.read()
    with lzma.open(os.path.join(d, "spec.dat.xz"), "rt") as f:
        spec = f.read()
    return tda, rots, spec


def test_golden_cases_present():
    assert len(GOLD) >= 13


@pytest.mark.parametrize("path", GOLD, ids=lambda p: os.path.basename(os.path.dirname(p)))
def test_oracle_matches_reference_outputs(path):
    tda, rots, spec = golden(path)
    r, s = OG.run(tda)
    assert r == rots
    assert s == spec


@pytest.mark.skipif(not os.path.exists(REF), reason="reference g_spec not built (make -C oracle ref)")
def test_oracle_matches_reference_program(tmp_path):
    import make_gspec
    texts = [make_gspec.synthetic(100 + q, 5 + 7 * q, heads, 0.5, 14.0, degenerate=q % 3)
             for q, heads in enumerate([["NM", "UV"], ["VELO"], ["NM", "CM", "UV", "LOG"],
                                        ["NM", "INV", "MIX"], ["CM", "VELO"]])]
    for q, text in enumerate(texts):
        d = tmp_path / str(q)
        d.mkdir()
        (d / "tda.dat").write_text(text)
        with open(d / "tda.dat") as fin:
            subprocess.run([REF], stdin=fin, stdout=subprocess.DEVNULL, cwd=d, check=True)
        r, s = OG.run(text)
        assert r == (d / "rots.dat").read_text()
        assert s == (d / "spec.dat").read_text()


def test_fortran_edit_descriptors():
    # values and strings as the reference program writes them (tests/golden/gspec/wide)
    assert S.fmt_e(0.42e-315) == "    0.4200-315"
    assert S.fmt_e(0.0) == "    0.0000E+00"
    assert S.fmt_e(-0.0) == "   -0.0000E+00"
    assert S.fmt_e(5e-7) == "    0.5000E-06"
    assert S.fmt_e(-123456.0) == "   -0.1235E+06"
    assert S.fmt_e(0.99996) == "    0.1000E+01"
    assert S.fmt_f(62.3459354, 16, 6) == "       62.345935"
    assert S.fmt_f(-1e-9, 13, 6) == "    -0.000000"
    assert S.fmt_f(1.5e12, 16, 6) == "*" * 16
    assert S.fmt_list_real8(18.0152) == "   18.015200000000000     "
    assert S.fmt_list_real8(178.2) == "   178.19999999999999     "
    for v in [0.0, -0.0, 1e-320, 3.14159e-7, 0.099995, 9.99949e10, -2.5e-101, 7e200]:
        assert S.fmt_e(v) == OG.fmt_e(v)


def test_number_reader_matches_oracle():
    rng = np.random.default_rng(5)
    lines = ["   1   14.0839     0.003718     0.024292    -0.000001    -0.000002",
             "   18.015200000000000     ", "  0.5", " -.25 7 1.5E-3 2D+02 -3.0E-5x 12345678901234.5",
             "abc -def - 4", " 0.2046E-63 0.4200-315 1.23456789012345", "x" * 130 + " 5",
             " . -. 1. .5 -0.000000 0012 -3 1234567890.1234567890", " 7" + " " * 117 + "8",
             "1\t2 3", "   1   14.0839" + " " * 110 + "9"]
    for _ in range(200):
        toks = []
        for _ in range(rng.integers(1, 7)):
            v = rng.normal() * 10.0 ** rng.integers(-8, 8)
            toks.append(rng.choice(["%.6f", "%.4E", "%g", "%.12f", "%d"]) % v)
        lines.append(" ".join(toks))
    for line in lines:
        xx = [0.0] * 52
        n = OG.readl(line, xx)
        got = S.read_numbers(line)
        assert len(got) == n, line
        assert got == xx[1:n + 1], line


@pytest.mark.parametrize("path", GOLD, ids=lambda p: os.path.basename(os.path.dirname(p)))
def test_host_side_matches_oracle(path):
    tda, rots, _ = golden(path)
    g = S.parse_gspec(tda)
    assert g.rots == rots
    assert g.npts == 2000
    assert len(g.x) == len(g.y) == len(g.wei)


def test_sort_ties_follow_the_reference():
    # dsort takes the last of equal minima first; keys 3,1,1,2,1 (labels 0..4)
    perm = S._selection_order(np.array([3.0, 1.0, 1.0, 2.0, 1.0]))
    ew = [3.0, 1.0, 1.0, 2.0, 1.0]
    lab = [0, 1, 2, 3, 4]
    OG._dsort(5, lab, [0.0] * 5, [0.0] * 5, ew)
    assert perm.tolist() == lab


def test_grid_size_grows_with_long_inputs():
    body = "\n".join("%4d%10.4f%13.6f%13.6f%13.6f%13.6f" % (i % 9999 + 1, 3.0 + i * 1e-4,
                                                              0.1, 0.1, 0.0, 0.0)
                     for i in range(100100))
    g = S.parse_gspec(" UV\n WIDTH\n  0.20\n DATXY\n" + body + "\n")
    assert g.npts == -(-100104 // 50)
    assert len(g.x) == 100100


def water_rows(name):
    rows = []
    with open(os.path.join(WATER, name)) as f:
        lines = f.read().splitlines()
    at = next(i for i, l in enumerate(lines) if "DATXY" in l)
    for l in lines[at + 1:]:
        rows.append([float(t) for t in l.split()])
    return np.array(rows)


@pytest.mark.parametrize("case", ["EXCI_TEST1.dat", "EXCI_TEST2.dat", "EXCI_TEST3.dat"])
def test_tda_dat_reproduces_reference_file(case):
    """The velocity-corrected layout (apbtrafo.f:115-128, 259-260), rebuilt from the rows of
    the reference's file, is that file byte for byte (header, molar mass, row format)."""
    mol = read_molden(os.path.join(WATER, "water_sto3g.molden"))
    rows = water_rows(case)
    omega = rows[:, 1] / S.EV_PRINT
    text = S.tda_dat(omega, rows[:, 2:6], molecular_weight(mol.Z), velcorr=True)
    with open(os.path.join(WATER, case)) as f:
        assert text == f.read()


def test_tda_dat_print_layout():
    """print_tdadat (stda.f:2235-2265): UV header, roots above thr left out, energies of the
    real*4 eigenvalues times the real*4 constant."""
    omega = np.array([0.25, 0.5, 0.75])
    props = np.array([[0.1, 0.2, -3.0, 4.0], [0.0, 0.0, 0.0, -0.0], [1.0, 1.0, 1.0, 1.0]])
    text = S.tda_dat(omega, props, 18.0152, velcorr=False, thr=0.6)
    lines = text.splitlines()
    assert lines[:3] == [" NM", " UV", " MMASS"]
    assert lines[-3] == " DATXY"
    e1 = np.float32(np.float32(0.25) * np.float32(27.21139))
    assert lines[-2] == "   1%10.4f%13.6f%13.6f%13.6f%13.6f" % (e1, 0.1, 0.2, -3.0, 4.0)
    assert lines[-1].startswith("   2   13.6057")
    # g_spec reads the strengths back from the columns the header selects (UV: fL)
    g = S.parse_gspec(text)
    assert np.allclose(sorted(g.y.tolist()), [0.0, 0.1], rtol=1e-15, atol=0)


def test_native_formatter_matches_the_oracle():
    """spec.dat lines from the library (stda_hip_format_curve, host only) equal the oracle's
    Fortran formats on values that reach every branch: subnormals and three-digit exponents,
    signed zeros, rounding carries, F overflow, non-finite numbers, and a batch long enough
    to be split over threads."""
    rng = np.random.default_rng(7)
    r = np.concatenate([rng.normal(size=9000) * 10.0 ** rng.integers(-320, 300, 9000),
                        [0.0, -0.0, 1e-320, 5e-324, 0.99996, -0.099995, 9.99949e10, np.inf,
                         -np.inf, np.nan]])
    e = rng.normal(size=len(r)) * 10.0 ** rng.integers(-3, 12, len(r))
    got = S._spec_text(e, r)
    assert got == "".join(OG.fmt_line(a, b) + "\n" for a, b in zip(e.tolist(), r.tolist()))


@pytest.mark.parametrize("case", [1, 2, 3])
def test_beta_hrs_file_reproduces_reference_file(case):
    """beta_HRS (linear_response.f:695: list-directed real*4 energy in eV and real*8 beta)
    rebuilt from the frequencies of the wavelength file and the file's own beta values is
    the reference's BETA_TEST file byte for byte."""
    sys.path.insert(0, HERE)
    from conftest import beta_test_omegas
    with open(os.path.join(WATER, f"BETA_TEST{case}.txt")) as f:
        ref = f.read()
    vals = [float(l.split()[1]) for l in ref.splitlines()]
    assert S.beta_hrs_file(beta_test_omegas(), vals) == ref


@pytest.mark.parametrize("case", [1, 2, 3])
def test_tpa_abs_file_reproduces_reference_file(case):
    """2PA-abs (format (I3,F7.3,4F20.3)) rebuilt from the rows of the reference's 2PA_TEST
    file (energies converted back to Hartree) is that file byte for byte."""
    with open(os.path.join(WATER, f"2PA_TEST{case}.txt")) as f:
        ref = f.read()
    rows = [[float(t) for t in l.split()] for l in ref.splitlines()]
    eci = [r[1] / 27.21139 for r in rows]
    assert S.tpa_abs_file(eci, [r[2:6] for r in rows]) == ref
