"""Spectrum broadening on the GPU (stda_hip_gspec through the package's g_spec front end).

Bar: the files are the reference program's, character for character --
rots.dat and spec.dat of every input under tests/golden/gspec (written by the
reference g_spec itself) and of fresh inputs run through the oracle
restatement (oracle/gspec.py, pinned by those fixtures).  The kernel keeps
the reference's operation order and constants, so its FP64 values differ
from the reference's only where exp/log10 round differently (an ulp); the
curves are also compared as numbers, at 1e-13 relative.
"""
import glob
import lzma
import os
import re
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))

import gspec as OG          # noqa: E402  (oracle)
import make_gspec           # noqa: E402

pytestmark = pytest.mark.gpu

GOLD = sorted(glob.glob(os.path.join(HERE, "golden", "gspec", "*", "tda.dat")))


def say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def h():
    from stda_amd.core import StdaHip
    with StdaHip(0) as ctx:
        yield ctx


def golden(path):
    d = os.path.dirname(path)
    with open(path) as f:
        tda = f.read()
    with open(os.path.join(d, "rots.dat")) as f:
        rots = f.read()
    with lzma.open(os.path.join(d, "spec.dat.xz"), "rt") as f:
        spec = f.read()
    return tda, rots, spec


def fortran_float(t):
    """A number of an E edit descriptor, also the three-digit exponent form 0.1234-100."""
    return float(re.sub(r"(\d)([-+]\d{3})$", r"\1E\2", t))


def parse_spec(text):
    return np.array([[fortran_float(t) for t in l.split()] for l in text.splitlines()])


def test_reference_files_one_call(h):
    """All golden inputs in one call (one launch per grid/flag group)."""
    from stda_amd import spectrum as S
    say(f"gspec: {len(GOLD)} golden inputs")
    cases = [golden(p) for p in GOLD]
    out = S.gspec_many(h, [c[0] for c in cases])
    for p, (tda, rots, spec), (r, s) in zip(GOLD, cases, out):
        name = os.path.basename(os.path.dirname(p))
        assert r == rots, name
        bad = [(a, b) for a, b in zip(s.splitlines(), spec.splitlines()) if a != b]
        assert not bad and len(s) == len(spec), (name, bad[:3])


def test_each_reference_file_alone(h):
    from stda_amd import spectrum as S
    for p in GOLD:
        tda, rots, spec = golden(p)
        r, s = S.gspec(h, tda)
        assert (r, s) == (rots, spec), os.path.basename(os.path.dirname(p))


def test_batch_against_oracle(h):
    """48 fresh inputs of mixed options and sizes in one call, against the oracle."""
    from stda_amd import spectrum as S
    rng = np.random.default_rng(9)
    heads = [["NM", "UV"], ["NM", "VELO"], ["UV"], ["NM", "UV", "VELO"], ["CM", "UV"],
             ["NM", "INV", "VELO"], ["NM", "UV", "LOG"]]
    texts = [make_gspec.synthetic(1000 + q, int(rng.integers(1, 40)), heads[q % len(heads)],
                                  0.5, float(rng.uniform(5, 30)), degenerate=q % 4)
             for q in range(48)]
    say("gspec: 48 synthetic inputs")
    out = S.gspec_many(h, texts)
    for q, (text, (r, s)) in enumerate(zip(texts, out)):
        ro, so = OG.run(text)
        assert r == ro, q
        a, b = parse_spec(s), parse_spec(so)
        np.testing.assert_array_equal(a[:, 0], b[:, 0])
        np.testing.assert_allclose(a[:, 1], b[:, 1], rtol=2e-4, atol=0)
        assert s == so, q


def test_curves_as_numbers(h):
    """The raw FP64 curve against the oracle's loop, 1e-13 relative (no printing)."""
    from stda_amd import spectrum as S
    texts = [make_gspec.synthetic(2000 + q, 25, ["NM", "UV", "VELO"] if q % 2 else ["VELO"])
             for q in range(6)]
    inputs = [S.parse_gspec(t) for t in texts]
    curves = S.broaden(h, inputs)
    for g, (e, r) in zip(inputs, curves):
        ref = np.zeros(g.npts)
        uv = bool(g.flags & S.FLAG_UV)
        f1 = OG.F1_UV if uv else OG.F1_CD
        for j in range(1, g.npts + 1):
            ee = g.emin + j * g.step
            acc = 0.0
            for xi, yi, wi in zip(g.x.tolist(), g.y.tolist(), g.wei.tolist()):
                f3 = 1.0 / (f1 * OG.SQRT_PI * wi)
                f2 = xi * (OG.DIP * yi / (3.67493e-2 * xi)) if uv else xi * yi
                acc = acc + f3 * f2 * np.exp(-(1.0 / (wi * wi)) * ((ee - xi) * (ee - xi)))
            ref[j - 1] = acc * g.rfaktor
        scale = np.abs(ref).max()
        np.testing.assert_allclose(r, ref, rtol=1e-13, atol=1e-13 * scale)


def test_many_spectra_one_launch(h):
    """512 spectra of 20 lines (the size of a mixed-molecule batch): sampled against the
    oracle, the kernel's statistics count one launch."""
    from stda_amd import spectrum as S
    texts = [make_gspec.synthetic(3000 + q, 20, ["NM", "UV", "VELO"]) for q in range(512)]
    inputs = [S.parse_gspec(t) for t in texts]
    before = h.kernel_stats("gspec")["launches"]
    curves = S.broaden(h, inputs)
    assert h.kernel_stats("gspec")["launches"] - before == 1
    for q in (0, 1, 255, 511):
        so = OG.run(texts[q])[1]
        assert S._spec_text(*curves[q]) == so, q


def test_spectrum_of_a_gpu_solve(h):
    """Water sTDA on the GPU -> tda.dat -> g_spec: the package's files equal the oracle's on
    the same tda.dat (and the tda.dat rows are those of the reference's own file to its
    printed precision)."""
    import oracle as O  # noqa: F401  (oracle path set up by conftest)
    from stda_amd import spectrum as S
    from stda_amd.core import StdaHip
    from stda_amd.molden import read_molden
    from stda_amd.prep import prepare
    water = os.path.join(HERE, "golden", "water")
    mol = read_molden(os.path.join(water, "water_sto3g.molden"))
    ci = prepare(mol, 20.0, 1.0)
    with StdaHip(0) as hs:
        hs.setup(ci, 20.0, 1.0)
        sol = hs.solve_tda(None)
    tda = S.tda_dat(sol.omega, sol.props, ci.xmolw, velcorr=True)
    with open(os.path.join(water, "EXCI_TEST1.dat")) as f:
        ref_lines = f.read().splitlines()
    got_lines = tda.splitlines()
    assert got_lines[:13] == ref_lines[:13]
    a = np.array([[float(t) for t in l.split()] for l in got_lines[13:]])
    b = np.array([[float(t) for t in l.split()] for l in ref_lines[13:]])
    np.testing.assert_allclose(a[:, 1], b[:, 1], atol=1.01e-4)
    np.testing.assert_allclose(a[:, 2:4], b[:, 2:4], atol=2.5e-6)
    assert S.gspec(h, tda) == OG.run(tda)
