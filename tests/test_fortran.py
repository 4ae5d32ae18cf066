"""The Fortran side of the boundary: the ISO_C_BINDING module fortran/stda_hip.F90 and the
harness fortran/tda_harness.F90, compiled with ROCm's flang in both integer models (4-byte
default integers with libstda_hip.so; -fdefault-integer-8 with libstda_hip_ilp64.so), which
pass the arrays as std2's `stda` holds them at stda.f:452 and write tda.dat as
apbtrafo.f:115-128,265 does.

CPU: the module and harness compile, the integer models of caller and library agree, and
an error status stops with the library's message.  GPU: the water fixtures EXCI_TEST1-3
(the reference's tda.dat of tests/run_tests.sh:22-68) through the harness."""
import os
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, ROOT
from fortran_io import HARNESS, read_tda_dat, run_harness, write_input
from stda_amd.prep import prepare

FLANG = "/opt/rocm/lib/llvm/bin/flang"
WATER = {1: (1.0, 20.0, -100.0, -100.0), 2: (0.5, 30.0, -100.0, -100.0), 3: (1.0, 30.0, 3.0, 1.0)}


@pytest.mark.parametrize("model", ["lp64", "ilp64"])
def test_module_compiles_and_index_models_agree(model, tmp_path):
    if not os.path.exists(FLANG):
        pytest.skip("flang not present")
    flags = ["-fdefault-integer-8", "-DSTDA_HIP_ILP64"] if model == "ilp64" else []
    r = subprocess.run([FLANG, *flags, "-fsyntax-only", "-module-dir", str(tmp_path),
                        os.path.join(ROOT, "fortran", "stda_hip.F90")],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    assert os.path.exists(HARNESS[model]), "build the harness first (python -m stda_amd.build)"
    out = subprocess.run([HARNESS[model], "--index-size"], capture_output=True, text=True,
                         timeout=60).stdout
    size = 4 if model == "lp64" else 8
    assert out.split() == ["library", str(size), "module", str(size), "default", "integer",
                           str(size)]


def test_error_status_stops_with_the_message(water, tmp_path):
    """Without a device (this container) the context cannot be created: the binding
    prints the library's message and stops with a nonzero status, as std2 stops."""
    from stda_amd.core import load_library
    ci = prepare(water, 20.0, 1.0)
    inp = str(tmp_path / "in.bin")
    write_input(inp, ci, 20.0, 1.0)
    L = load_library()
    r = run_harness("lp64", inp, str(tmp_path / "tda.dat"))
    if r.returncode == 0:
        pytest.skip("a device is present")
    assert L.stda_hip_error_string(3).decode() in r.stdout
    assert not os.path.exists(tmp_path / "tda.dat")


@pytest.mark.gpu
@pytest.mark.parametrize("model", ["lp64", "ilp64"])
@pytest.mark.parametrize("case", [1, 2, 3])
def test_water_fixtures_through_fortran(water, case, model, tmp_path):
    from stda_amd.core import StdaHip
    ax, thr, al, be = WATER[case]
    print(f"water EXCI_TEST{case} through the {model} Fortran harness", flush=True)
    ci = prepare(water, thr, ax)
    inp, out = str(tmp_path / "in.bin"), str(tmp_path / "tda.dat")
    write_input(inp, ci, thr, ax, alphak=al, betaj=be)
    r = run_harness(model, inp, out)
    assert r.returncode == 0, r.stdout + r.stderr
    head, rows = read_tda_dat(out)
    ehead, erows = read_tda_dat(os.path.join(GOLDEN, "water", f"EXCI_TEST{case}.dat"))
    assert [h for h in head if h.isalpha()] == [h for h in ehead if h.isalpha()]
    assert rows.shape == erows.shape
    np.testing.assert_array_equal(rows[:, 0], erows[:, 0])
    assert np.all(np.abs(rows[:, 1] - erows[:, 1]) <= 0.5e-4 + 1e-6)       # F10.4 energies
    assert np.all(np.abs(rows[:, 2:4] - erows[:, 2:4]) <= 2.5e-6)          # fL, fV (f13.6)
    # RL, Rv(corr) vanish for achiral water; the reference binary prints its FP32
    # round-off there (|R| <= 2.5e-5 in the fixtures)
    assert np.all(np.abs(rows[:, 4:] - erows[:, 4:]) <= 2.5e-5)
    # the same numbers as the Python host layer over the same library
    with StdaHip(0, ilp64=(model == "ilp64")) as h:
        h.setup(ci, thr, ax, alphak=al, betaj=be)
        sol = h.solve_tda(None)
    np.testing.assert_allclose(rows[:, 1], np.round(sol.omega * 27.21139, 4), atol=1e-9)
    np.testing.assert_allclose(rows[:, 2:], np.round(sol.props[:, :4], 6), atol=1e-9)
