"""Input files of the Fortran harness (fortran/tda_harness.F90): the boundary arrays that
std2's `stda` holds at stda.f:452, in the stream layout the harness documents."""
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HARNESS = {"lp64": os.path.join(ROOT, "fortran", "tda_harness"),
           "ilp64": os.path.join(ROOT, "fortran", "tda_harness_ilp64")}


def write_input(path, ci, thr_ev, ax, alphak=-100.0, betaj=-100.0, thrp=1e-4, fthr_ev=-1.0,
                nroots=0, velcorr=True, rpa=False):
    with open(path, "wb") as fh:
        np.array([1396982849, 1], dtype="<i4").tofile(fh)
        np.array([ci.ncent, ci.nao, ci.no, ci.nv, nroots, int(velcorr), int(rpa)],
                 dtype="<i4").tofile(fh)
        np.array([thr_ev, thrp, fthr_ev, ax, alphak, betaj, ci.xmolw], dtype="<f8").tofile(fh)
        np.asarray(ci.clow, dtype="<f8").ravel(order="F").tofile(fh)
        np.asarray(ci.epsi, dtype="<f8").tofile(fh)
        np.asarray(ci.iaoat, dtype="<i4").tofile(fh)
        np.asarray(ci.xyz, dtype="<f8").ravel(order="F").tofile(fh)
        np.asarray(ci.moments, dtype="<f8").ravel(order="C").tofile(fh)


def run_harness(model, inp, out, timeout=300):
    return subprocess.run([HARNESS[model], inp, out], capture_output=True, text=True,
                          timeout=timeout)


def read_tda_dat(path):
    """Header keywords and the DATXY rows of a tda.dat file."""
    lines = open(path).read().splitlines()
    k = next(i for i, ln in enumerate(lines) if ln.strip() == "DATXY")
    head = [ln.strip() for ln in lines[:k + 1]]
    rows = [[float(x) for x in ln.split()] for ln in lines[k + 1:] if ln.strip()]
    return head, np.array(rows)
