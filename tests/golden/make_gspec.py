"""Writes the g_spec fixtures under tests/golden/gspec/ with the reference program itself.

Each case is a tda.dat-style input (the water EXCI_TEST1 output of the
reference's own test suite, and seeded synthetic inputs that switch on the
program's keywords: UV/CD, VELO/MIX, NM/CM, LOG, INV, AU, ETHR, FORB, widely
spread lines whose curve reaches subnormal values, degenerate energies).  The
reference program g_spec is built from its source by ``make -C oracle ref``
(oracle/_ref/g_spec) and run on each input; its rots.dat and spec.dat are
stored next to the input (spec.dat xz-compressed).

Run from the repository root after ``make -C oracle ref``:
    python tests/golden/make_gspec.py
"""

from __future__ import annotations

import lzma
import os
import shutil
import subprocess
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
OUT = os.path.join(ROOT, "tests", "golden", "gspec")
REF = os.path.join(ROOT, "oracle", "_ref", "g_spec")


def synthetic(seed: int, n: int, head: list, emin: float = 1.0, emax: float = 12.0,
              degenerate: int = 0, extra: tuple = ()) -> str:
    """A tda.dat-like input: header keywords, then n DATXY rows (i4, F10.4, 4f13.6)."""
    rng = np.random.default_rng(seed)
    e = np.sort(rng.uniform(emin, emax, n))
    if degenerate:
        e[1:1 + degenerate] = e[0]
    out = [" " + h for h in head]
    out += [" MMASS", "   123.45600000000000     ", " LFAKTOR", "  0.5", " RFAKTOR", "  1.0",
            " WIDTH", "  0.20", " SHIFT", "  0.00", *extra, " DATXY"]
    for i in range(n):
        f = rng.uniform(0, 0.5, 4) * rng.choice([-1, 1], 4) * np.array([0.5, 0.5, 50, 50])
        f[:2] = np.abs(f[:2])
        out.append("%4d%10.4f%13.6f%13.6f%13.6f%13.6f" % (i + 1, e[i], *f))
    return "\n".join(out) + "\n"


def cases() -> dict:
    water = open(os.path.join(ROOT, "tests", "golden", "water", "EXCI_TEST1.dat")).read()
    return {
        "water_exci1": water,
        "cd_velo": synthetic(1, 30, ["NM", "VELO"]),
        "uv_velo": synthetic(2, 30, ["NM", "UV", "VELO"]),
        "uv_length_ev": synthetic(3, 25, ["UV"]),
        "nm_cm": synthetic(6, 15, ["NM", "CM", "VELO"]),
        "log": synthetic(7, 20, ["NM", "UV", "LOG"]),
        "inv": synthetic(8, 20, ["NM", "INV", "VELO"]),
        "wide": synthetic(9, 6, ["NM", "UV"], 0.5, 60.0),
        "degenerate": synthetic(10, 12, ["NM", "UV", "VELO"], degenerate=3),
        "ethr": synthetic(11, 20, ["NM", "UV"], extra=(" ETHR", "  0.3")),
        "forb": synthetic(12, 20, ["NM", "UV"], extra=(" FORB", "  0.01")),
        "au": synthetic(13, 10, ["AU", "UV"], 0.1, 0.5),
        "mix": synthetic(14, 10, ["NM", "MIX"]),
    }


def main() -> None:
    if not os.path.exists(REF):
        raise SystemExit("build the reference program first: make -C oracle ref")
    os.makedirs(OUT, exist_ok=True)
    for name, text in cases().items():
        with tempfile.TemporaryDirectory() as d:
            with open(os.path.join(d, "tda.dat"), "w") as f:
                f.write(text)
            with open(os.path.join(d, "tda.dat")) as fin:
                subprocess.run([REF], stdin=fin, stdout=subprocess.DEVNULL, cwd=d, check=True)
            dst = os.path.join(OUT, name)
            os.makedirs(dst, exist_ok=True)
            shutil.copy(os.path.join(d, "tda.dat"), os.path.join(dst, "tda.dat"))
            shutil.copy(os.path.join(d, "rots.dat"), os.path.join(dst, "rots.dat"))
            with open(os.path.join(d, "spec.dat"), "rb") as f:
                data = f.read()
            with open(os.path.join(dst, "spec.dat.xz"), "wb") as f:
                f.write(lzma.compress(data, preset=9))
        print(name)


if __name__ == "__main__":
    main()
