"""Gaussian basis-set data for the synthetic molecules.

Two families of Cartesian s/p contractions are provided:

* ``minimal``  -- STO-3G-type: every shell is the 3-Gaussian least-squares fit
  of a Slater function, scaled by the standard Slater exponent zeta of the
  shell (alpha = alpha(zeta=1) * zeta**2).  For O and H this reproduces the
  contraction held in the reference's water test fixture
  (tests/water_sto3g.molden, 130.7093214 = 2.227660584 * 7.66**2).
* ``split``    -- a 3-21G-type split-valence set: a 3-Gaussian core s shell,
  an inner 2-Gaussian valence sp shell and an outer 1-Gaussian sp shell
  (H: inner 2-Gaussian s + outer 1-Gaussian s).  Exponents are those of the
  carbon/hydrogen set scaled by (zeta/zeta_C)**2 for N and O.

The values only have to be a sensible, deterministic basis for synthetic
inputs; they are not claimed to be any published basis to the last digit.
Each shell is ``(kind, [(alpha, c_s[, c_p]), ...])`` with ``kind`` in
``{"s", "p", "sp"}`` and contraction coefficients referring to normalised
primitives, the convention of molden files.
"""

from __future__ import annotations

# 3-Gaussian fits of Slater functions with zeta = 1 (Hehre, Stewart, Pople 1969).
_STO3G_1S = [(2.227660584, 0.1543289673), (0.4057711562, 0.5353281423),
             (0.1098175104, 0.4446345422)]
_STO3G_2SP = [(0.9942027990, -0.09996722919, 0.1559162750),
              (0.2310313333, 0.3995128261, 0.6076837186),
              (0.07513858573, 0.7001154689, 0.3919573931)]

# Slater exponents (1s, 2sp) per element.
_ZETA = {1: (1.24, None), 6: (5.67, 1.72), 7: (6.67, 1.95), 8: (7.66, 2.25)}

ELEMENT_SYMBOL = {1: "h", 6: "c", 7: "n", 8: "o"}
SYMBOL_Z = {v: k for k, v in ELEMENT_SYMBOL.items()}


def _scaled(prims, zeta):
    z2 = zeta * zeta
    return [(a * z2,) + tuple(c) for a, *c in prims]


def minimal_shells(z: int):
    """STO-3G-type shells of element ``z``."""
    z1s, z2sp = _ZETA[z]
    shells = [("s", _scaled(_STO3G_1S, z1s))]
    if z2sp is not None:
        shells.append(("sp", _scaled(_STO3G_2SP, z2sp)))
    return shells


# 3-21G-type carbon and hydrogen data; N and O exponents are scaled.
_SV_C_CORE = [(172.2560, 0.0617669), (25.91090, 0.358794), (5.533350, 0.700713)]
_SV_C_IN = [(3.664980, -0.395897, 0.236460), (0.770545, 1.215840, 0.860619)]
_SV_C_OUT = [(0.195857, 1.0, 1.0)]
_SV_H_IN = [(5.447178, 0.156285), (0.824547, 0.904691)]
_SV_H_OUT = [(0.183192, 1.0)]


def split_shells(z: int):
    """3-21G-type split-valence shells of element ``z``."""
    if z == 1:
        return [("s", list(_SV_H_IN)), ("s", list(_SV_H_OUT))]
    zc1, zc2 = _ZETA[6]
    z1, z2 = _ZETA[z]
    f1 = (z1 / zc1) ** 2
    f2 = (z2 / zc2) ** 2
    core = [(a * f1, c) for a, c in _SV_C_CORE]
    inner = [(a * f2, cs, cp) for a, cs, cp in _SV_C_IN]
    outer = [(a * f2, cs, cp) for a, cs, cp in _SV_C_OUT]
    return [("s", core), ("sp", inner), ("sp", outer)]


def shells_for(z: int, family: str):
    if family == "minimal":
        return minimal_shells(z)
    if family == "split":
        return split_shells(z)
    raise ValueError(f"unknown basis family {family!r}")


def n_cart(kind: str) -> int:
    return {"s": 1, "p": 3, "sp": 4, "d": 6}[kind]
