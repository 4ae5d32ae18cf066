"""Molden wavefunction files: reader and writer for Cartesian s/p/sp/d bases.

Reading follows the rules of the reference's molden reader for the
"TERACHEM/GAUSSIAN" input style (``-sty 3``), the style its tests use
(readbasmold.f:217-740, normalize.f:23-123):

* ``[Atoms] Angs`` coordinates are divided by 0.52917721092 (readbasmold.f:273);
  ``[Atoms] AU`` coordinates are taken as Bohr.  The nuclear charge is the
  third field of each atom line.
* Contraction coefficients in the file refer to normalised primitives; each
  primitive coefficient is divided by the square root of the self-overlap of
  the unnormalised Cartesian primitive (readbasmold.f:380-395, 420-433) and
  then every contracted Cartesian function is rescaled to unit self-overlap
  (``normalize`` with a Cartesian basis, normalize.f:36-70).  The net result,
  which is what this module implements, is that every contracted Cartesian AO
  x^l y^m z^n sum_k d_k exp(-a_k r^2) has unit norm.
* AO order per shell: s; p_x p_y p_z; sp = s then p; Cartesian d as
  xx yy zz xy xz yz (readbasmold.f:426-470, the molden order).
* MO blocks ``Ene=``, ``Occup=`` and ``index coefficient`` lines; missing
  coefficients are zero (readbasmold.f:583-640).  Energies are in Hartree.

Spherical bases ([5D], [5D7F], ...) are rejected, as in readbasmold.f:93-106.
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from .basis import ELEMENT_SYMBOL

BOHR_ANGSTROM = 0.52917721092  # readbasmold.f:273

_CART = {
    "s": [(0, 0, 0)],
    "p": [(1, 0, 0), (0, 1, 0), (0, 0, 1)],
    "d": [(2, 0, 0), (0, 2, 0), (0, 0, 2), (1, 1, 0), (1, 0, 1), (0, 1, 1)],
}


def _dfact(n: int) -> float:
    """(n)!! for n >= -1."""
    r = 1.0
    while n > 1:
        r *= n
        n -= 2
    return r


def prim_self_overlap(alpha, lmn):
    """Self-overlap of the unnormalised primitive x^l y^m z^n exp(-alpha r^2)."""
    l, m, n = lmn
    L = l + m + n
    a2 = 2.0 * np.asarray(alpha, dtype=np.float64)
    return ((math.pi / a2) ** 1.5 * _dfact(2 * l - 1) * _dfact(2 * m - 1)
            * _dfact(2 * n - 1) / (2.0 * a2) ** L)


@dataclass
class Molecule:
    """A closed-shell molecule with a contracted Cartesian AO basis and MOs.

    Attributes are in atomic units.  ``ao_*`` arrays have one entry per AO;
    primitive arrays are indexed by ``ao_prim_ptr[k]:ao_prim_ptr[k+1]``.
    ``prim_coef`` holds the final coefficients d_k of *unnormalised*
    primitives, such that every AO has unit self-overlap.
    """

    Z: np.ndarray                       # (Na,) nuclear charges
    xyz: np.ndarray                     # (Na, 3) Bohr
    ao_atom: np.ndarray                 # (nao,) 0-based atom of each AO
    ao_lmn: np.ndarray                  # (nao, 3) Cartesian exponents
    ao_prim_ptr: np.ndarray             # (nao+1,)
    prim_alpha: np.ndarray              # (nprim,)
    prim_coef: np.ndarray               # (nprim,)
    C: np.ndarray = None                # (nao, nmo) MO coefficients
    eps: np.ndarray = None              # (nmo,) Hartree
    occ: np.ndarray = None              # (nmo,)
    shells: list = field(default_factory=list)  # per atom: [(kind, prims)]

    @property
    def natoms(self) -> int:
        return int(self.Z.shape[0])

    @property
    def nao(self) -> int:
        return int(self.ao_atom.shape[0])


def build_ao_basis(Z, xyz, shells_per_atom):
    """Expand shells into contracted, unit-normalised Cartesian AOs."""
    ao_atom, ao_lmn, ptr, alph, coef = [], [], [0], [], []
    for iat, shells in enumerate(shells_per_atom):
        for kind, prims in shells:
            parts = [("s", 1), ("p", 2)] if kind == "sp" else [(kind, 1)]
            for comp, col in parts:
                a = np.array([p[0] for p in prims], dtype=np.float64)
                c_norm = np.array([p[col] for p in prims], dtype=np.float64)
                for lmn in _CART[comp]:
                    d = c_norm / np.sqrt(prim_self_overlap(a, lmn))
                    # contracted self-overlap, then unit normalisation
                    l, m, n = lmn
                    L = l + m + n
                    pa = a[:, None] + a[None, :]
                    s = (np.outer(d, d) * (math.pi / pa) ** 1.5
                         * _dfact(2 * l - 1) * _dfact(2 * m - 1) * _dfact(2 * n - 1)
                         / (2.0 * pa) ** L).sum()
                    d = d / math.sqrt(s)
                    ao_atom.append(iat)
                    ao_lmn.append(lmn)
                    alph.extend(a.tolist())
                    coef.extend(d.tolist())
                    ptr.append(len(alph))
    return Molecule(Z=np.asarray(Z, dtype=np.int64), xyz=np.asarray(xyz, dtype=np.float64),
                    ao_atom=np.asarray(ao_atom, dtype=np.int64),
                    ao_lmn=np.asarray(ao_lmn, dtype=np.int64),
                    ao_prim_ptr=np.asarray(ptr, dtype=np.int64),
                    prim_alpha=np.asarray(alph), prim_coef=np.asarray(coef),
                    shells=[list(s) for s in shells_per_atom])


def _f(tok: str) -> float:
    return float(tok.replace("D", "E").replace("d", "e"))


def read_molden(path: str) -> Molecule:
    """Read a closed-shell Cartesian molden file (input style ``-sty 3``)."""
    with open(path) as fh:
        lines = fh.read().splitlines()
    low = [ln.strip() for ln in lines]
    for ln in low:
        if ln.upper() in ("[5D]", "[5D10F]", "[5D7F]"):
            raise ValueError("program requires Cartesian basis")
    i = next(k for k, ln in enumerate(low) if ln.startswith("[Atoms]"))
    unit = low[i].split()[1] if len(low[i].split()) > 1 else "AU"
    scale = 1.0 / BOHR_ANGSTROM if unit.lower().startswith("ang") else 1.0
    Z, xyz = [], []
    i += 1
    while i < len(low) and low[i] and not low[i].startswith("["):
        t = low[i].split()
        Z.append(int(t[2]))
        xyz.append([_f(t[3]) * scale, _f(t[4]) * scale, _f(t[5]) * scale])
        i += 1
    na = len(Z)
    i = next(k for k, ln in enumerate(low) if ln.startswith("[GTO]")) + 1
    shells_per_atom = [[] for _ in range(na)]
    while i < len(low) and not low[i].startswith("["):
        t = low[i].split()
        if len(t) == 2 and t[0].isdigit():
            iat = int(t[0]) - 1
            i += 1
            while i < len(low) and low[i] and not low[i].startswith("["):
                t = low[i].split()
                kind, nprim = t[0].lower(), int(t[1])
                if kind not in ("s", "p", "sp", "d"):
                    raise ValueError(f"unsupported shell type {kind!r}")
                prims = []
                for k in range(nprim):
                    pt = low[i + 1 + k].split()
                    prims.append(tuple(_f(x) for x in pt))
                shells_per_atom[iat].append((kind, prims))
                i += 1 + nprim
        else:
            i += 1
    mol = build_ao_basis(Z, xyz, shells_per_atom)
    nao = mol.nao
    # MO section
    i = next(k for k, ln in enumerate(low) if ln.startswith("[MO]")) + 1
    eps, occ, cols = [], [], []
    cur = None
    while i < len(low):
        ln = low[i]
        if ln.startswith("["):
            break
        t = ln.split()
        if len(t) == 2 and t[0] == "Ene=":
            eps.append(_f(t[1]))
        elif len(t) == 2 and t[0] == "Occup=":
            occ.append(_f(t[1]))
            cur = np.zeros(nao)
            cols.append(cur)
        elif len(t) == 2 and cur is not None:
            try:
                idx = int(t[0])
            except ValueError:
                idx = 0
            if 1 <= idx <= nao:
                cur[idx - 1] = _f(t[1])
        i += 1
    mol.C = np.array(cols).T.copy() if cols else np.zeros((nao, 0))
    mol.eps = np.asarray(eps, dtype=np.float64)
    mol.occ = np.asarray(occ, dtype=np.float64)
    return mol


def write_molden(path: str, mol: Molecule) -> None:
    """Write ``mol`` as a Cartesian molden file (Angstrom coordinates).

    MOs are written in the order held in ``mol`` (the synthetic generator
    keeps them in ascending energy, which the reference's selection sort of
    the orbital window requires to be a no-op: stda.f:333 with
    sutda.f:1789-1816 sorts the energies together with the caller's full
    coefficient array rather than the windowed copy it uses afterwards).
    """
    out = ["[Molden Format]", "[Atoms] Angs"]
    for k in range(mol.natoms):
        x, y, z = mol.xyz[k] * BOHR_ANGSTROM
        out.append(f"{ELEMENT_SYMBOL.get(int(mol.Z[k]), 'x').upper():<3s} {k+1:6d} {int(mol.Z[k]):3d} "
                   f"{x:22.15e} {y:22.15e} {z:22.15e}")
    out.append("[GTO]")
    for k, shells in enumerate(mol.shells):
        out.append(f"{k+1:6d} 0")
        for kind, prims in shells:
            out.append(f" {kind} {len(prims)} 1.00")
            for p in prims:
                out.append("  " + " ".join(f"{v:.16e}" for v in p))
        out.append("")
    out.append("[MO]")
    for j in range(mol.C.shape[1]):
        out.append(" Sym= 1")
        out.append(f" Ene= {mol.eps[j]:.16e}")
        out.append(" Spin= Alpha")
        out.append(f" Occup= {mol.occ[j]:.6f}")
        col = mol.C[:, j]
        out.extend(f"{mu+1:6d} {col[mu]:.16e}" for mu in range(col.shape[0]))
    with open(path, "w") as fh:
        fh.write("\n".join(out) + "\n")
