"""Upstream preparation: from a molecule to the FP64 inputs of the core.

This is the part of ``stda`` that runs before the hot path (stda.f:142-452):

1. orbital window (stda.f:142-307): deps = 2 (1 + 0.8 ax) thr;
   occupied MOs with eps > LUMO - deps, virtual MOs with eps < HOMO + deps;
2. ``sort_vec`` (stda.f:333, sutda.f:1789-1816), including its quirk of
   permuting the caller's *full* coefficient array, not the windowed copy;
3. AO -> MO transformation of the dipole-length, angular-momentum and
   velocity integrals (``onetri``, onetri.f:20-55; ``shrink``, stda.f:1695-1706);
4. Loewdin orthogonalisation C' = S^1/2 C (``makel``, stda.f:1615-1691;
   stda.f:430-452) with the |<c'|c'> - 1| > 1e-2 norm check.

The result, :class:`CoreInputs`, is exactly what the excited-state core
consumes at its FP64 boundary.  Arrays use the reference's conventions:
column-major ``clow(nao, moci)``, 1-based ``iaoat``, ``xyz(4, ncent)`` in Bohr
with the nuclear charge in row 4, packed upper-triangle MO moments of length
moci (moci + 1) / 2 where element lin(i, j) (i >= j, 1-based) holds M(j, i).
"""

from __future__ import annotations

import warnings
from dataclasses import dataclass, field

import numpy as np

from .aoints import one_electron
from .constants import HARTREE_EV, molecular_weight
from .molden import Molecule


class StdaInputError(RuntimeError):
    """Raised where the reference stops with a message."""


@dataclass
class CoreInputs:
    ncent: int
    nao: int
    no: int
    nv: int
    clow: np.ndarray        # (nao, moci) Fortran order
    epsi: np.ndarray        # (moci,)
    iaoat: np.ndarray       # (nao,) 1-based atom of each AO, int32
    xyz: np.ndarray         # (4, ncent) Bohr + nuclear charge
    moments: np.ndarray     # (9, moci (moci+1)/2): xl yl zl xm ym zm xv yv zv
    xmolw: float = 0.0
    meta: dict = field(default_factory=dict)

    @property
    def moci(self) -> int:
        return self.no + self.nv


def packed_index(i, j):
    """0-based packed index of the pair (i, j) (0-based MO indices)."""
    a = np.maximum(i, j)
    b = np.minimum(i, j)
    return b + a * (a + 1) // 2


def shrink(M: np.ndarray) -> np.ndarray:
    """Packed upper triangle: element lin(i,j) holds M(j,i), j <= i (stda.f:1695-1706)."""
    n = M.shape[0]
    iu = np.triu_indices(n)          # rows j <= cols i, row-major over (j, i)
    out = np.empty(n * (n + 1) // 2)
    out[packed_index(iu[1], iu[0])] = M[iu]
    return out


def _selection_sort_quirk(eps, C_full, moci):
    """sort_vec: selection sort of eps(1:moci), swapping columns of C_full."""
    eps = eps.copy()
    swapped = False
    for i in range(moci - 1):
        k = i
        pp = eps[i]
        for j in range(i + 1, moci):
            if eps[j] < pp:
                k = j
                pp = eps[j]
        if k != i:
            eps[k] = eps[i]
            eps[i] = pp
            C_full[:, [i, k]] = C_full[:, [k, i]]
            swapped = True
    return eps, swapped


def prepare(mol: Molecule, thr_ev: float, ax: float, ints: dict | None = None,
            bfw: bool = False, device=None) -> CoreInputs:
    """Build the core's FP64 inputs for molecule ``mol`` (see module docstring).

    ``device``: a :class:`stda_amd.core.StdaHip` context; the AO integrals (when ``ints``
    is not given: ``stda_hip_ao_ints``), the moment transforms and the Loewdin
    orthogonalisation (steps 3-4: ``stda_hip_transform``) then run on it, otherwise with
    numpy on the host."""
    eps, occ, C = mol.eps, mol.occ, mol.C
    thr = thr_ev / HARTREE_EV
    deps = 2.0 * (1.0 + 0.8 * ax) * thr
    occ_mask = occ > 1.990
    vir_mask = occ < 0.010
    omax = eps[occ_mask].max() if occ_mask.any() else -1e42
    vmin = eps[vir_mask].min() if vir_mask.any() else 1e42
    othr = vmin - deps
    vthr = deps + omax
    osel = np.nonzero(occ_mask & (eps > othr))[0]
    vsel = np.nonzero(vir_mask & (eps < vthr))[0]
    no, nv = len(osel), len(vsel)
    if no == 0 or nv == 0:
        raise StdaInputError("no CSF, increase energy threshold (-e option)")
    sel = np.concatenate([osel, vsel])
    ca = C[:, sel].copy()
    epsi = eps[sel].copy()
    C_scratch = C.copy()
    epsi, swapped = _selection_sort_quirk(epsi, C_scratch, len(sel))
    if swapped:
        warnings.warn("MO energies of the window are not ascending: the reference's "
                      "sort_vec pairs energies and orbitals inconsistently here, and "
                      "so does this restatement")
    if device is not None and ints is None:
        from .core import StdaHipError
        try:
            clow, mom = device.prepare_ao(mol, ca, bfw=bfw)
        except StdaHipError as err:
            raise StdaInputError(str(err)) from err
        return _core_inputs(mol, no, nv, clow, epsi, mom, othr, vthr)
    if ints is None:
        ints = one_electron(mol)
    S = ints["S"]
    if device is not None:
        from .core import StdaHipError
        ao9 = [ints["R"][k] for k in range(3)] + [ints["M"][k] for k in range(3)] + \
              [ints["V"][k] for k in range(3)]
        try:
            clow, mom = device.transform(ca, S, ao9, bfw=bfw)
        except StdaHipError as err:
            raise StdaInputError(str(err)) from err
        return _core_inputs(mol, no, nv, clow, epsi, mom, othr, vthr)
    mom = np.empty((9, len(sel) * (len(sel) + 1) // 2))
    for k in range(3):
        mom[k] = shrink(ca.T @ ints["R"][k] @ ca)
        mom[3 + k] = shrink(ca.T @ ints["M"][k] @ ca)
        mom[6 + k] = shrink(ca.T @ ints["V"][k] @ ca)
    e, U = np.linalg.eigh(S)
    if (e < 0).any():
        if not bfw:
            raise StdaInputError("negative eigenvalue of S: must stop in S^1/2")
        e = np.where(e < 0, 0.0, e)
    X = (U * np.sqrt(e)) @ U.T
    clow = X @ ca
    nrm = (clow * clow).sum(axis=0)
    bad = np.nonzero(np.abs(nrm - 1.0) > 1e-2)[0]
    if len(bad):
        raise StdaInputError(f"internal MO norm error (MO {bad[0]+1}, norm {nrm[bad[0]]})")
    return _core_inputs(mol, no, nv, clow, epsi, mom, othr, vthr)


def _core_inputs(mol, no, nv, clow, epsi, mom, othr, vthr) -> CoreInputs:
    xyz = np.zeros((4, mol.natoms))
    xyz[:3] = mol.xyz.T
    xyz[3] = mol.Z
    return CoreInputs(ncent=mol.natoms, nao=mol.nao, no=no, nv=nv,
                      clow=np.asfortranarray(clow), epsi=epsi,
                      iaoat=(mol.ao_atom + 1).astype(np.int32), xyz=np.asfortranarray(xyz),
                      moments=mom, xmolw=molecular_weight(mol.Z),
                      meta={"othr": othr, "vthr": vthr})
