"""xTB input path: the ``wfn.xtb`` wavefunction file and the ``-xtb`` parameter set.

``wfn.xtb`` is a Fortran unformatted sequential file (records framed by 4-byte
length markers, the gfortran/flang default) whose record sequence the
reference's reader defines (readxtb0 / readxtb, readxtb.f:38-200):

1. ``imethod`` (1 = RKS, 2 = UKS)
2. ``ncent, nbf, nmo, nprims``
3. ``ncent`` records: the atom names (character*2)
4. per atom four records: x, y, z (real*8, Bohr), then the nuclear charge (integer)
5. ``nprims`` records each of ``ipty`` (angular type of the primitive), then of
   ``ipat`` (its atom, 1-based), then of ``ipao`` (its contracted AO, 1-based)
6. ``exip(1:nprims)``, then ``cxip(1:nprims)`` (exponents; coefficients that
   include the primitive normalisation: std2 writes ``fnorm`` = 1 for xTB
   input, main.f:832-838)
7. RKS: ``occ(1:nmo)``, ``eps(1:nmo)``, ``cmo(nbf*nmo)`` (column-major nbf x nmo)

A contracted AO is the sum of its primitives (intslvm.f:66-77 maps AOs to atoms
through ``ipao``/``ipat``).  The angular types are those of the reference's
integral code (``lx``/``my``/``nz`` of intpack.f90:1212-1217): 1 s; 2-4 x, y, z;
5-10 xx, yy, zz, xy, xz, yz; 11-20 xxx, yyy, zzz, xxy, xxz, xyy, yyz, xzz, yzz,
xyz.  With fewer MOs than AOs the MOs come from a spherical basis expanded in the
Cartesian AOs (readxtb.f:193-197); the coefficients are used as they stand.

``XTB_PARAMETERS`` is what ``std2 -xtb`` sets (main.f:198-219): a_x 0.5,
alpha 2.0, beta 4.0 and the K(ia,ia) shift (shftmax 0.5 eV, width 0.1 eV,
steepness 4).
"""

from __future__ import annotations

import numpy as np

from .molden import Molecule

XTB_PARAMETERS = dict(ax=0.5, alphak=2.0, betaj=4.0, kshift=(0.5, 0.1, 4.0))

# Cartesian exponents of the angular types 1..20 (intpack.f90:1212-1217, minus one)
IPTY_LMN = [(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1),
            (2, 0, 0), (0, 2, 0), (0, 0, 2), (1, 1, 0), (1, 0, 1), (0, 1, 1),
            (3, 0, 0), (0, 3, 0), (0, 0, 3), (2, 1, 0), (2, 0, 1), (1, 2, 0), (0, 2, 1),
            (1, 0, 2), (0, 1, 2), (1, 1, 1)]
_LMN_IPTY = {lmn: k + 1 for k, lmn in enumerate(IPTY_LMN)}


class XtbFormatError(ValueError):
    pass


class _Records:
    """Sequential unformatted records with 4-byte markers."""

    def __init__(self, data: bytes):
        self.d = data
        self.p = 0

    def next(self) -> bytes:
        if self.p + 4 > len(self.d):
            raise XtbFormatError("wfn.xtb ends before the record sequence does")
        n = int(np.frombuffer(self.d, "<i4", 1, self.p)[0])
        end = self.p + 4 + n
        if n < 0 or end + 4 > len(self.d) or int(np.frombuffer(self.d, "<i4", 1, end)[0]) != n:
            raise XtbFormatError(f"bad record marker at byte {self.p}")
        rec = self.d[self.p + 4:end]
        self.p = end + 4
        return rec

    def ints(self, count=None):
        rec = self.next()
        # default integers: 4 bytes, or 8 bytes from a writer built with -i8
        size = 4 if count is None else len(rec) // count
        if size not in (4, 8) or (count is not None and size * count != len(rec)) \
                or (count is None and len(rec) not in (4, 8)):
            raise XtbFormatError("integer record of unexpected length")
        if count is None:
            return int(np.frombuffer(rec, "<i4" if len(rec) == 4 else "<i8")[0])
        return np.frombuffer(rec, "<i4" if size == 4 else "<i8").astype(np.int64)

    def reals(self, count):
        rec = self.next()
        if len(rec) != 8 * count:
            raise XtbFormatError("real*8 record of unexpected length")
        return np.frombuffer(rec, "<f8").copy()


def read_wfn_xtb(path: str) -> Molecule:
    """Read a closed-shell (RKS) ``wfn.xtb`` into a :class:`Molecule`."""
    try:
        data = open(path, "rb").read()
    except FileNotFoundError:
        raise XtbFormatError("input file not found") from None   # readxtb.f:51-52
    r = _Records(data)
    imethod = r.ints()
    if imethod == 2:
        raise XtbFormatError("UKS wfn.xtb: the excited-state core here is restricted (RKS)")
    if imethod != 1:
        raise XtbFormatError(f"unknown method flag {imethod}")
    ncent, nbf, nmo, nprims = (int(v) for v in r.ints(4))
    if min(ncent, nbf, nmo, nprims) < 1:
        raise XtbFormatError("read error")                           # main.f:684-685
    for _ in range(ncent):
        r.next()                                                     # atom names
    xyz = np.zeros((ncent, 3))
    Z = np.zeros(ncent, dtype=np.int64)
    for i in range(ncent):
        for j in range(3):
            xyz[i, j] = r.reals(1)[0]
        Z[i] = r.ints()
    ipty = np.array([r.ints() for _ in range(nprims)])
    ipat = np.array([r.ints() for _ in range(nprims)])
    ipao = np.array([r.ints() for _ in range(nprims)])
    exip = r.reals(nprims)
    cxip = r.reals(nprims)
    occ = r.reals(nmo)
    eps = r.reals(nmo)
    cmo = r.reals(nbf * nmo).reshape(nbf, nmo, order="F")
    if ipty.min() < 1 or ipty.max() > len(IPTY_LMN):
        raise XtbFormatError("angular type outside 1..20")
    if ipat.min() < 1 or ipat.max() > ncent or ipao.min() < 1 or ipao.max() > nbf:
        raise XtbFormatError("primitive atom or AO index out of range")
    # primitives grouped by AO, in file order within each AO
    order = np.argsort(ipao, kind="stable")
    counts = np.bincount(ipao - 1, minlength=nbf)
    if np.any(counts == 0):
        raise XtbFormatError("an AO without primitives")
    ptr = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
    ao_atom = np.zeros(nbf, dtype=np.int64)
    ao_lmn = np.zeros((nbf, 3), dtype=np.int64)
    for mu in range(nbf):
        prims = order[ptr[mu]:ptr[mu + 1]]
        if np.unique(ipat[prims]).size != 1 or np.unique(ipty[prims]).size != 1:
            raise XtbFormatError(f"AO {mu + 1} mixes atoms or angular types")
        ao_atom[mu] = ipat[prims[0]] - 1
        ao_lmn[mu] = IPTY_LMN[ipty[prims[0]] - 1]
    return Molecule(Z=Z, xyz=xyz, ao_atom=ao_atom, ao_lmn=ao_lmn, ao_prim_ptr=ptr,
                    prim_alpha=exip[order], prim_coef=cxip[order], C=cmo, eps=eps, occ=occ,
                    shells=[])


def write_wfn_xtb(path: str, mol: Molecule, names=None) -> None:
    """Write ``mol`` as an RKS ``wfn.xtb`` in the record sequence read above."""
    def rec(payload: bytes, fh):
        n = np.array([len(payload)], dtype="<i4").tobytes()
        fh.write(n + payload + n)

    def i4(*v):
        return np.array(v, dtype="<i4").tobytes()

    ncent, nbf = len(mol.Z), mol.nao
    nmo = mol.C.shape[1]
    nprims = int(mol.ao_prim_ptr[-1])
    ao_of = np.repeat(np.arange(nbf), np.diff(mol.ao_prim_ptr))
    with open(path, "wb") as fh:
        rec(i4(1), fh)
        rec(i4(ncent, nbf, nmo, nprims), fh)
        for i in range(ncent):
            nm = (names[i] if names else "x").ljust(2)[:2]
            rec(nm.encode(), fh)
        for i in range(ncent):
            for j in range(3):
                rec(np.array([mol.xyz[i, j]], dtype="<f8").tobytes(), fh)
            rec(i4(int(mol.Z[i])), fh)
        for p in range(nprims):
            rec(i4(_LMN_IPTY[tuple(int(x) for x in mol.ao_lmn[ao_of[p]])]), fh)
        for p in range(nprims):
            rec(i4(int(mol.ao_atom[ao_of[p]]) + 1), fh)
        for p in range(nprims):
            rec(i4(int(ao_of[p]) + 1), fh)
        rec(np.asarray(mol.prim_alpha, dtype="<f8").tobytes(), fh)
        rec(np.asarray(mol.prim_coef, dtype="<f8").tobytes(), fh)
        rec(np.asarray(mol.occ, dtype="<f8").tobytes(), fh)
        rec(np.asarray(mol.eps, dtype="<f8").tobytes(), fh)
        rec(np.asarray(mol.C, dtype="<f8").ravel(order="F").tobytes(), fh)
