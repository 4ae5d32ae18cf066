"""Deterministic synthetic molecules for tests and benchmarks.

Every molecule is closed shell with Cartesian s/p AOs.  MO coefficients are
C = S^{-1/2} U with U a Haar-random orthogonal matrix (QR of a Gaussian
matrix with the sign of diag(R) folded in), so that C^T S C = 1 and the
Loewdin orbitals S^{1/2} C equal U.  Occupied energies are uniform in
[-20, -5] eV and virtual energies uniform in [-1, 15] eV; MOs are stored in
ascending energy, occupied first, so the reference's orbital sort is a no-op.
All randomness comes from ``numpy.random.Generator(PCG64(seed))``.

Workloads:

* :func:`polyene` -- all-trans planar C_nH_{n+2} zig-zag chain, C-C bonds
  1.40 +- 0.05 Angstrom (uniform jitter), C-H 1.08 Angstrom, minimal basis.
  C14H16 with 37 occupied MOs is the small CPU-runnable case.
* :func:`cluster` -- ``natoms`` atoms with C:H:N:O = 45:45:5:5 placed
  uniformly at random in a cube (density 0.1 atoms per cubic Angstrom) with
  every pair at least 1.0 Angstrom apart; split-valence or minimal basis;
  ``nocc`` occupied MOs (default: half the electron count).
"""

from __future__ import annotations

import math

import numpy as np

from .basis import shells_for
from .constants import BOHR_ANGSTROM, HARTREE_EV
from .molden import build_ao_basis, Molecule
from .aoints import one_electron


def haar_orthogonal(n: int, rng: np.random.Generator) -> np.ndarray:
    G = rng.standard_normal((n, n))
    Q, R = np.linalg.qr(G)
    return Q * np.sign(np.diag(R))[None, :]


def _finish(mol: Molecule, nocc: int, rng: np.random.Generator, ints=None):
    nao = mol.nao
    if ints is None:
        ints = one_electron(mol)
    U = haar_orthogonal(nao, rng)
    e, V = np.linalg.eigh(ints["S"])
    Sm12 = (V / np.sqrt(e)) @ V.T
    mol.C = Sm12 @ U
    eo = np.sort(rng.uniform(-20.0, -5.0, nocc)) / HARTREE_EV
    ev = np.sort(rng.uniform(-1.0, 15.0, nao - nocc)) / HARTREE_EV
    mol.eps = np.concatenate([eo, ev])
    mol.occ = np.concatenate([np.full(nocc, 2.0), np.zeros(nao - nocc)])
    return mol, ints


def polyene(ncarbon: int = 14, nocc: int = 37, seed: int = 0, basis: str = "minimal"):
    """Planar all-trans polyene C_nH_{n+2}; returns (Molecule, AO integrals)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    pos = [np.zeros(3)]
    for k in range(1, ncarbon):
        b = 1.40 + rng.uniform(-0.05, 0.05)
        d = math.radians(30.0 if k % 2 else -30.0)   # alternating: C-C-C = 120 deg
        pos.append(pos[-1] + b * np.array([math.cos(d), math.sin(d), 0.0]))
    carbons = np.array(pos)
    Hs = []
    for k in range(ncarbon):
        nb = [carbons[j] for j in (k - 1, k + 1) if 0 <= j < ncarbon]
        c = carbons[k]
        if len(nb) == 2:
            v = 2 * c - nb[0] - nb[1]
            v /= np.linalg.norm(v)
            Hs.append(c + 1.08 * v)
        else:
            u = c - nb[0]
            u /= np.linalg.norm(u)
            for s in (+1, -1):
                rot = np.array([[math.cos(s * math.pi / 3), -math.sin(s * math.pi / 3), 0],
                                [math.sin(s * math.pi / 3), math.cos(s * math.pi / 3), 0],
                                [0, 0, 1]])
                Hs.append(c + 1.08 * (rot @ u))
    xyz = np.vstack([carbons, np.array(Hs)]) / BOHR_ANGSTROM
    Z = [6] * ncarbon + [1] * len(Hs)
    mol = build_ao_basis(Z, xyz, [shells_for(z, basis) for z in Z])
    return _finish(mol, nocc, rng)


def cluster(natoms: int, seed: int = 0, basis: str = "split", nocc: int | None = None,
            density: float = 0.1, min_dist: float = 1.0):
    """Random organic-like cluster; returns (Molecule, AO integrals)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    counts = np.floor(np.array([0.45, 0.45, 0.05, 0.05]) * natoms).astype(int)
    counts[0] += natoms - counts.sum()
    Z = np.repeat([6, 1, 7, 8], counts)
    Z = Z[rng.permutation(natoms)]
    L = (natoms / density) ** (1.0 / 3.0)
    pts = np.empty((natoms, 3))
    n = 0
    while n < natoms:
        c = rng.uniform(0.0, L, 3)
        if n == 0 or np.min(np.sum((pts[:n] - c) ** 2, axis=1)) >= min_dist ** 2:
            pts[n] = c
            n += 1
    xyz = pts / BOHR_ANGSTROM
    mol = build_ao_basis(Z, xyz, [shells_for(int(z), basis) for z in Z])
    if nocc is None:
        nocc = int(Z.sum()) // 2
    return _finish(mol, nocc, rng)


def mixed_sizes(n: int, lo: int = 10, hi: int = 200, seed: int = 0) -> np.ndarray:
    """Atom counts log-uniform in [lo, hi] (rounded), one per molecule."""
    rng = np.random.Generator(np.random.PCG64(seed))
    return np.rint(np.exp(rng.uniform(math.log(lo), math.log(hi), n))).astype(int)


def _mixed_one(args):
    from .prep import prepare
    natoms, seed, basis, thr, ax = args
    mol, ints = cluster(int(natoms), seed=int(seed), basis=basis)
    return prepare(mol, thr, ax, ints)


def _mixed_worker_init():
    # one BLAS thread per worker process: the pool itself is the parallelism
    from threadpoolctl import threadpool_limits
    threadpool_limits(1)


def mixed_batch(n: int, lo: int = 10, hi: int = 200, base_seed: int = 1000,
                basis: str = "minimal", thr: float = 7.0, ax: float = 0.25, workers: int = 0):
    """``n`` independent molecules for batched/served solves.

    Molecule ``i`` is :func:`cluster` with ``natoms = mixed_sizes(n, lo, hi,
    base_seed)[i]``, ``seed = base_seed + i``, the given basis and half the
    electron count occupied, prepared for the energy window ``thr`` (eV) and
    ``ax``.  Returns the list of boundary inputs
    (:class:`stda_amd.prep.CoreInputs`) in molecule order.  ``workers`` > 1
    generates them in that many processes (the AO integrals dominate).
    """
    sizes = mixed_sizes(n, lo, hi, base_seed)
    tasks = [(int(sizes[i]), base_seed + i, basis, thr, ax) for i in range(n)]
    if workers and workers > 1:
        import multiprocessing as mp
        # largest first so the pool's tail is short
        order = sorted(range(n), key=lambda i: -sizes[i])
        with mp.get_context("fork").Pool(workers, initializer=_mixed_worker_init) as pool:
            res = pool.map(_mixed_one, [tasks[i] for i in order], chunksize=1)
        out = [None] * n
        for i, r in zip(order, res):
            out[i] = r
        return out
    return [_mixed_one(t) for t in tasks]
