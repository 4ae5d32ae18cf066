"""One-electron integrals over contracted Cartesian Gaussian AOs.

Computes, for the AO basis of a :class:`~stda_amd.molden.Molecule`:

* ``S``   overlap <mu|nu>;
* ``R``   dipole-length integrals in the reference's sign convention,
          R_k = -<mu| r_k |nu> with the origin at (0,0,0)
          (libcint.f:252-352 stores ``-buf`` of ``cint1e_r_cart``; the common
          origin env(2:4) is zero, libcint.f:103);
* ``V``   velocity integrals V_k = -<mu| d/dr_k |nu> (``cint1e_p_cart``,
          libcint.f:653-756; the operator is -i nabla with the imaginary unit
          dropped, which makes the velocity and length forms of the rotatory
          strength agree in sign through the hypervirial relation);
* ``M``   angular-momentum integrals M_k = <mu| (r x nabla)_k |nu>, origin 0,
          the quantity ``opam`` forms (intpack.f90:1208-1283, used by
          intslvm2 at intslvm.f:213-332).

All matrices are returned full (nao x nao).  ``V`` and ``M`` are
antisymmetric; the reference stores their lower triangle and blows them up
antisymmetrically (onetri.f:56-84), so only the sign convention of the lower
triangle matters, and it is the one stated above.

The primitive integrals use the Obara-Saika recurrences for 1-D overlaps,
vectorised over primitive pairs with numpy.
"""

from __future__ import annotations

import math

import numpy as np

from .molden import Molecule


def _prims(mol: Molecule):
    ptr = mol.ao_prim_ptr
    nprim = int(ptr[-1])
    ao_of = np.repeat(np.arange(mol.nao), np.diff(ptr))
    center = mol.xyz[mol.ao_atom[ao_of]]
    lmn = mol.ao_lmn[ao_of]
    return nprim, ao_of, center, lmn, mol.prim_alpha, mol.prim_coef


def one_electron(mol: Molecule, want=("S", "R", "V", "M"), block: int = 192):
    """Return a dict with the requested AO matrices (see module docstring)."""
    nprim, ao_of, cen, lmn, alpha, coef = _prims(mol)
    nao = mol.nao
    ptr = mol.ao_prim_ptr
    lmax = int(lmn.max()) if nprim else 0
    L1, L2 = lmax + 1, lmax + 2
    out = {}
    if "S" in want:
        out["S"] = np.zeros((nao, nao))
    for key in ("R", "V", "M"):
        if key in want:
            out[key] = np.zeros((3, nao, nao))
    bj = alpha[None, :]
    Bj = cen[None, :, :]
    for i0 in range(0, nprim, block):
        i1 = min(nprim, i0 + block)
        ai = alpha[i0:i1, None]
        Ai = cen[i0:i1, None, :]
        p = ai + bj
        inv2p = 0.5 / p
        mu = ai * bj / p
        P = (ai[..., None] * Ai + bj[..., None] * Bj) / p[..., None]
        s_d, x_d, g_d = [], [], []
        for d in range(3):
            XPA = P[..., d] - Ai[..., d]
            XPB = P[..., d] - Bj[..., d]
            XAB = Ai[..., d] - Bj[..., d]
            S = [[None] * (L2 + 1) for _ in range(L1)]
            S[0][0] = np.sqrt(math.pi / p) * np.exp(-mu * XAB * XAB)
            for t in range(L1):
                if t > 0:
                    prev = S[t - 1][0]
                    S[t][0] = XPA * prev + ((t - 1) * S[t - 2][0] if t >= 2 else 0.0) * inv2p
                for u in range(1, L2 + 1):
                    acc = XPB * S[t][u - 1]
                    if t > 0:
                        acc = acc + t * S[t - 1][u - 1] * inv2p
                    if u > 1:
                        acc = acc + (u - 1) * S[t][u - 2] * inv2p
                    S[t][u] = acc
            la = lmn[i0:i1, d][:, None]
            lb = lmn[:, d][None, :]
            T = np.stack([np.stack(row) for row in S])       # (L1, L2+1, nb, np)
            ii = np.broadcast_to(la, p.shape)
            jj = np.broadcast_to(lb, p.shape)
            r_idx = np.arange(p.shape[0])[:, None]
            c_idx = np.arange(p.shape[1])[None, :]
            s_ = T[ii, jj, r_idx, c_idx]
            sp1 = T[ii, jj + 1, r_idx, c_idx]
            sm1 = np.where(jj > 0, T[ii, np.maximum(jj - 1, 0), r_idx, c_idx], 0.0)
            s_d.append(s_)
            x_d.append(sp1 + Bj[..., d] * s_)
            g_d.append(jj * sm1 - 2.0 * bj * sp1)
        w = coef[i0:i1, None] * coef[None, :]

        def contract(v):
            v = np.add.reduceat(v * w, ptr[:-1], axis=1)          # ket -> AO
            return v

        rows = slice(int(ao_of[i0]), int(ao_of[i1 - 1]) + 1)

        def accumulate(dst, v):
            vk = contract(v)
            # bra primitives -> AOs (this block may split an AO; add.at handles it)
            np.add.at(dst, ao_of[i0:i1], vk)

        sx, sy, sz = s_d
        if "S" in want:
            accumulate(out["S"], sx * sy * sz)
        if "R" in want:
            accumulate(out["R"][0], -(x_d[0] * sy * sz))
            accumulate(out["R"][1], -(sx * x_d[1] * sz))
            accumulate(out["R"][2], -(sx * sy * x_d[2]))
        if "V" in want:
            accumulate(out["V"][0], -(g_d[0] * sy * sz))
            accumulate(out["V"][1], -(sx * g_d[1] * sz))
            accumulate(out["V"][2], -(sx * sy * g_d[2]))
        if "M" in want:
            accumulate(out["M"][0], sx * (x_d[1] * g_d[2] - g_d[1] * x_d[2]))
            accumulate(out["M"][1], sy * (x_d[2] * g_d[0] - g_d[2] * x_d[0]))
            accumulate(out["M"][2], sz * (x_d[0] * g_d[1] - g_d[0] * x_d[1]))
        del rows
    return out
