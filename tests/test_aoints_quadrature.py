"""The one-electron AO integrals against brute-force quadrature on a 3-D grid.

Conventions (the reference's, as the integral driver stores them):
* dipole length R_k = -<mu| r_k |nu>, origin 0 (libcint.f:252-352 stores -buf of
  cint1e_r_cart);
* velocity V_k = -<mu| d/dr_k |nu> (cint1e_p_cart, libcint.f:653-756: libcint's
  p = -i nabla with the imaginary unit dropped);
* angular momentum M_k = <mu| (r x nabla)_k |nu>, origin 0, the derivative acting on the
  ket (opam, intpack.f90:1208-1283: v(1) = pre s2 (sy s9 - sz s6) with s9, s6 the ket
  derivatives of bip, intpack.f90:1289-1345, and sy, sz the moments).

The grid values of every AO and of its analytic gradient come from the definition of a
contracted Cartesian Gaussian, independently of the Obara-Saika recurrences in
stda_amd/aoints.py; a uniform grid of spacing h integrates these smooth, rapidly
decaying products to ~exp(-pi^2 / (p h^2)) (p the largest exponent sum), far below the
tolerance.  A wrong sign, a factor, a derivative on the bra instead of the ket, or
swapped components in r x nabla each fail one of the checks."""
import numpy as np
import pytest

from stda_amd.aoints import one_electron
from stda_amd.molden import build_ao_basis


@pytest.fixture(scope="module")
def tiny():
    # three centres (Bohr), diffuse s, sp, p and d shells so that a 0.16 Bohr grid suffices
    xyz = np.array([[0.0, 0.0, 0.0], [1.3, -0.4, 0.7], [-0.6, 1.1, -0.9]])
    shells = [
        [("s", [(1.1, 0.6), (0.35, 0.5)]), ("sp", [(0.9, 0.4, 0.5), (0.3, 0.7, 0.6)])],
        [("d", [(0.8, 1.0)]), ("s", [(0.5, 1.0)])],
        [("p", [(0.7, 0.8), (0.25, 0.4)])],
    ]
    mol = build_ao_basis([6, 8, 1], xyz, shells)
    h = 0.16
    ax = np.arange(-9.0, 9.0 + 1e-9, h)
    X, Y, Z = np.meshgrid(ax, ax, ax, indexing="ij")
    R = np.stack([X.ravel(), Y.ravel(), Z.ravel()])          # (3, npts)
    nao = mol.nao
    phi = np.zeros((nao, R.shape[1]))
    grad = np.zeros((3, nao, R.shape[1]))
    for mu in range(nao):
        A = mol.xyz[mol.ao_atom[mu]]
        lmn = mol.ao_lmn[mu]
        d = R - A[:, None]
        r2 = (d * d).sum(0)
        for p in range(mol.ao_prim_ptr[mu], mol.ao_prim_ptr[mu + 1]):
            a, c = mol.prim_alpha[p], mol.prim_coef[p]
            g = c * np.exp(-a * r2)
            poly = [d[k] ** lmn[k] for k in range(3)]
            phi[mu] += g * poly[0] * poly[1] * poly[2]
            for k in range(3):
                dk = (lmn[k] * d[k] ** (lmn[k] - 1) if lmn[k] > 0 else 0.0) - 2.0 * a * d[k] ** (lmn[k] + 1)
                others = np.prod([poly[j] for j in range(3) if j != k], axis=0)
                grad[k, mu] += g * dk * others
    return mol, R, phi, grad, h ** 3


def test_overlap_and_dipole(tiny):
    mol, R, phi, grad, w = tiny
    ints = one_electron(mol)
    S = (phi * w) @ phi.T
    np.testing.assert_allclose(np.diag(S), 1.0, atol=1e-9)         # unit-normalised AOs
    np.testing.assert_allclose(ints["S"], S, atol=1e-9)
    for k in range(3):
        ref = -((phi * R[k] * w) @ phi.T)
        np.testing.assert_allclose(ints["R"][k], ref, atol=1e-9, err_msg=f"dipole {k}")


def test_velocity(tiny):
    mol, R, phi, grad, w = tiny
    V = one_electron(mol)["V"]
    for k in range(3):
        ref = -((phi * w) @ grad[k].T)                                 # -<mu| d_k |nu>
        np.testing.assert_allclose(V[k], ref, atol=1e-9, err_msg=f"velocity {k}")
        np.testing.assert_allclose(V[k], -V[k].T, atol=1e-9)
    assert np.abs(V).max() > 0.1


def test_angular_momentum(tiny):
    mol, R, phi, grad, w = tiny
    M = one_electron(mol)["M"]
    x, y, z = R
    gx, gy, gz = grad
    refs = [y * gz - z * gy, z * gx - x * gz, x * gy - y * gx]        # (r x nabla) on the ket
    for k in range(3):
        ref = (phi * w) @ refs[k].T
        np.testing.assert_allclose(M[k], ref, atol=1e-9, err_msg=f"angular momentum {k}")
        np.testing.assert_allclose(M[k], -M[k].T, atol=1e-9)
    assert np.abs(M).max() > 0.1
