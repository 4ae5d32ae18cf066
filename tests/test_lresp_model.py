"""The library's response-solver loop (restated in numpy, tests/lresp_model.py) against
the oracle's dense solve of the reference's response equations (linear_response.f:90-131).

The cases are chosen where the loop used to fail: a frequency above ~200 RPA poles
(M(w) indefinite, Galerkin residuals oscillate), a subspace capped far below what the
solve needs (restarts), and a cap too small to hold the poles (the loop must report
non-convergence instead of returning an unconverged alpha)."""
import numpy as np
import pytest

import oracle as O
from lresp_model import lresp_model
from stda_amd import synth
from stda_amd.prep import prepare


def _system(kind):
    if kind == "polyene":
        mol, ints = synth.polyene()
        thr = 7.0
    else:
        mol, ints = synth.cluster(30, seed=5, basis="minimal")
        thr = 10.0
    ci = prepare(mol, thr, 0.25, ints)
    inter = O.intermediates(ci, 0.25)
    sel = O.select(ci, inter, thr)
    apb, amb = O.build_apb_amb(ci, inter, sel, 0.25)
    n = sel["nci"]
    APB, AMB = O.unpack_sym(apb, n), O.unpack_sym(amb, n)
    mu = np.ascontiguousarray(O.csf_moments(ci, sel["iconf"])[0:3].T)
    return ci, sel, APB, AMB, mu


@pytest.fixture(scope="module")
def many_poles():
    return _system("cluster")


@pytest.fixture(scope="module")
def polyene_sys():
    return _system("polyene")


def _dense_alpha(ci, sel, APB, AMB, omegas):
    ref = O.lresp(ci, sel["iconf"], APB, AMB, omegas)
    return np.stack([a for _, a in ref])


def _rpa_omegas(APB, AMB):
    L = np.linalg.cholesky(AMB)
    return np.sqrt(np.linalg.eigvalsh(L.T @ APB @ L))


def test_above_many_poles(many_poles):
    """w = 0.34 Eh lies above 199 RPA poles of this 2301-CSF system; the uncapped loop
    converges well before the space is complete and alpha matches the dense solve."""
    ci, sel, APB, AMB, mu = many_poles
    omegas = [0.0, 0.2, 0.3, 0.34]
    w = _rpa_omegas(APB, AMB)
    assert (w < 0.34).sum() > 150
    alpha, status, info = lresp_model(lambda X: (APB @ X, AMB @ X), sel["ed"], mu, omegas,
                                      maxdim=2400)
    assert status == "ok" and info["restarts"] == 0
    assert info["size"] < len(sel["ed"]) // 2
    ref = _dense_alpha(ci, sel, APB, AMB, omegas)
    for f in range(len(omegas)):
        np.testing.assert_allclose(alpha[f], ref[f], rtol=0, atol=1e-8 * np.abs(ref[f]).max())


def test_forced_restarts(polyene_sys):
    """Capped at 48 vectors (the library's STDA_HIP_LRESP_MAXDIM), the loop restarts
    dozens of times, keeps the solutions and the low RPA pairs, and converges."""
    ci, sel, APB, AMB, mu = polyene_sys
    omegas = [0.0, 0.05, 0.1, 0.15]
    alpha, status, info = lresp_model(lambda X: (APB @ X, AMB @ X), sel["ed"], mu, omegas,
                                      maxdim=48)
    assert status == "ok" and info["restarts"] > 5
    ref = _dense_alpha(ci, sel, APB, AMB, omegas)
    for f in range(len(omegas)):
        np.testing.assert_allclose(alpha[f], ref[f], rtol=0, atol=1e-8 * np.abs(ref[f]).max())


def test_cap_below_poles_reports_not_converged(many_poles):
    """A cap of 100 vectors cannot hold the ~200 poles below 0.34 Eh: the loop must end
    as not converged (never as converged with a wrong alpha)."""
    ci, sel, APB, AMB, mu = many_poles
    alpha, status, info = lresp_model(lambda X: (APB @ X, AMB @ X), sel["ed"], mu,
                                      [0.0, 0.2, 0.3, 0.34], maxdim=100, maxit=60)
    assert status == "not converged" and alpha is None
