"""The matrix-core GEMM behind the operator contractions (stda_hip_gemm, gemm.hip) against
numpy, every transpose combination and ragged edge, in its three arithmetic modes.

Bounds are elementwise, relative to (|op(A)| |op(B)|)_ij: FP64 1e-13 (K-independent for
these K); FP32 K * 2^-23; split-bfloat16 (hi*hi + hi*lo + lo*hi, FP32 accumulation)
K * 2^-23 + 3 * 2^-17, and an order of magnitude below what one bfloat16 product gives,
so that a dropped correction term fails."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SHAPES = [(64, 64, 16), (77, 130, 45), (1, 1, 1), (200, 33, 3), (130, 257, 700), (5, 9, 0)]


@pytest.fixture(scope="module")
def hip():
    from stda_amd.core import StdaHip
    with StdaHip(0) as h:
        yield h


def _operands(rng, m, n, k, ta, tb):
    A = rng.standard_normal((k, m) if ta else (m, k))
    B = rng.standard_normal((n, k) if tb else (k, n))
    C = rng.standard_normal((m, n))
    return A, B, C


@pytest.mark.parametrize("mode", ["fp64", "fp32", "bf16"])
@pytest.mark.parametrize("ta,tb", [(0, 0), (0, 1), (1, 0), (1, 1)])
def test_gemm_against_numpy(hip, mode, ta, tb):
    rng = np.random.default_rng(11 + 2 * ta + tb)
    print(f"gemm {mode} ta={ta} tb={tb}", flush=True)
    for m, n, k in SHAPES:
        A, B, C = _operands(rng, m, n, k, ta, tb)
        opA, opB = (A.T if ta else A), (B.T if tb else B)
        alpha, beta = 0.75, -0.5
        got = hip.gemm(A, B, C, alpha=alpha, beta=beta, ta=bool(ta), tb=bool(tb), mode=mode)
        if mode == "fp64":
            ref = alpha * (opA @ opB) + beta * C
            scale = np.abs(alpha) * (np.abs(opA) @ np.abs(opB)) + abs(beta) * np.abs(C) + 1e-300
            bound = 1e-13
        else:
            f = lambda x: x.astype(np.float32).astype(np.float64)
            ref = alpha * (f(opA) @ f(opB)) + beta * f(C)
            scale = np.abs(alpha) * (np.abs(f(opA)) @ np.abs(f(opB))) + abs(beta) * np.abs(C) + 1e-30
            bound = (max(k, 1) + 2) * 2.0 ** -23 + (3 * 2.0 ** -17 if mode == "bf16" else 0.0)
        err = np.abs(got - ref) / scale
        assert err.max() <= bound, (m, n, k, err.max(), bound)
        if mode == "bf16" and k >= 16:
            # one bfloat16 product would be off by ~2^-9 relative: the split must do better
            assert err.max() < 2.0 ** -12, (m, n, k, err.max())


def test_bf16_precision_solvers(hip):
    """The split-bfloat16 precision mode of the solvers (bfloat16 phase, FP32 phase, FP64
    refinement): roots within 1 meV and fL/fV within 1e-4 of FP64, per-root error reported;
    alpha (which finishes in FP64) within 1e-6 relative."""
    from stda_amd import synth
    from stda_amd.prep import prepare
    EV = 27.211385050
    mol, ints = synth.cluster(30, seed=3, basis="minimal")
    ci = prepare(mol, 10.0, 0.25, ints)
    hip.setup(ci, 10.0, 0.25)
    hip.set_precision("fp64")
    t64 = hip.solve_tda(12, vectors=False)
    r64 = hip.solve_rpa(8)
    a64 = hip.lresp([0.0, 0.05])
    hip.set_precision("bf16")
    try:
        tb = hip.solve_tda(12, vectors=False)
        rb = hip.solve_rpa(8)
        ab = hip.lresp([0.0, 0.05], tol=1e-8)
    finally:
        hip.set_precision("fp64")
    de_t = np.abs(tb.omega - t64.omega) * EV
    de_r = np.abs(rb.omega - r64.omega) * EV
    print(f"bf16 TDA per-root |dE| (eV): {np.array2string(de_t, precision=1)}", flush=True)
    print(f"bf16 RPA per-root |dE| (eV): {np.array2string(de_r, precision=1)}", flush=True)
    assert de_t.max() <= 1e-3 and de_r.max() <= 1e-3
    assert np.abs(tb.props[:, :2] - t64.props[:, :2]).max() <= 1e-4
    assert np.abs(rb.props[:, :2] - r64.props[:, :2]).max() <= 1e-4
    for (w1, a1), (w2, a2) in zip(ab, a64):
        assert w1 == w2
        np.testing.assert_allclose(a1, a2, rtol=0, atol=1e-6 * np.abs(a2).max())
