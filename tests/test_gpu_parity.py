"""Parity of the HIP path (through the C ABI) with the FP64 oracle.

Tolerances (plain FP64 path): excitation energies 1e-6 eV, oscillator
strengths 1e-6 absolute, rotatory strengths 2e-5 relative, polarizability tensors
1e-6 relative to the largest component; the CSF selection (iconf) must be identical and the
PT2-corrected diagonal equal to 1e-10 Hartree.
"""
import os
import sys

import numpy as np
import pytest

from conftest import GOLDEN, beta_test_omegas, read_datxy, read_table
import oracle as O
from stda_amd.prep import prepare

pytestmark = pytest.mark.gpu

EV = 27.211385050
E_TOL = 1e-6 / EV          # 1e-6 eV in Hartree
F_TOL = 1e-6
R_RTOL = 2e-5              # rotatory strengths (|R| up to ~1e3): relative


def assert_props(p, ref):
    """fL, fV within 1e-6 absolute; RL, RV within 2e-5 relative (+1e-6).

    With a Davidson residual of 1e-8 the eigenvectors carry errors of order
    |r| / gap; R = -235.722 mu.m amplifies them, hence the relative bound."""
    np.testing.assert_allclose(p[:, :2], ref[:, :2], atol=F_TOL, rtol=0)
    np.testing.assert_allclose(p[:, 2:], ref[:, 2:], atol=F_TOL, rtol=R_RTOL)


def _hip():
    from stda_amd.core import StdaHip
    return StdaHip(0)


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


WATER = {1: (1.0, 20.0, -100.0, -100.0), 2: (0.5, 30.0, -100.0, -100.0), 3: (1.0, 30.0, 3.0, 1.0)}


@pytest.mark.parametrize("case", [1, 2, 3])
def test_water_tda_window(water, case):
    ax, thr, al, be = WATER[case]
    say(f"water TDA case {case}")
    ci = prepare(water, thr, ax)
    ref = O.run(ci, thr, ax, alphak=al, betaj=be)
    with _hip() as h:
        h.setup(ci, thr, ax, alphak=al, betaj=be)
        iconf, ed = h.csf()
        sol = h.solve_tda(None)
    np.testing.assert_array_equal(iconf, ref.iconf)
    np.testing.assert_allclose(ed, ref.ed, atol=1e-10)
    assert len(sol.omega) == len(ref.omega)
    np.testing.assert_allclose(sol.omega, ref.omega, atol=E_TOL)
    assert_props(sol.props, ref.props)
    exp = read_datxy(os.path.join(GOLDEN, "water", f"EXCI_TEST{case}.dat"))
    for k, row in enumerate(exp):
        assert abs(sol.omega[k] * 27.21139 - row[1]) <= 0.5e-4 + 1e-6
        assert abs(sol.props[k, 0] - row[2]) <= 2.5e-6


@pytest.mark.parametrize("case", [1, 2, 3])
def test_water_rpa_window(water, case):
    ax, thr, al, be = WATER[case]
    say(f"water RPA case {case}")
    ci = prepare(water, thr, ax)
    ref = O.run(ci, thr, ax, alphak=al, betaj=be, rpa=True)
    with _hip() as h:
        h.setup(ci, thr, ax, alphak=al, betaj=be)
        sol = h.solve_rpa(None)
    assert len(sol.omega) == len(ref.omega)
    np.testing.assert_allclose(sol.omega, ref.omega, atol=E_TOL)
    assert_props(sol.props, ref.props)


@pytest.fixture(scope="module")
def polyene():
    from stda_amd import synth
    mol, ints = synth.polyene()
    return prepare(mol, 7.0, 0.25, ints)


@pytest.fixture(scope="module")
def cluster30():
    from stda_amd import synth
    mol, ints = synth.cluster(30, seed=3, basis="minimal")
    return prepare(mol, 10.0, 0.25, ints)


def _compare_selection(h, ref):
    iconf, ed = h.csf()
    assert h.nprim == ref.nex and h.nci == ref.nci
    np.testing.assert_array_equal(iconf, ref.iconf)
    np.testing.assert_allclose(ed, ref.ed, atol=1e-10)


def _compare_vectors(X, Xref):
    """|cos| of the angle between matching columns (sign and norm free)."""
    ov = np.abs(np.sum(X * Xref, axis=0)) / (np.linalg.norm(X, axis=0) *
                                             np.linalg.norm(Xref, axis=0))
    np.testing.assert_allclose(ov, 1.0, atol=1e-6)


def test_polyene_tda_lowest_roots(polyene):
    say("polyene TDA k=10")
    ci = polyene
    ref = O.run(ci, 7.0, 0.25, nroots=10)
    with _hip() as h:
        h.setup(ci, 7.0, 0.25)
        _compare_selection(h, ref)
        sol = h.solve_tda(10)
    np.testing.assert_allclose(sol.omega, ref.omega, atol=E_TOL)
    assert_props(sol.props, ref.props)
    np.testing.assert_allclose(sol.alpha_sos, ref.alpha_sos, rtol=0,
                               atol=1e-6 * np.abs(ref.alpha_sos).max())
    _compare_vectors(sol.X, ref.X)


def test_polyene_tda_window(polyene):
    say("polyene TDA window")
    ci = polyene
    ref = O.run(ci, 7.0, 0.25)
    with _hip() as h:
        h.setup(ci, 7.0, 0.25)
        sol = h.solve_tda(None)
    assert len(sol.omega) == len(ref.omega)
    np.testing.assert_allclose(sol.omega, ref.omega, atol=E_TOL)
    assert_props(sol.props, ref.props)


def test_polyene_explicit_matrices(polyene):
    say("polyene explicit A, A+-B, 0.5B and matrix-free products")
    ci = polyene
    ax = 0.25
    it = O.intermediates(ci, ax)
    sel = O.select(ci, it, 7.0)
    A = O.build_a(ci, it, sel)
    apb, amb = O.build_apb_amb(ci, it, sel, ax)
    bh = O.build_bhalf(ci, it, sel, ax)
    n = sel["nci"]
    rng = np.random.default_rng(0)
    V = rng.standard_normal((n, 5))
    with _hip() as h:
        h.setup(ci, 7.0, ax)
        Ah = h.build_a()
        apbh, ambh = h.build_apb_amb()
        bhh = h.build_bhalf()
        mv = [h.matvec(w, V) for w in range(4)]
    np.testing.assert_allclose(Ah, A, atol=1e-12)
    np.testing.assert_allclose(apbh, apb, atol=1e-12)
    np.testing.assert_allclose(ambh, amb, atol=1e-12)
    np.testing.assert_allclose(bhh, bh, atol=1e-12)
    mats = [A, O.unpack_sym(apb, n), O.unpack_sym(amb, n), O.unpack_sym(bh, n)]
    for w in range(4):
        np.testing.assert_allclose(mv[w], mats[w] @ V, atol=1e-11, err_msg=f"operator {w}")


def test_polyene_rpa(polyene):
    say("polyene RPA k=8")
    ci = polyene
    ref = O.run(ci, 7.0, 0.25, rpa=True, nroots=8)
    with _hip() as h:
        h.setup(ci, 7.0, 0.25)
        sol = h.solve_rpa(8)
    np.testing.assert_allclose(sol.omega, ref.omega, atol=E_TOL)
    assert_props(sol.props, ref.props)
    _compare_vectors(sol.X + sol.Y, ref.X + ref.Y)


def test_polyene_lresp(polyene):
    say("polyene lresp at 4 frequencies")
    ci = polyene
    omegas = [0.0, 0.05, 0.1, 0.15]
    ref = O.run(ci, 7.0, 0.25, rpa=True, resp_omegas=omegas)
    with _hip() as h:
        h.setup(ci, 7.0, 0.25)
        res = h.lresp(omegas)
    for (w, a), (wr, ar) in zip(res, ref.alpha_resp):
        assert w == wr
        np.testing.assert_allclose(a, ar, rtol=0, atol=1e-6 * np.abs(ar).max())


def test_lresp_restarts(cluster30, monkeypatch):
    """The response solver restarted (subspace capped far below its default
    size; the initial block, the current solutions and the space's RPA
    vectors below the frequencies are kept) reaches the same alpha(w) as the
    uncapped solve, below and between poles."""
    say("30-atom cluster lresp with forced restarts")
    omegas = [0.0, 0.1, 0.15]
    with _hip() as h:
        h.setup(cluster30, 10.0, 0.25)
        ref = h.lresp(omegas)
        monkeypatch.setenv("STDA_HIP_LRESP_MAXDIM", "48")
        res = h.lresp(omegas)
    for (w, a), (wr, ar) in zip(res, ref):
        assert w == wr
        np.testing.assert_allclose(a, ar, rtol=0, atol=1e-6 * np.abs(ar).max())


def test_device_reserve_nomem():
    """A device allocation that would leave less than the runtime's reserve
    free fails with STDA_HIP_ERR_NOMEM instead of taking the device's last
    memory.  The reserve is set beyond the device's size for a fresh
    process (it is read once per process), so every new device buffer
    refuses cleanly: the call returns status 4 and the process ends normally."""
    import subprocess
    import textwrap
    say("device reserve: NOMEM in a child process")
    code = textwrap.dedent("""
        import sys
        sys.path.insert(0, sys.argv[1])
        from stda_amd import synth
        from stda_amd.prep import prepare
        from stda_amd.core import StdaHip, StdaHipError
        mol, ints = synth.cluster(12, seed=7, basis="minimal")
        ci = prepare(mol, 7.0, 0.25, ints)
        try:
            with StdaHip(0) as h:
                h.setup(ci, 7.0, 0.25)
                h.solve_tda(5)
            print("status 0")
        except StdaHipError as e:
            print("status", e.status, str(e)[:160])
    """)
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    env = dict(os.environ, STDA_HIP_DEVICE_RESERVE_MB="1e12")
    r = subprocess.run([sys.executable, "-c", code, root], env=env, capture_output=True,
                       text=True, timeout=240, cwd=root)
    say(r.stdout.strip()[-200:])
    assert r.returncode == 0, r.stderr[-2000:]
    assert "status 4" in r.stdout, r.stdout


@pytest.mark.parametrize("case", [1, 2, 3])
def test_water_beta(water, case):
    """SHG beta and beta_HRS (std2 -resp 2) against the reference's BETA_TEST files (its
    FP32 noise: 1e-6 relative) and the full tensor against the oracle (1e-8 of its largest
    component)."""
    ax, thr, alk, bej = WATER[case]
    say(f"water beta case {case}")
    ci = prepare(water, thr, ax)
    om = beta_test_omegas()
    r = O.run(ci, thr, ax, alphak=alk, betaj=bej, rpa=True, resp_omegas=[0.0], keep_matrices=True)
    ref = O.beta_shg(ci, r.iconf, r.apb, r.amb, om)
    with _hip() as h:
        h.setup(ci, thr, ax, alphak=alk, betaj=bej)
        res = h.beta_shg(om)
    exp = read_table(os.path.join(GOLDEN, "water", f"BETA_TEST{case}.txt"))
    for (w, b, hr), (wr, br, hrr), row in zip(res, ref, exp):
        assert w == wr
        np.testing.assert_allclose(b, br, rtol=0, atol=1e-8 * np.abs(br).max())
        assert abs(hr["beta_hrs"] - hrr[2]) <= 1e-8 * hrr[2]
        assert abs(hr["dr"] - hrr[3]) <= 1e-7 * abs(hrr[3])
        assert abs(hr["beta_hrs"] - row[1]) <= 1e-6 * row[1]


@pytest.mark.parametrize("which", ["polyene", "cluster30"])
def test_beta_against_oracle(which, polyene, cluster30):
    """Larger systems: the C14H16 polyene (1034 of its 1040 occupied-virtual pairs selected:
    the pairs left out are zero-filled in the dense form and must drop out of both sums)
    and a 30-atom cluster (2301 CSFs)."""
    ci, thr = (polyene, 7.0) if which == "polyene" else (cluster30, 10.0)
    om = [0.0, 0.02, 0.04]
    say(f"{which} beta at {om}")
    r = O.run(ci, thr, 0.25, rpa=True, resp_omegas=[0.0], keep_matrices=True)
    ref = O.beta_shg(ci, r.iconf, r.apb, r.amb, om)
    with _hip() as h:
        h.setup(ci, thr, 0.25)
        assert h.nci == len(r.iconf)
        res = h.beta_shg(om)
    for (w, b, hr), (wr, br, hrr) in zip(res, ref):
        np.testing.assert_allclose(b, br, rtol=0, atol=1e-7 * np.abs(br).max())
        assert abs(hr["beta_hrs"] - hrr[2]) <= 1e-7 * hrr[2]


@pytest.mark.parametrize("case", [1, 2, 3])
def test_water_2pa(water, case):
    """2PA (std2 -2PA 3) with the states srpapack hands the routine in that mode: the
    deltas against the reference's 2PA_TEST files (printed F20.3, its FP32 noise) and sigma
    against the oracle (1e-8 of its largest element)."""
    ax, thr, alk, bej = WATER[case]
    say(f"water 2PA case {case}")
    ci = prepare(water, thr, ax)
    r = O.run(ci, thr, ax, alphak=alk, betaj=bej, rpa=True, keep_matrices=True)
    eci, X, Y = O.rpa_states_tpa_mode(r.apb, r.amb, 3)
    ref = O.tpa(ci, r.iconf, r.apb, r.amb, eci, X, Y, 3)
    with _hip() as h:
        h.setup(ci, thr, ax, alphak=alk, betaj=bej)
        res = h.tpa(eci, X, Y)
    exp = read_table(os.path.join(GOLDEN, "water", f"2PA_TEST{case}.txt"))
    for (e, sg, d), (er, sgr, dr), row in zip(res, ref, exp):
        np.testing.assert_allclose(sg, sgr, rtol=0, atol=1e-8 * np.abs(sgr).max())
        got = [d["par"], d["perp"], d["circ"], d["rho"]]
        np.testing.assert_allclose(got, dr, rtol=1e-7, atol=1e-12)
        for v, rv in zip(got, row[2:6]):
            assert abs(v - rv) <= 0.5e-3 + 1e-4 + 2e-4 * abs(rv)


def test_polyene_2pa_of_solved_states(polyene):
    """2PA of the polyene's four lowest states from the GPU RPA solve, against the oracle
    on the same states (sigma within 1e-7 of its largest element)."""
    say("polyene 2PA")
    ci = polyene
    r = O.run(ci, 7.0, 0.25, rpa=True, nroots=4, keep_matrices=True)
    with _hip() as h:
        h.setup(ci, 7.0, 0.25)
        sol = h.solve_rpa(4)
        res = h.tpa(sol.omega, sol.X, sol.Y)
    ref = O.tpa(ci, r.iconf, r.apb, r.amb, sol.omega, sol.X, sol.Y, 4)
    for (e, sg, d), (er, sgr, dr) in zip(res, ref):
        np.testing.assert_allclose(sg, sgr, rtol=0, atol=1e-7 * np.abs(sgr).max())


@pytest.mark.parametrize("variant", ["dual", "dual_all_atoms", "pt2_off", "dual_pt2_off"])
def test_selection_variants(variant, polyene, cluster30):
    """-dual thresholds and PT2 off (stda-rw_dual.f:658-853): the same CSF list as the
    oracle (identical iconf) and the same lowest roots (1e-6 eV).  ed to 1e-10 Hartree plus
    1e-9 relative: with two thresholds, secondaries of the non-chromophore MOs can lie below
    chromophore primaries, so PT2 denominators (de_s - de_p) pass near zero and the sums
    carry large terms whose rounding depends on the summation order."""
    say(f"selection variant {variant}")
    cases = [(polyene, 7.0, [1, 2], 5.0), (cluster30, 10.0, [1, 2, 3, 4], 6.0)]
    for ci, thr, atoms, thr2 in cases:
        if variant == "dual_all_atoms":
            atoms = list(range(1, ci.ncent + 1))
        dual = None if variant == "pt2_off" else (thr2, atoms)
        off = variant.endswith("pt2_off")
        ref = O.run(ci, thr, 0.25, nroots=5, dual=dual, pt2_off=off)
        with _hip() as h:
            h.set_selection(dual=dual, pt2_off=off)
            h.setup(ci, thr, 0.25)
            iconf, ed = h.csf()
            sol = h.solve_tda(5)
        np.testing.assert_array_equal(iconf, ref.iconf)
        np.testing.assert_allclose(ed, ref.ed, rtol=1e-9, atol=1e-10)
        np.testing.assert_allclose(sol.omega, ref.omega, rtol=0, atol=E_TOL)


@pytest.mark.parametrize("velcorr", [False, True])
def test_axis_resolved_properties(polyene, velcorr):
    """-aniso: per-axis fL, fV, RL, RV (Rv(corr) with velcorr) against the oracle, and the
    reference's stated sums f = (f_x + f_y + f_z)/3, R = R_x + R_y + R_z against the solve's
    own totals."""
    say(f"axis-resolved properties velcorr={velcorr}")
    ci = polyene
    ref = O.run(ci, 7.0, 0.25, nroots=6, velcorr=velcorr, keep_matrices=True)
    with _hip() as h:
        h.setup(ci, 7.0, 0.25)
        sol = h.solve_tda(6, velcorr=velcorr)
        pxyz = h.props_xyz(sol.omega, sol.X, velcorr=velcorr)
    np.testing.assert_allclose(pxyz[:, :, :2].sum(1) / 3.0, sol.props[:, :2], rtol=1e-10,
                               atol=1e-14)
    np.testing.assert_allclose(pxyz[:, :, 2:].sum(1), sol.props[:, 2:], rtol=1e-10, atol=1e-12)
    X = sol.X * np.sign((sol.X * ref.X).sum(0))[None, :]
    oref = O.properties_xyz(ci, ref.iconf, sol.omega, X, 2.0,
                            ref.bhalf if velcorr else None)
    np.testing.assert_allclose(pxyz[:, :, :2], oref[:, :, :2], atol=F_TOL, rtol=0)
    np.testing.assert_allclose(pxyz[:, :, 2:], oref[:, :, 2:], atol=F_TOL, rtol=R_RTOL)
    from stda_amd import spectrum as S
    tx, ty, tz = S.tda_dat_xyz(sol.omega, pxyz, ci.xmolw, velcorr=velcorr, thr=0.3)
    assert tx.splitlines()[1] == (" VELO" if velcorr else " UV")


@pytest.mark.parametrize("which", ["water", "polyene", "cluster30"])
def test_device_transforms(which, water):
    """stda_hip_transform (moments and Loewdin on the device) against the host numpy
    preparation: Loewdin orbitals and the nine packed moment arrays to 1e-12 (of the
    largest element); for water the GPU-prepared inputs also reproduce EXCI_TEST1."""
    from stda_amd import synth
    from stda_amd.aoints import one_electron
    say(f"device transforms {which}")
    if which == "water":
        mol, thr, ax = water, 20.0, 1.0
        ints = one_electron(mol)
    elif which == "polyene":
        (mol, ints), thr, ax = synth.polyene(), 7.0, 0.25
    else:
        (mol, ints), thr, ax = synth.cluster(30, seed=3, basis="minimal"), 10.0, 0.25
    host = prepare(mol, thr, ax, ints)
    with _hip() as h:
        dev = prepare(mol, thr, ax, ints, device=h)
    np.testing.assert_allclose(dev.clow, host.clow, rtol=0, atol=1e-12)
    np.testing.assert_allclose(dev.moments, host.moments, rtol=0,
                               atol=1e-12 * np.abs(host.moments).max())
    if which == "water":
        r = O.run(dev, thr, ax)
        exp = read_datxy(os.path.join(GOLDEN, "water", "EXCI_TEST1.dat"))
        for k, row in enumerate(exp):
            assert abs(r.omega[k] * O.HARTREE_EV_PRINT - row[1]) <= 0.5e-4 + 1e-6


@pytest.mark.parametrize("which", ["water", "polyene", "cluster30"])
def test_device_ao_integrals(which, water):
    """stda_hip_ao_ints against the host Obara-Saika integrals (aoints.one_electron): S, the
    dipole, angular-momentum and velocity matrices to 1e-13 of their largest element; and
    the whole preparation on the device (integrals, transforms, Loewdin) against the host's."""
    from stda_amd import synth
    from stda_amd.aoints import one_electron
    say(f"device AO integrals {which}")
    if which == "water":
        mol, thr, ax = water, 20.0, 1.0
    elif which == "polyene":
        mol, thr, ax = synth.polyene()[0], 7.0, 0.25
    else:
        mol, thr, ax = synth.cluster(30, seed=3, basis="minimal")[0], 10.0, 0.25
    ref = one_electron(mol)
    with _hip() as h:
        got = h.ao_ints(mol)
        dev = prepare(mol, thr, ax, device=h)
    for key in ("S", "R", "M", "V"):
        np.testing.assert_allclose(got[key], ref[key], rtol=0,
                                   atol=1e-13 * max(1.0, np.abs(ref[key]).max()), err_msg=key)
    host = prepare(mol, thr, ax, ref)
    np.testing.assert_allclose(dev.clow, host.clow, rtol=0, atol=1e-12)
    np.testing.assert_allclose(dev.moments, host.moments, rtol=0,
                               atol=1e-12 * np.abs(host.moments).max())


def test_device_transform_negative_overlap(water):
    """A negative eigenvalue of S stops as the reference does ('must stop in S^1/2')."""
    from stda_amd.aoints import one_electron
    from stda_amd.prep import StdaInputError
    ints = dict(one_electron(water))
    e, U = np.linalg.eigh(ints["S"])
    ints["S"] = ints["S"] - 2.0 * e[0] * np.outer(U[:, 0], U[:, 0])
    with _hip() as h:
        with pytest.raises(StdaInputError, match="negative eigenvalue"):
            prepare(water, 20.0, 1.0, ints, device=h)


def test_stda_dropin_branches(water):
    """The stda() drop-in's -resp (beta), -2PA and -dual/-rw 1 branches return what the
    entry points return."""
    from stda_amd.core import stda
    say("stda() drop-in branches")
    ci = prepare(water, 20.0, 1.0)
    om = beta_test_omegas()
    betas, csf = stda(ci, 20.0, 1e-4, 1.0, beta_omegas=om)
    ref = O.run(ci, 20.0, 1.0, rpa=True, resp_omegas=[0.0], keep_matrices=True)
    for (w, b, hr), (wr, br, hrr) in zip(betas, O.beta_shg(ci, ref.iconf, ref.apb, ref.amb, om)):
        assert abs(hr["beta_hrs"] - hrr[2]) <= 1e-8 * hrr[2]
    tp, _ = stda(ci, 20.0, 1e-4, 1.0, rpa=True, tpa_states=2)
    assert len(tp) == 2 and all(d["par"] >= 0.0 for _, _, d in tp)
    sol, (iconf, ed) = stda(ci, 20.0, 1e-4, 1.0, nroots=3, pt2_off=True)
    r2 = O.run(ci, 20.0, 1.0, nroots=3, pt2_off=True)
    np.testing.assert_array_equal(iconf, r2.iconf)
    np.testing.assert_allclose(sol.omega, r2.omega, rtol=0, atol=E_TOL)


def test_sos_rotation_mirror_image():
    """The SOS optical rotation table (sosor.f) from GPU solves on device-prepared inputs:
    equal to the oracle's, and the negative for the mirror image (x -> -x)."""
    import dataclasses
    from stda_amd import spectrum as S
    from stda_amd import synth
    say("SOS rotation, mirror image")
    mol, _ = synth.cluster(12, seed=5, basis="minimal")
    sign = np.where(mol.ao_lmn[:, 0] % 2 == 1, -1.0, 1.0)
    xyz = mol.xyz.copy()
    xyz[:, 0] *= -1.0
    mirror = dataclasses.replace(mol, xyz=xyz, C=mol.C * sign[:, None])
    vals = []
    with _hip() as h:
        for m in (mol, mirror):
            ci = prepare(m, 7.0, 0.25, device=h)
            h.setup(ci, 7.0, 0.25)
            sol = h.solve_tda(10)
            got = np.array([x[2] for x in S.sos_rotation(sol.omega, sol.props[:, 3], ci.xmolw)])
            ref = O.run(ci, 7.0, 0.25, nroots=10)
            want = np.array([x[2] for x in S.sos_rotation(ref.omega, ref.props[:, 3], ci.xmolw)])
            np.testing.assert_allclose(got, want, rtol=1e-5, atol=1e-4)
            vals.append(got)
    np.testing.assert_allclose(vals[0], -vals[1], rtol=1e-5, atol=1e-4)


def test_cluster_selection_and_roots(cluster30):
    say("30-atom cluster TDA k=12")
    ci = cluster30
    ref = O.run(ci, 10.0, 0.25, nroots=12)
    with _hip() as h:
        h.setup(ci, 10.0, 0.25)
        _compare_selection(h, ref)
        sol = h.solve_tda(12)
    np.testing.assert_allclose(sol.omega, ref.omega, atol=E_TOL)
    assert_props(sol.props, ref.props)


@pytest.fixture(scope="module")
def cluster60():
    from stda_amd import synth
    mol, ints = synth.cluster(60, seed=1004, basis="minimal")
    return prepare(mol, 7.0, 0.25, ints)


@pytest.mark.parametrize("store_max", [None, "0"])
def test_pt2_pass2_forms(cluster60, monkeypatch, store_max):
    """PT2 pass 2 (the primaries' lowering by the rejected secondaries, 640
    of 5686 here) from the stored pass-1 terms (default) and by recomputing
    the couplings of the rejected rows (the form used when the matrix of
    pass-1 terms exceeds its memory limit, forced with a limit of 0 bytes):
    both give the oracle's selection and PT2-lowered diagonal."""
    if store_max is not None:
        monkeypatch.setenv("STDA_HIP_PT2_STORE_MAX", store_max)
    ci = cluster60
    r = O.select(ci, O.intermediates(ci, 0.25), 7.0)
    assert r["nexpt"] - r["npt"] > 100
    ref = type("Sel", (), {"nex": r["nex"], "nci": r["nci"], "iconf": r["iconf"], "ed": r["ed"]})
    with _hip() as h:
        h.setup(ci, 7.0, 0.25)
        assert h.kernel_stats("pt2_cols")["launches"] >= 1
        _compare_selection(h, ref)


def test_mixed_precision_selection_is_fp64(cluster60):
    """Mixed precision runs PT2 pass 1 on FP32 operands and evaluates again in
    FP64 every row tile with a secondary near the threshold: the CSF list is
    the FP64 one, and the PT2-lowered diagonal (pass 2 from the FP32-derived
    terms) agrees to 1e-8 Hartree."""
    ci = cluster60
    with _hip() as h:
        h.setup(ci, 7.0, 0.25)
        ic64, ed64 = h.csf()
        h.set_precision("mixed")
        h.setup(ci, 7.0, 0.25)
        assert h.kernel_stats("pt2_rows_f32")["launches"] == 1
        ic32, ed32 = h.csf()
    np.testing.assert_array_equal(ic32, ic64)
    np.testing.assert_allclose(ed32, ed64, atol=1e-8, rtol=0)


def test_many_roots_against_dense():
    """50 roots: a Davidson subspace of up to 300 vectors, beyond the shapes
    of the hand-written tall-skinny kernels (split-K Gram products, V C with
    more than 32 columns, Cholesky QR of blocks of up to 50 directions).
    Reference: the lowest eigenvalues of the explicit A of the same setup
    (A itself is held to the oracle by test_polyene_explicit_matrices)."""
    from stda_amd import synth
    say("40-atom cluster TDA k=50 against the dense eigenvalues")
    mol, ints = synth.cluster(40, seed=41, basis="minimal")
    ci = prepare(mol, 7.0, 0.25, ints)
    with _hip() as h:
        h.setup(ci, 7.0, 0.25)
        say(f"  nci={h.nci}")
        A = h.build_a()
        sol = h.solve_tda(50, tol=1e-8)
    w = np.linalg.eigvalsh(A)
    k = len(sol.omega)
    assert k == min(50, A.shape[0])
    np.testing.assert_allclose(sol.omega, w[:k], atol=E_TOL)


@pytest.mark.parametrize("system", ["polyene", "cluster30"])
def test_kshift_xtb_parameters(system):
    """xTB mode (std2 -xtb, main.f:198-219): a_x 0.5, alpha 2, beta 4 and the K(ia,ia)
    shift (shftmax 0.5 eV, width 0.1 eV, steepness 4; stda.f:1750-1768) in the screening
    and PT2 selection: identical CSF list, ed and roots as the oracle."""
    from stda_amd import synth
    say(f"{system} xTB parameters with the K shift")
    if system == "polyene":
        mol, ints = synth.polyene()
        thr = 7.0
    else:
        mol, ints = synth.cluster(30, seed=3, basis="minimal")
        thr = 10.0
    ci = prepare(mol, thr, 0.5, ints)
    ks = (0.5, 0.1, 4.0)
    ref = O.run(ci, thr, 0.5, alphak=2.0, betaj=4.0, nroots=8, kshift=ks)
    plain = O.run(ci, thr, 0.5, alphak=2.0, betaj=4.0, nroots=8)
    assert ref.nex != plain.nex or not np.array_equal(ref.iconf, plain.iconf)
    with _hip() as h:
        h.setup(ci, thr, 0.5, alphak=2.0, betaj=4.0, kshift=ks)
        _compare_selection(h, ref)
        sol = h.solve_tda(8)
    np.testing.assert_allclose(sol.omega, ref.omega, atol=E_TOL)
    assert_props(sol.props, ref.props)


def test_triplet_and_ax0(polyene):
    say("polyene triplet TDA and ax=0 RPA")
    ci = polyene
    ref = O.run(ci, 7.0, 0.25, nroots=6, triplet=True)
    with _hip() as h:
        h.setup(ci, 7.0, 0.25, triplet=True)
        _compare_selection(h, ref)
        sol = h.solve_tda(6)
    np.testing.assert_allclose(sol.omega, ref.omega, atol=E_TOL)
    ref0 = O.run(ci, 7.0, 0.0, rpa=True, nroots=6)
    with _hip() as h:
        h.setup(ci, 7.0, 0.0)
        sol0 = h.solve_rpa(6)
    np.testing.assert_allclose(sol0.omega, ref0.omega, atol=E_TOL)
    assert_props(sol0.props, ref0.props)


def test_no_csf_error(polyene):
    from stda_amd.core import StdaHipError
    with _hip() as h:
        with pytest.raises(StdaHipError) as e:
            h.setup(polyene, 0.01, 0.25)
    assert e.value.status == 2


# ---------------------------------------------------------------- mixed precision
MIXED_E_TOL = 1e-3 / EV      # every root within 1 meV of the FP64 oracle
MIXED_F_TOL = 1e-4           # oscillator strengths within 1e-4


def test_mixed_precision_tda_rpa(polyene, cluster30):
    say("mixed precision TDA / RPA / response")
    for ci, thr, k in ((polyene, 7.0, 10), (cluster30, 10.0, 12)):
        ref = O.run(ci, thr, 0.25, nroots=k)
        with _hip() as h:
            h.setup(ci, thr, 0.25)
            h.set_precision("mixed")
            sol = h.solve_tda(k)
        err = np.abs(sol.omega - ref.omega)
        say(f"  TDA nci={ref.nci}: max |dE| = {err.max() * EV:.3e} eV, "
            f"max |df| = {np.abs(sol.props[:, 0] - ref.props[:, 0]).max():.3e}")
        assert err.max() < MIXED_E_TOL
        np.testing.assert_allclose(sol.props[:, :2], ref.props[:, :2], atol=MIXED_F_TOL, rtol=0)
    ref = O.run(polyene, 7.0, 0.25, rpa=True, nroots=8)
    with _hip() as h:
        h.setup(polyene, 7.0, 0.25)
        h.set_precision("mixed")
        sol = h.solve_rpa(8)
        res = h.lresp([0.0, 0.05])
    assert np.abs(sol.omega - ref.omega).max() < MIXED_E_TOL
    np.testing.assert_allclose(sol.props[:, :2], ref.props[:, :2], atol=MIXED_F_TOL, rtol=0)
    refa = O.run(polyene, 7.0, 0.25, rpa=True, resp_omegas=[0.0, 0.05])
    for (w, a), (wr, ar) in zip(res, refa.alpha_resp):
        np.testing.assert_allclose(a, ar, rtol=0, atol=1e-4 * np.abs(ar).max())
