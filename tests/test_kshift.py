"""The K(ia,ia)-dependent diagonal shift of xTB mode (kshift_to_ediag, stda.f:1750-1768;
applied in the screening loop, stda.f:610-614; parameters of std2 -xtb, main.f:198-219:
a_x 0.5, alpha 2, beta 4, shftmax 0.5 eV, shftwidth 0.1 eV, shftsteep 4).

The oracle's screening with the shift is checked against an independent numpy evaluation
of the screening energies over the whole occupied x virtual grid, against the shift's
limits (zero maximum, large K) and against the reference's selection rule and sort."""
import numpy as np
import pytest

import oracle as O
from stda_amd.prep import prepare

EV_AU = 0.036749320          # the conversion kshift_to_ediag uses (stda.f:1763-1764)
XTB = dict(ax=0.5, alphak=2.0, betaj=4.0)
KS = (0.5, 0.1, 4.0)


@pytest.fixture(scope="module")
def poly_xtb():
    from stda_amd import synth
    mol, ints = synth.polyene()
    ci = prepare(mol, 7.0, XTB["ax"], ints)
    inter = O.intermediates(ci, XTB["ax"], XTB["alphak"], XTB["betaj"])
    return ci, inter


def _grid_energies(ci, inter, ak=2.0):
    """de(io, a) = eps_a - eps_i - (ii|aa)_J + ak (ia|ia)_K and K(ia,ia), pair-major."""
    no, nv = ci.no, ci.nv
    eps = np.asarray(ci.epsi)
    K = np.einsum("ap,ap->p", inter["qia"], inter["pia"]).reshape(no, nv)
    gap = eps[no:][None, :] - eps[:no][:, None]
    J = inter["jdiag"].T                      # (no, nv)
    return gap - J + ak * K, K


def test_shift_values_and_selection(poly_xtb):
    ci, inter = poly_xtb
    thr = 7.0 / 27.211385050
    de0, K = _grid_energies(ci, inter)
    shift = KS[0] * EV_AU / (1.0 + (K / (KS[1] * EV_AU)) ** KS[2])
    de = de0 + shift
    sel = O.select(ci, inter, 7.0, kshift=KS, pt2_off=True)
    io, iv = sel["iconf"][:, 0] - 1, sel["iconf"][:, 1] - 1 - ci.no
    # the primaries are exactly the shifted energies <= thr, in ascending energy
    expect = set(zip(*np.nonzero(de <= thr)))
    assert set(zip(io.tolist(), iv.tolist())) == expect
    np.testing.assert_allclose(sel["ed"], de[io, iv], rtol=0, atol=1e-14)
    assert np.all(np.diff(sel["ed"]) >= 0)
    # the shift of every primary lies in (0, shftmax]
    s = sel["ed"] - de0[io, iv]
    assert np.all(s > 0) and np.all(s <= KS[0] * EV_AU * (1 + 1e-12))
    # the shift moves CSFs with small K across the threshold: the selection differs
    plain = O.select(ci, inter, 7.0, pt2_off=True)
    assert sel["nex"] < plain["nex"]
    # secondaries (PT2 window) carry the same shift
    full = O.select(ci, inter, 7.0, kshift=KS)
    ko, kv = full["kconf"][:, 0] - 1, full["kconf"][:, 1] - 1 - ci.no
    np.testing.assert_allclose(full["edpt"], de[ko, kv], rtol=0, atol=1e-14)


def test_zero_maximum_is_no_shift(poly_xtb):
    ci, inter = poly_xtb
    a = O.select(ci, inter, 7.0, kshift=(0.0, 0.1, 4.0))
    b = O.select(ci, inter, 7.0)
    assert np.array_equal(a["iconf"], b["iconf"])
    np.testing.assert_array_equal(a["ed"], b["ed"])


def test_large_k_unshifted(poly_xtb):
    """With a vanishing width every CSF has K >> width: the shift is below 1e-12 Eh."""
    ci, inter = poly_xtb
    a = O.select(ci, inter, 7.0, kshift=(0.5, 1e-8, 4.0), pt2_off=True)
    b = O.select(ci, inter, 7.0, pt2_off=True)
    assert np.array_equal(a["iconf"], b["iconf"])
    np.testing.assert_allclose(a["ed"], b["ed"], rtol=0, atol=1e-12)
