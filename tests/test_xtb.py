"""The xTB input path: the wfn.xtb reader (readxtb.f:38-200) and the -xtb parameters
(main.f:198-219).  The reference holds no wfn.xtb fixture; the reader is pinned by a file
assembled byte by byte here from the record sequence of readxtb.f (independently of the
writer), by a round trip through the writer, and by the core inputs of a synthetic molecule
read back from wfn.xtb being identical to those of the molecule itself."""
import struct

import numpy as np
import pytest

import oracle as O
from stda_amd import synth
from stda_amd.prep import prepare
from stda_amd.xtb import XTB_PARAMETERS, XtbFormatError, read_wfn_xtb, write_wfn_xtb


def _rec(payload):
    return struct.pack("<i", len(payload)) + payload + struct.pack("<i", len(payload))


def test_hand_assembled_file(tmp_path):
    # H at the origin with one s AO of two primitives; O at (0, 0, 1.8) Bohr with an s AO and
    # px, py, pz AOs; the p primitives are listed before the O s primitive (ipao decides)
    ipty = [1, 1, 2, 3, 4, 1]
    ipat = [1, 1, 2, 2, 2, 2]
    ipao = [1, 1, 3, 4, 5, 2]
    exip = [3.4, 0.62, 1.1, 1.1, 1.1, 7.6]
    cxip = [0.28, 0.27, 1.3, 1.31, 1.32, 2.9]
    nmo, nbf = 5, 5
    C = np.arange(1.0, 26.0).reshape(nbf, nmo, order="F") / 10.0
    occ = [2.0, 2.0, 0.0, 0.0, 0.0]
    eps = [-0.9, -0.4, 0.1, 0.3, 0.5]
    b = _rec(struct.pack("<i", 1)) + _rec(struct.pack("<4i", 2, nbf, nmo, 6))
    b += _rec(b"h ") + _rec(b"o ")
    for xyz, z in (((0.0, 0.0, 0.0), 1), ((0.0, 0.0, 1.8), 8)):
        for v in xyz:
            b += _rec(struct.pack("<d", v))
        b += _rec(struct.pack("<i", z))
    for arr in (ipty, ipat, ipao):
        for v in arr:
            b += _rec(struct.pack("<i", v))
    for arr in (exip, cxip, occ, eps):
        b += _rec(struct.pack(f"<{len(arr)}d", *arr))
    b += _rec(C.ravel(order="F").tobytes())
    path = tmp_path / "wfn.xtb"
    path.write_bytes(b)
    mol = read_wfn_xtb(str(path))
    assert list(mol.Z) == [1, 8]
    np.testing.assert_array_equal(mol.xyz, [[0, 0, 0], [0, 0, 1.8]])
    assert list(mol.ao_atom) == [0, 1, 1, 1, 1]
    assert [tuple(x) for x in mol.ao_lmn] == [(0, 0, 0), (0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1)]
    assert list(mol.ao_prim_ptr) == [0, 2, 3, 4, 5, 6]
    np.testing.assert_array_equal(mol.prim_alpha, [3.4, 0.62, 7.6, 1.1, 1.1, 1.1])
    np.testing.assert_array_equal(mol.prim_coef, [0.28, 0.27, 2.9, 1.3, 1.31, 1.32])
    np.testing.assert_array_equal(mol.C, C)
    np.testing.assert_array_equal(mol.eps, eps)
    np.testing.assert_array_equal(mol.occ, occ)


def test_d_and_f_types_and_errors(tmp_path):
    from stda_amd.xtb import IPTY_LMN
    # the d and f tables of the reference's integral code (intpack.f90 lx/my/nz, minus one)
    assert IPTY_LMN[4:10] == [(2, 0, 0), (0, 2, 0), (0, 0, 2), (1, 1, 0), (1, 0, 1), (0, 1, 1)]
    assert IPTY_LMN[19] == (1, 1, 1) and all(sum(t) == 3 for t in IPTY_LMN[10:])
    p = tmp_path / "wfn.xtb"
    p.write_bytes(_rec(struct.pack("<i", 2)))
    with pytest.raises(XtbFormatError, match="UKS"):
        read_wfn_xtb(str(p))
    p.write_bytes(_rec(struct.pack("<i", 1))[:-1])
    with pytest.raises(XtbFormatError, match="marker"):
        read_wfn_xtb(str(p))
    with pytest.raises(XtbFormatError, match="not found"):
        read_wfn_xtb(str(tmp_path / "missing"))


def test_round_trip_and_core_inputs(tmp_path):
    mol, ints = synth.cluster(14, seed=7, basis="split")
    path = str(tmp_path / "wfn.xtb")
    write_wfn_xtb(path, mol)
    back = read_wfn_xtb(path)
    for f in ("Z", "xyz", "ao_atom", "ao_lmn", "ao_prim_ptr", "prim_alpha", "prim_coef", "C",
              "eps", "occ"):
        np.testing.assert_array_equal(getattr(back, f), getattr(mol, f), err_msg=f)
    ax = XTB_PARAMETERS["ax"]
    a = prepare(mol, 10.0, ax, ints)
    b = prepare(back, 10.0, ax)
    for f in ("clow", "epsi", "iaoat", "xyz", "moments"):
        np.testing.assert_allclose(getattr(b, f), getattr(a, f), rtol=0, atol=1e-12, err_msg=f)
    # the -xtb run of the oracle on the read molecule: the K shift is active
    kw = dict(alphak=XTB_PARAMETERS["alphak"], betaj=XTB_PARAMETERS["betaj"])
    r1 = O.run(b, 10.0, ax, nroots=5, kshift=XTB_PARAMETERS["kshift"], **kw)
    r0 = O.run(a, 10.0, ax, nroots=5, kshift=XTB_PARAMETERS["kshift"], **kw)
    np.testing.assert_allclose(r1.omega, r0.omega, atol=1e-12)
