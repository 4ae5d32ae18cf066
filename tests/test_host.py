"""Host-side logic without a GPU: synthetic inputs, molden round trip, the
boundary preparation, and the serving scheduler's bookkeeping."""
import os

import numpy as np

from stda_amd import synth
from stda_amd.molden import read_molden, write_molden
from stda_amd.prep import prepare
from stda_amd.serve import Job, footprint_bytes, poisson_arrivals


def test_mixed_sizes_deterministic_and_bounded():
    a = synth.mixed_sizes(512, 10, 200, 1000)
    b = synth.mixed_sizes(512, 10, 200, 1000)
    np.testing.assert_array_equal(a, b)
    assert a.min() >= 10 and a.max() <= 200
    # log-uniform: about half of the molecules below the geometric mean
    gm = np.sqrt(10 * 200)
    assert 0.4 < np.mean(a < gm) < 0.6


def test_mixed_batch_parallel_equals_serial():
    s = synth.mixed_batch(4, lo=10, hi=20, base_seed=3)
    p = synth.mixed_batch(4, lo=10, hi=20, base_seed=3, workers=2)
    for x, y in zip(s, p):
        assert (x.ncent, x.nao, x.no, x.nv) == (y.ncent, y.nao, y.no, y.nv)
        np.testing.assert_array_equal(x.clow, y.clow)
        np.testing.assert_array_equal(x.moments, y.moments)


def test_loewdin_orbitals_orthonormal():
    mol, ints = synth.cluster(12, seed=4, basis="minimal")
    ci = prepare(mol, 7.0, 0.25, ints)
    C = np.asarray(ci.clow)
    np.testing.assert_allclose(C.T @ C, np.eye(C.shape[1]), atol=1e-10)
    # AO -> atom map is 1-based and covers every atom
    assert ci.iaoat.min() == 1 and ci.iaoat.max() == ci.ncent


def test_molden_round_trip(tmp_path):
    mol, ints = synth.polyene()
    path = os.path.join(tmp_path, "p.molden")
    write_molden(path, mol)
    m2 = read_molden(path)
    assert m2.nao == mol.nao
    np.testing.assert_allclose(m2.eps, mol.eps, rtol=0, atol=1e-12)
    np.testing.assert_allclose(np.abs(m2.C), np.abs(mol.C), rtol=0, atol=1e-10)
    ci1 = prepare(mol, 7.0, 0.25, ints)
    ci2 = prepare(m2, 7.0, 0.25)
    np.testing.assert_allclose(np.abs(ci1.clow), np.abs(ci2.clow), atol=1e-9)


def test_footprint_grows_with_size():
    small = synth.mixed_batch(1, lo=10, hi=10, base_seed=5)[0]
    big = synth.mixed_batch(1, lo=60, hi=60, base_seed=5)[0]
    fs = footprint_bytes(Job(0, "spectrum", small))
    fb = footprint_bytes(Job(1, "spectrum", big))
    fp = footprint_bytes(Job(2, "polar", big, omegas=(0.0, 0.1)))
    assert 0 < fs < fb and fp > 0


def test_poisson_arrivals():
    t = poisson_arrivals(20000, 500.0, seed=1)
    assert np.all(np.diff(t) > 0)
    rate = len(t) / t[-1]
    assert abs(rate - 500.0) / 500.0 < 0.03


def test_primary_order_is_the_reference_selection_sort(tmp_path):
    """stda_amd/csrc/selsort.h reproduces the reference's selection sort of the primary CSFs
    (stda.f:651-668: first strict minimum of the rest, then a swap) on keys with many ties,
    where it differs from a stable sort."""
    import ctypes
    import shutil
    import subprocess
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    src = tmp_path / "w.cpp"
    src.write_text('#include "selsort.h"\n'
                   'extern "C" void order(int n, const double *k, int *out) {\n'
                   '  std::vector<double> key(k, k + n); std::vector<int32_t> it(n);\n'
                   '  for (int i = 0; i < n; ++i) it[i] = i;\n'
                   '  stda::reference_selection_order(key, it);\n'
                   '  for (int i = 0; i < n; ++i) out[i] = it[i]; }\n')
    so = tmp_path / "w.so"
    csrc = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "stda_amd",
                        "csrc")
    subprocess.run(["g++", "-O1", "-std=c++17", "-shared", "-fPIC", "-I", csrc, str(src), "-o",
                    str(so)], check=True)
    lib = ctypes.CDLL(str(so))

    def reference(keys):
        ed = list(keys)
        items = list(range(len(ed)))
        for i in range(len(ed) - 1):
            k, pp = i, ed[i]
            for j in range(i + 1, len(ed)):
                if ed[j] >= pp:
                    continue
                k, pp = j, ed[j]
            if k == i:
                continue
            ed[k], ed[i] = ed[i], pp
            items[i], items[k] = items[k], items[i]
        return items

    rng = np.random.default_rng(3)
    differs = 0
    for n in [1, 2, 3, 7, 64, 257, 1000]:
        for _ in range(5):
            keys = rng.integers(0, max(2, n // 8), n).astype(np.float64)
            out = np.zeros(n, dtype=np.int32)
            lib.order(n, keys.ctypes.data_as(ctypes.c_void_p), out.ctypes.data_as(ctypes.c_void_p))
            ref = reference(keys)
            assert out.tolist() == ref
            differs += ref != np.argsort(keys, kind="stable").tolist()
    assert differs > 0
