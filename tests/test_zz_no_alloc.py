"""stda_hip_gspec_noalloc (no device allocation, fixed staging area per GPU) against
stda_hip_gspec through the C ABI: the same kernel on the same inputs gives identical curves,
on two calls in a row (the second smaller, so stale contents of the shared area would show),
and a call beyond the fixed area is refused with STDA_HIP_ERR_TOO_LARGE."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

UV, NM, CM, LOG = 1, 2, 4, 8


def _lines(rng, nspec, maxlines):
    counts = rng.integers(0, maxlines + 1, size=nspec)
    off = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
    nl = int(off[-1])
    x = rng.uniform(1.0, 12.0, nl)
    x[rng.uniform(size=nl) < 0.05] = 0.0          # lines g_spec skips (|x| <= 1e-6)
    y = rng.uniform(-0.5, 2.0, nl)
    w = rng.uniform(0.1, 0.5, nl)
    emin = rng.uniform(0.5, 2.0, nspec)
    step = rng.uniform(0.002, 0.01, nspec)
    return off, x, y, w, emin, step


def _call(L, h, fn, args, nspec, npts, flags):
    off, x, y, w, emin, step = args
    e = np.full(nspec * npts, -12345.0)      # sentinel: every slot must be written
    r = np.full(nspec * npts, -12345.0)
    p = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    st = getattr(L, fn)(h, nspec, p(off), p(x), p(y), p(w), p(emin), p(step), npts, flags, 1.0,
                        p(e), p(r))
    return st, e, r


def test_noalloc_equals_gspec_twice():
    from stda_amd.core import StdaHip
    rng = np.random.default_rng(7)
    with StdaHip(0) as h:
        L = h._L
        for nspec, npts, maxl, flags in ((96, 2000, 60, UV | NM), (7, 500, 9, LOG)):
            print(f"gspec_noalloc: {nspec} spectra x {npts} points, flags {flags}", flush=True)
            args = _lines(rng, nspec, maxl)
            st0, e0, r0 = _call(L, h._h, "stda_hip_gspec", args, nspec, npts, flags)
            st1, e1, r1 = _call(L, h._h, "stda_hip_gspec_noalloc", args, nspec, npts, flags)
            assert st0 == 0 and st1 == 0, (st0, st1)
            np.testing.assert_array_equal(e1, e0)
            np.testing.assert_array_equal(r1, r0)
            # every grid point written (NaN where LOG meets a negative CD-type sum, as in
            # the reference program)
            assert not np.any(e1 == -12345.0) and not np.any(r1 == -12345.0)


def test_noalloc_rejects_beyond_the_area():
    from stda_amd.core import StdaHip
    rng = np.random.default_rng(8)
    nspec, npts = 1200, 2000        # 2 x 1200 x 2000 doubles of output alone: 38.4 MB
    args = _lines(rng, nspec, 2)
    with StdaHip(0) as h:
        st, _, _ = _call(h._L, h._h, "stda_hip_gspec_noalloc", args, nspec, npts, UV)
        assert st == 9                           # STDA_HIP_ERR_TOO_LARGE
        assert b"fixed area" in h._L.stda_hip_last_error(h._h)
