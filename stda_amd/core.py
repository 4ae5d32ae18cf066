"""Python host layer over the C ABI of ``libstda_hip.so``.

:class:`StdaHip` wraps one device context.  :func:`stda` mirrors the
reference driver routine ``stda`` (stda.f:53-1471): from the molecule's
boundary inputs (see :mod:`stda_amd.prep`) through CSF selection to the
spectrum, with the same parameter names and meanings (``thr`` in eV, ``thrp``,
``ax``, ``alphak``/``betaj`` with -100 meaning "default", ``fthr``), plus
``nroots`` for the root count (None reproduces the reference's energy
window).  Errors raise :class:`StdaHipError` carrying the message the
reference stops with.

There is no CPU fallback: if the library is missing the import of this
module's loader raises.
"""

from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass, field

import numpy as np

from .constants import HARTREE_EV, HARTREE_EV_PRINT
from .prep import CoreInputs

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libstda_hip.so")
# the build with 64-bit index arrays (stda_hip_int = int64_t: iaoat, iconf, -dual atoms)
LIB_PATH_ILP64 = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libstda_hip_ilp64.so")

C_API = [
    "stda_hip_create", "stda_hip_destroy", "stda_hip_error_string", "stda_hip_last_error",
    "stda_hip_setup", "stda_hip_get_csf", "stda_hip_solve_tda", "stda_hip_solve_rpa",
    "stda_hip_lresp", "stda_hip_build_a", "stda_hip_build_apb_amb", "stda_hip_build_bhalf",
    "stda_hip_matvec", "stda_hip_timings", "stda_hip_load", "stda_hip_select",
    "stda_hip_kernel_stats", "stda_hip_set_precision", "stda_hip_batch_create",
    "stda_hip_batch_destroy", "stda_hip_batch_last_error", "stda_hip_batch_solve_tda",
    "stda_hip_batch_timings", "stda_hip_batch_kernel_bytes", "stda_hip_batch_configure",
    "stda_hip_batch_phase_times", "stda_hip_batch_kernel_stats", "stda_hip_gspec",
    "stda_hip_beta_shg", "stda_hip_tpa", "stda_hip_set_selection", "stda_hip_props_xyz",
    "stda_hip_format_curve", "stda_hip_transform", "stda_hip_ao_ints",
    "stda_hip_prepare_ao", "stda_hip_index_size", "stda_hip_gspec_noalloc", "stda_hip_gemm",
    "stda_hip_batch_submit", "stda_hip_batch_result", "stda_hip_batch_stop",
    "stda_hip_batch_configure_wide", "stda_hip_batch_wide_count", "stda_hip_batch_set_precision",
]


class StdaHipError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"[{status}] {message}")
        self.status = status


class _System(ctypes.Structure):
    _fields_ = [("ncent", ctypes.c_int32), ("nao", ctypes.c_int32), ("no", ctypes.c_int32),
                ("nv", ctypes.c_int32), ("clow", ctypes.c_void_p), ("epsi", ctypes.c_void_p),
                ("iaoat", ctypes.c_void_p), ("xyz", ctypes.c_void_p),
                ("moments", ctypes.c_void_p)]


class _Params(ctypes.Structure):
    _fields_ = [("thr_ev", ctypes.c_double), ("thrp", ctypes.c_double),
                ("fthr_ev", ctypes.c_double), ("ax", ctypes.c_double),
                ("alphak", ctypes.c_double), ("betaj", ctypes.c_double),
                ("triplet", ctypes.c_int32), ("kshift", ctypes.c_int32),
                ("kshift_max_ev", ctypes.c_double), ("kshift_width_ev", ctypes.c_double),
                ("kshift_steep", ctypes.c_double)]


_libs = {}


def load_library(path: str = LIB_PATH) -> ctypes.CDLL:
    """Load ``libstda_hip.so`` (or the ILP64 build); raises if it has not been built."""
    if path in _libs:
        return _libs[path]
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is missing: build it with `python -m stda_amd.build` "
                           "(there is no CPU fallback)")
    L = ctypes.CDLL(path)
    P, I32, I64, F64 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64, ctypes.c_double
    L.stda_hip_index_size.argtypes = []
    L.stda_hip_create.argtypes = [ctypes.POINTER(P), ctypes.c_int]
    L.stda_hip_destroy.argtypes = [P]
    L.stda_hip_error_string.restype = ctypes.c_char_p
    L.stda_hip_error_string.argtypes = [ctypes.c_int]
    L.stda_hip_last_error.restype = ctypes.c_char_p
    L.stda_hip_last_error.argtypes = [P]
    L.stda_hip_setup.argtypes = [P, ctypes.POINTER(_System), ctypes.POINTER(_Params),
                                 ctypes.POINTER(I64), ctypes.POINTER(I64), ctypes.POINTER(I64)]
    L.stda_hip_get_csf.argtypes = [P, P, P]
    L.stda_hip_solve_tda.argtypes = [P, I32, I32, I32, F64, ctypes.POINTER(I32), P, P, P, P]
    L.stda_hip_solve_rpa.argtypes = [P, I32, I32, F64, ctypes.POINTER(I32), P, P, P, P, P]
    L.stda_hip_lresp.argtypes = [P, I32, P, F64, P, P]
    L.stda_hip_build_a.argtypes = [P, P]
    L.stda_hip_build_apb_amb.argtypes = [P, P, P]
    L.stda_hip_build_bhalf.argtypes = [P, P]
    L.stda_hip_matvec.argtypes = [P, I32, I32, P, P]
    L.stda_hip_gspec.argtypes = [P, I32, P, P, P, P, P, P, I32, I32, F64, P, P]
    L.stda_hip_gspec_noalloc.argtypes = [P, I32, P, P, P, P, P, P, I32, I32, F64, P, P]
    L.stda_hip_beta_shg.argtypes = [P, I32, P, F64, P, P]
    L.stda_hip_tpa.argtypes = [P, I32, P, P, P, I64, F64, P, P]
    L.stda_hip_set_selection.argtypes = [P, F64, I32, P, I32]
    L.stda_hip_props_xyz.argtypes = [P, I32, P, P, I64, I32, P]
    L.stda_hip_format_curve.argtypes = [I64, P, P, P]
    L.stda_hip_transform.argtypes = [P, I32, I32, P, P, P, I32, P, P]
    L.stda_hip_ao_ints.argtypes = [P, I32, P, P, P, P, P, P, P]
    L.stda_hip_prepare_ao.argtypes = [P, I32, P, P, P, P, P, I32, P, I32, P, P]
    L.stda_hip_timings.argtypes = [P, P]
    L.stda_hip_set_precision.argtypes = [P, I32]
    L.stda_hip_gemm.argtypes = [P, I32, I32, I32, I32, I32, I32, F64, P, I64, P, I64, F64, P, I64]
    L.stda_hip_load.argtypes = [P, ctypes.POINTER(_System), ctypes.POINTER(_Params)]
    L.stda_hip_select.argtypes = [P, ctypes.POINTER(I64), ctypes.POINTER(I64), ctypes.POINTER(I64)]
    L.stda_hip_kernel_stats.argtypes = [P, ctypes.c_char_p, ctypes.POINTER(F64),
                                        ctypes.POINTER(F64), ctypes.POINTER(F64)]
    L.stda_hip_batch_create.argtypes = [ctypes.POINTER(P), ctypes.c_int, ctypes.c_int]
    L.stda_hip_batch_destroy.argtypes = [P]
    L.stda_hip_batch_last_error.restype = ctypes.c_char_p
    L.stda_hip_batch_last_error.argtypes = [P]
    L.stda_hip_batch_solve_tda.argtypes = [P, I32, ctypes.POINTER(_System),
                                           ctypes.POINTER(_Params), I32, F64, I32, P, P, P, P]
    L.stda_hip_batch_timings.argtypes = [P, P, P]
    L.stda_hip_batch_kernel_bytes.argtypes = [P, ctypes.POINTER(F64)]
    L.stda_hip_batch_submit.argtypes = [P, ctypes.POINTER(_System), ctypes.POINTER(_Params), I32,
                                        F64, I32, ctypes.POINTER(I64)]
    L.stda_hip_batch_result.argtypes = [P, I64, I32, ctypes.POINTER(I32), ctypes.POINTER(I32), P,
                                        P, ctypes.POINTER(I32)]
    L.stda_hip_batch_stop.argtypes = [P]
    L.stda_hip_batch_configure.argtypes = [P, I64, I32]
    L.stda_hip_batch_configure_wide.argtypes = [P, I32, I64]
    L.stda_hip_batch_set_precision.argtypes = [P, I32]
    L.stda_hip_batch_wide_count.argtypes = [P, ctypes.POINTER(I64)]
    L.stda_hip_batch_phase_times.argtypes = [P, P]
    L.stda_hip_batch_kernel_stats.argtypes = [P, ctypes.c_char_p, ctypes.POINTER(F64),
                                              ctypes.POINTER(F64), ctypes.POINTER(F64)]
    _libs[path] = L
    return L


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


@dataclass
class Solution:
    omega: np.ndarray                    # Hartree
    X: np.ndarray = None                 # (nci, k)
    Y: np.ndarray = None                 # (nci, k) RPA only
    props: np.ndarray = None             # (k, 4): fL, fV, RL, RV or Rv(corr)
    alpha_sos: np.ndarray = None         # (3, 3)
    timings_ms: dict = field(default_factory=dict)

    def tda_dat_rows(self):
        """Rows of tda.dat's DATXY block as the reference formats them."""
        return [(k + 1, round(float(self.omega[k]) * HARTREE_EV_PRINT, 4),
                 *[round(float(x), 6) for x in self.props[k]]) for k in range(len(self.omega))]


def _alpha3(a6):
    xx, xy, yy, xz, yz, zz = a6
    return np.array([[xx, xy, xz], [xy, yy, yz], [xz, yz, zz]])


class StdaHip:
    """One device context of the excited-state core."""

    def __init__(self, device: int = 0, ilp64: bool = False):
        """``ilp64`` selects the library built with 64-bit index arrays."""
        self._L = load_library(LIB_PATH_ILP64 if ilp64 else LIB_PATH)
        self._idx = np.int64 if ilp64 else np.int32
        if self._L.stda_hip_index_size() != np.dtype(self._idx).itemsize:
            raise RuntimeError("library index size does not match the requested model")
        self._h = ctypes.c_void_p()
        self._check(self._L.stda_hip_create(ctypes.byref(self._h), device))
        self.nci = self.nprim = self.nsec = 0
        self._keep = []

    def close(self):
        if self._h:
            self._L.stda_hip_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, st):
        if st != 0:
            msg = self._L.stda_hip_error_string(st).decode()
            detail = self._L.stda_hip_last_error(self._h).decode() if self._h else ""
            raise StdaHipError(st, f"{msg}: {detail}" if detail else msg)

    # ---------------------------------------------------------------- setup
    def set_selection(self, dual=None, pt2_off: bool = False):
        """Selection variants for the next setup (stda_hip_set_selection): dual = (thr2_ev,
        atoms 1-based) as std2 -dual reads them from its 'dual' file; pt2_off: primaries only."""
        atoms = np.ascontiguousarray([] if dual is None else dual[1], dtype=self._idx)
        thr2 = 0.0 if dual is None else float(dual[0])
        self._check(self._L.stda_hip_set_selection(self._h, thr2, len(atoms),
                                                   _ptr(atoms) if len(atoms) else None,
                                                   1 if pt2_off else 0))

    def setup(self, ci: CoreInputs, thr: float, ax: float, thrp: float = 1e-4,
              alphak: float = -100.0, betaj: float = -100.0, fthr: float | None = None,
              triplet: bool = False, kshift=None):
        """Stage 1: gamma, charges, screening, PT2 selection (stda.f:454-790)."""
        clow = np.asfortranarray(ci.clow, dtype=np.float64)
        eps = np.ascontiguousarray(ci.epsi, dtype=np.float64)
        iaoat = np.ascontiguousarray(ci.iaoat, dtype=self._idx)
        xyz = np.asfortranarray(ci.xyz, dtype=np.float64)
        mom = np.ascontiguousarray(ci.moments, dtype=np.float64) if ci.moments is not None else None
        self._keep = [clow, eps, iaoat, xyz, mom]
        sysd = _System(ci.ncent, ci.nao, ci.no, ci.nv, _ptr(clow), _ptr(eps), _ptr(iaoat),
                       _ptr(xyz), _ptr(mom))
        ks = kshift or (0.5, 0.1, 4.0)
        par = _Params(thr, thrp, -1.0 if fthr is None else fthr, ax, alphak, betaj,
                      1 if triplet else 0, 1 if kshift else 0, ks[0], ks[1], ks[2])
        self._check(self._L.stda_hip_load(self._h, ctypes.byref(sysd), ctypes.byref(par)))
        self.thr = thr / HARTREE_EV
        return self.select()

    def load(self, ci: CoreInputs, thr: float, ax: float, **kw):
        """Copy the system to the device only (see :meth:`select`)."""
        clow = np.asfortranarray(ci.clow, dtype=np.float64)
        eps = np.ascontiguousarray(ci.epsi, dtype=np.float64)
        iaoat = np.ascontiguousarray(ci.iaoat, dtype=self._idx)
        xyz = np.asfortranarray(ci.xyz, dtype=np.float64)
        mom = np.ascontiguousarray(ci.moments, dtype=np.float64) if ci.moments is not None else None
        self._keep = [clow, eps, iaoat, xyz, mom]
        sysd = _System(ci.ncent, ci.nao, ci.no, ci.nv, _ptr(clow), _ptr(eps), _ptr(iaoat),
                       _ptr(xyz), _ptr(mom))
        ks = kw.get("kshift") or (0.5, 0.1, 4.0)
        fthr = kw.get("fthr")
        par = _Params(thr, kw.get("thrp", 1e-4), -1.0 if fthr is None else fthr, ax,
                      kw.get("alphak", -100.0), kw.get("betaj", -100.0),
                      1 if kw.get("triplet") else 0, 1 if kw.get("kshift") else 0,
                      ks[0], ks[1], ks[2])
        self._check(self._L.stda_hip_load(self._h, ctypes.byref(sysd), ctypes.byref(par)))
        self.thr = thr / HARTREE_EV

    def select(self):
        """Device pipeline on the loaded system: gamma, charges, screening, PT2."""
        nci, nprim, nsec = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        self._check(self._L.stda_hip_select(self._h, ctypes.byref(nci), ctypes.byref(nprim),
                                            ctypes.byref(nsec)))
        self.nci, self.nprim, self.nsec = nci.value, nprim.value, nsec.value
        return self.nci, self.nprim, self.nsec

    def set_precision(self, precision: str):
        """'fp64' (default), 'mixed' (FP32 contractions, FP64 Davidson numerics
        and a final FP64 Rayleigh-Ritz refinement) or 'bf16' (split-bfloat16
        contractions first, then as 'mixed')."""
        code = {"fp64": 0, "mixed": 1, "bf16": 2}[precision]
        self._check(self._L.stda_hip_set_precision(self._h, code))

    def gemm(self, A, B, C=None, alpha=1.0, beta=0.0, ta=False, tb=False, mode="fp64"):
        """C = alpha op(A) op(B) + beta C with the library's matrix-core GEMM
        (stda_hip_gemm); mode 'fp64', 'fp32' or 'bf16' (split-bfloat16)."""
        A = np.asfortranarray(A, dtype=np.float64)
        B = np.asfortranarray(B, dtype=np.float64)
        m = A.shape[1] if ta else A.shape[0]
        k = A.shape[0] if ta else A.shape[1]
        n = B.shape[0] if tb else B.shape[1]
        C = np.zeros((m, n), order="F") if C is None else np.array(C, dtype=np.float64, order="F")
        code = {"fp64": 0, "fp32": 1, "bf16": 2}[mode]
        self._check(self._L.stda_hip_gemm(self._h, code, int(ta), int(tb), m, n, k, alpha, _ptr(A),
                                          max(1, A.shape[0]), _ptr(B), max(1, B.shape[0]), beta,
                                          _ptr(C), max(1, m)))
        return C

    def kernel_stats(self, name: str):
        ms, n, w = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        self._check(self._L.stda_hip_kernel_stats(self._h, name.encode(), ctypes.byref(ms),
                                                  ctypes.byref(n), ctypes.byref(w)))
        return {"ms": ms.value, "launches": n.value, "work": w.value}

    def csf(self):
        iconf = np.zeros(2 * self.nci, dtype=self._idx)
        ed = np.zeros(self.nci)
        self._check(self._L.stda_hip_get_csf(self._h, _ptr(iconf), _ptr(ed)))
        return np.stack([iconf[:self.nci], iconf[self.nci:]], axis=1), ed

    # ---------------------------------------------------------------- solvers
    def solve_tda(self, nroots: int | None, velcorr: bool = True, tol: float = 1e-8,
                  vectors: bool = True) -> Solution:
        cap = self.nci if nroots is None else min(nroots, self.nci)
        omega = np.zeros(cap)
        X = np.zeros((self.nci, cap), order="F") if vectors else None
        props = np.zeros((cap, 5))
        a6 = np.zeros(6)
        nf = ctypes.c_int32()
        self._check(self._L.stda_hip_solve_tda(self._h, 0 if nroots is None else nroots, cap,
                                               1 if velcorr else 0, tol, ctypes.byref(nf),
                                               _ptr(omega), _ptr(X), _ptr(props), _ptr(a6)))
        k = nf.value
        return Solution(omega=omega[:k], X=None if X is None else X[:, :k],
                        props=props[:k, :4], alpha_sos=_alpha3(a6), timings_ms=self.timings())

    def solve_rpa(self, nroots: int | None, tol: float = 1e-8, vectors: bool = True) -> Solution:
        cap = self.nci if nroots is None else min(nroots, self.nci)
        omega = np.zeros(cap)
        X = np.zeros((self.nci, cap), order="F") if vectors else None
        Y = np.zeros((self.nci, cap), order="F") if vectors else None
        props = np.zeros((cap, 5))
        a6 = np.zeros(6)
        nf = ctypes.c_int32()
        self._check(self._L.stda_hip_solve_rpa(self._h, 0 if nroots is None else nroots, cap, tol,
                                               ctypes.byref(nf), _ptr(omega), _ptr(X), _ptr(Y),
                                               _ptr(props), _ptr(a6)))
        k = nf.value
        return Solution(omega=omega[:k], X=None if X is None else X[:, :k],
                        Y=None if Y is None else Y[:, :k], props=props[:k, :4],
                        alpha_sos=_alpha3(a6), timings_ms=self.timings())

    def lresp(self, omegas, tol: float = 1e-10, vectors: bool = False):
        """alpha(w) for each frequency (Hartree): list of (w, 3x3).

        ``tol`` bounds the residual norm of the response equations; near a
        pole alpha amplifies it, so the default is tighter than the
        eigensolvers' 1e-8 (alpha within 1e-6 relative of the exact solve)."""
        w = np.ascontiguousarray(omegas, dtype=np.float64)
        nf = len(w)
        alpha = np.zeros(9 * nf)
        xpy = np.zeros(self.nci * 3 * nf) if vectors else None
        self._check(self._L.stda_hip_lresp(self._h, nf, _ptr(w), tol, _ptr(alpha), _ptr(xpy)))
        out = [(float(w[f]), alpha[9 * f:9 * f + 9].reshape(3, 3, order="F")) for f in range(nf)]
        if vectors:
            return out, xpy.reshape(self.nci, 3 * nf, order="F")
        return out

    # ---------------------------------------------------------------- explicit matrices
    def build_a(self):
        A = np.zeros((self.nci, self.nci), order="F")
        self._check(self._L.stda_hip_build_a(self._h, _ptr(A)))
        return A

    def build_apb_amb(self):
        npk = self.nci * (self.nci + 1) // 2
        apb, amb = np.zeros(npk), np.zeros(npk)
        self._check(self._L.stda_hip_build_apb_amb(self._h, _ptr(apb), _ptr(amb)))
        return apb, amb

    def build_bhalf(self):
        b = np.zeros(self.nci * (self.nci + 1) // 2)
        self._check(self._L.stda_hip_build_bhalf(self._h, _ptr(b)))
        return b

    def matvec(self, which: int, V: np.ndarray) -> np.ndarray:
        V = np.asfortranarray(V, dtype=np.float64).reshape(self.nci, -1, order="F")
        out = np.zeros_like(V, order="F")
        self._check(self._L.stda_hip_matvec(self._h, which, V.shape[1], _ptr(V), _ptr(out)))
        return out

    def beta_shg(self, omegas, tol: float = 1e-10):
        """SHG beta(-2w; w, w) per frequency (stda_hip_beta_shg): [(w, beta (3,3,3), hrs)] with
        hrs = {"beta_hrs", "dr", "b2_zzz", "b2_xzz"}; beta[i, j, k] as the reference's
        beta(i+1, j+1, k+1)."""
        om = np.ascontiguousarray(omegas, dtype=np.float64)
        b = np.zeros(27 * len(om))
        h = np.zeros(4 * len(om))
        self._check(self._L.stda_hip_beta_shg(self._h, len(om), _ptr(om), tol, _ptr(b), _ptr(h)))
        out = []
        for f, w in enumerate(om.tolist()):
            t = b[27 * f:27 * (f + 1)].reshape((3, 3, 3), order="F")
            hv = h[4 * f:4 * (f + 1)]
            out.append((w, t, {"beta_hrs": hv[0], "dr": hv[1], "b2_zzz": hv[2], "b2_xzz": hv[3]}))
        return out

    def ao_ints(self, mol):
        """One-electron AO integrals on the device (stda_hip_ao_ints), in the layout of
        :func:`stda_amd.aoints.one_electron`: {"S": (nao, nao), "R", "M", "V": (3, nao, nao)}."""
        nao = mol.nao
        ptr = np.ascontiguousarray(mol.ao_prim_ptr, dtype=np.int64)
        al = np.ascontiguousarray(mol.prim_alpha, dtype=np.float64)
        co = np.ascontiguousarray(mol.prim_coef, dtype=np.float64)
        cen = np.ascontiguousarray(mol.xyz[mol.ao_atom], dtype=np.float64)
        lmn = np.ascontiguousarray(mol.ao_lmn, dtype=np.int32)
        S = np.zeros((nao, nao), order="F")
        ao9 = np.zeros((9, nao, nao))
        self._check(self._L.stda_hip_ao_ints(self._h, nao, _ptr(ptr), _ptr(al), _ptr(co),
                                             _ptr(cen), _ptr(lmn), _ptr(S), _ptr(ao9)))
        # ao9[q] holds a column-major matrix: transpose the C-order view
        m = [ao9[q].T for q in range(9)]
        return {"S": S, "R": np.stack(m[0:3]), "M": np.stack(m[3:6]), "V": np.stack(m[6:9])}

    def prepare_ao(self, mol, ca, bfw: bool = False):
        """AO integrals, moment transforms and Loewdin on the device without a host round
        trip of the AO matrices (stda_hip_prepare_ao): (clow, moments)."""
        ca = np.asfortranarray(ca, dtype=np.float64)
        nao, moci = ca.shape
        ptr = np.ascontiguousarray(mol.ao_prim_ptr, dtype=np.int64)
        al = np.ascontiguousarray(mol.prim_alpha, dtype=np.float64)
        co = np.ascontiguousarray(mol.prim_coef, dtype=np.float64)
        cen = np.ascontiguousarray(mol.xyz[mol.ao_atom], dtype=np.float64)
        lmn = np.ascontiguousarray(mol.ao_lmn, dtype=np.int32)
        clow = np.zeros((nao, moci), order="F")
        mom = np.zeros((9, moci * (moci + 1) // 2))
        self._check(self._L.stda_hip_prepare_ao(self._h, nao, _ptr(ptr), _ptr(al), _ptr(co),
                                                _ptr(cen), _ptr(lmn), moci, _ptr(ca),
                                                1 if bfw else 0, _ptr(clow), _ptr(mom)))
        return clow, mom

    def transform(self, ca, S, ao9, bfw: bool = False):
        """Device upstream transforms (stda_hip_transform): (clow (nao, moci), moments (9, npk))
        from the window's MO coefficients ca, the overlap S and the nine AO moment matrices."""
        ca = np.asfortranarray(ca, dtype=np.float64)
        nao, moci = ca.shape
        Sf = np.asfortranarray(S, dtype=np.float64)
        a9 = np.ascontiguousarray(np.stack([np.asfortranarray(a, dtype=np.float64).ravel(order="F")
                                            for a in ao9]))
        clow = np.zeros((nao, moci), order="F")
        mom = np.zeros((9, moci * (moci + 1) // 2))
        self._check(self._L.stda_hip_transform(self._h, nao, moci, _ptr(ca), _ptr(Sf), _ptr(a9),
                                               1 if bfw else 0, _ptr(clow), _ptr(mom)))
        return clow, mom

    def props_xyz(self, omega, X, velcorr: bool = True):
        """Axis-resolved fL, fV, RL, RV of TDA states (stda_hip_props_xyz): (k, 3, 4)."""
        w = np.ascontiguousarray(omega, dtype=np.float64)
        Xf = np.asfortranarray(X, dtype=np.float64).reshape(self.nci, -1, order="F")
        out = np.zeros(12 * len(w))
        self._check(self._L.stda_hip_props_xyz(self._h, len(w), _ptr(w), _ptr(Xf), self.nci,
                                               1 if velcorr else 0, _ptr(out)))
        return out.reshape(len(w), 3, 4)

    def tpa(self, eci, X, Y, tol: float = 1e-10):
        """Two-photon absorption of the given RPA states (stda_hip_tpa): [(E, sigma (3,3),
        delta)] with delta = {"par", "perp", "circ", "rho"}."""
        e = np.ascontiguousarray(eci, dtype=np.float64)
        Xf = np.asfortranarray(X, dtype=np.float64).reshape(self.nci, -1, order="F")
        Yf = np.asfortranarray(Y, dtype=np.float64).reshape(self.nci, -1, order="F")
        k = len(e)
        sig = np.zeros(9 * k)
        d = np.zeros(4 * k)
        self._check(self._L.stda_hip_tpa(self._h, k, _ptr(e), _ptr(Xf), _ptr(Yf), self.nci, tol,
                                         _ptr(sig), _ptr(d)))
        return [(float(e[q]), sig[9 * q:9 * q + 9].reshape((3, 3), order="F"),
                 dict(zip(["par", "perp", "circ", "rho"], d[4 * q:4 * q + 4].tolist())))
                for q in range(k)]

    def gspec_curves(self, offsets, x, y, wei, emin, step, npts: int, flags: int,
                     rfaktor: float):
        """Broadened curves of len(offsets)-1 line spectra (stda_hip_gspec): (e, r), each
        (nspec, npts)."""
        off = np.ascontiguousarray(offsets, dtype=np.int64)
        arr = [np.ascontiguousarray(a, dtype=np.float64) for a in (x, y, wei, emin, step)]
        nspec = len(off) - 1
        e = np.zeros((nspec, npts))
        r = np.zeros((nspec, npts))
        self._check(self._L.stda_hip_gspec(self._h, nspec, _ptr(off), *[_ptr(a) for a in arr],
                                           npts, flags, rfaktor, _ptr(e), _ptr(r)))
        return e, r

    def timings(self):
        t = np.zeros(6)
        self._L.stda_hip_timings(self._h, _ptr(t))
        return dict(zip(["intermediates", "screen_sort", "pt2", "solve", "properties", "setup_total"],
                        t.tolist()))


def stda(ci: CoreInputs, thr: float, thrp: float, ax: float, alphak: float = -100.0,
         betaj: float = -100.0, fthr: float | None = None, nroots: int | None = None,
         rpa: bool = False, triplet: bool = False, velcorr: bool | None = None,
         resp_omegas=None, tol: float = 1e-8, device: int = 0, beta_omegas=None,
         tpa_states: int | None = None, dual=None, pt2_off: bool = False, kshift=None):
    """Drop-in for the reference's ``stda`` hot path (stda.f:454-1471).

    Returns ``(Solution | list-of-alpha | list-of-beta | list-of-2PA, csf)`` where
    ``csf`` = (iconf, ed).  ``resp_omegas`` (Hartree) selects the linear-response
    branch of ``-aresp`` (stda.f:905-920) instead of the eigenvalue problem;
    ``beta_omegas`` the SHG hyperpolarizability of ``-resp`` (lresp1, stda.f:878-903);
    ``tpa_states`` = N the two-photon absorption of the N lowest RPA states of
    ``-2PA N`` (stda.f:1034-1048), here of the properly normalised states of the RPA
    solve (the reference's srpapack forms them with doubled diagonals in that mode;
    :meth:`StdaHip.tpa` takes any states).  ``dual`` = (thr2_ev, atoms) and ``pt2_off``
    are the selection variants of ``-dual`` / ``-rw 1`` (:meth:`StdaHip.set_selection`).
    ``kshift`` = (shftmax eV, width eV, steepness) switches on the K(ia,ia) shift of
    xTB mode (``dokshift``); ``stda_amd.xtb.XTB_PARAMETERS`` holds the whole ``-xtb`` set
    (a_x, alphak, betaj, kshift) for an input read with ``stda_amd.xtb.read_wfn_xtb``.
    """
    if velcorr is None:
        velcorr = not rpa                 # main.f:102, 662
    with StdaHip(device) as h:
        if dual is not None or pt2_off:
            h.set_selection(dual=dual, pt2_off=pt2_off)
        h.setup(ci, thr, ax, thrp=thrp, alphak=alphak, betaj=betaj, fthr=fthr, triplet=triplet,
                kshift=kshift)
        csf = h.csf()
        if beta_omegas is not None:
            return h.beta_shg(beta_omegas), csf
        if tpa_states is not None:
            sol = h.solve_rpa(tpa_states, tol=tol)
            return h.tpa(sol.omega, sol.X, sol.Y), csf
        if resp_omegas is not None:
            return h.lresp(resp_omegas, tol=tol), csf
        if rpa:
            return h.solve_rpa(nroots, tol=tol), csf
        return h.solve_tda(nroots, velcorr=velcorr, tol=tol), csf


class StdaHipBatch:
    """Batched solves of many independent molecules (``stda_hip_batch_*``).

    ``workers`` device contexts, each on its own host thread and stream,
    prepare the molecules concurrently (selection, explicit FP64/FP32 A,
    0.5 B, CSF moments).  Molecules with at most ``nmax_batched`` CSFs and at
    most 24 roots join a lock-step solver as soon as they are ready: per
    iteration one launch computes the operator products of all of them and
    one advances each by a Davidson step (FP32 operator products, FP64
    subspace numerics, one FP64 Rayleigh-Ritz refinement).  Larger molecules
    take the single-molecule mixed path on their worker.  See
    :meth:`configure`.
    """

    def __init__(self, device: int = 0, workers: int = 8):
        self._L = load_library()
        self._b = ctypes.c_void_p()
        self._pending = {}
        st = self._L.stda_hip_batch_create(ctypes.byref(self._b), device, workers)
        if st != 0:
            raise StdaHipError(st, self._L.stda_hip_error_string(st).decode())

    def close(self):
        if self._b:
            self._L.stda_hip_batch_destroy(self._b)
            self._b = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def solve_tda(self, cis, thr: float, ax: float, nroots: int, tol: float = 1e-8,
                  velcorr: bool = True, thrp: float = 1e-4, fthr: float | None = None,
                  triplet: bool = False):
        """Lowest ``nroots`` sTDA roots of every molecule in ``cis``.

        Returns a list with, per molecule, ``(omega, props, status)``: omega in
        Hartree (nfound,), props (nfound, 4) = fL, fV, RL, Rv(corr), status 0
        or an ``STDA_HIP_ERR_*`` code.
        """
        nmol = len(cis)
        keep = []
        sysarr = (_System * max(nmol, 1))()
        pararr = (_Params * max(nmol, 1))()
        for i, ci in enumerate(cis):
            clow = np.asfortranarray(ci.clow, dtype=np.float64)
            eps = np.ascontiguousarray(ci.epsi, dtype=np.float64)
            iaoat = np.ascontiguousarray(ci.iaoat, dtype=np.int32)
            xyz = np.asfortranarray(ci.xyz, dtype=np.float64)
            mom = (np.ascontiguousarray(ci.moments, dtype=np.float64)
                   if ci.moments is not None else None)
            keep.append((clow, eps, iaoat, xyz, mom))
            sysarr[i] = _System(ci.ncent, ci.nao, ci.no, ci.nv, _ptr(clow), _ptr(eps),
                                _ptr(iaoat), _ptr(xyz), _ptr(mom))
            pararr[i] = _Params(thr, thrp, -1.0 if fthr is None else fthr, ax, -100.0, -100.0,
                                1 if triplet else 0, 0, 0.5, 0.1, 4.0)
        nf = np.zeros(max(nmol, 1), dtype=np.int32)
        om = np.zeros((max(nmol, 1), nroots))
        pr = np.zeros((max(nmol, 1), nroots, 5))
        stt = np.zeros(max(nmol, 1), dtype=np.int32)
        st = self._L.stda_hip_batch_solve_tda(self._b, nmol, sysarr, pararr, nroots, tol,
                                              1 if velcorr else 0, _ptr(nf), _ptr(om), _ptr(pr),
                                              _ptr(stt))
        if st != 0:
            raise StdaHipError(st, f"{self._L.stda_hip_error_string(st).decode()}: "
                                   f"{self._L.stda_hip_batch_last_error(self._b).decode()}")
        del keep
        return [(om[i, :nf[i]].copy(), pr[i, :nf[i], :4].copy(), int(stt[i]))
                for i in range(nmol)]

    def last_error(self) -> str:
        return self._L.stda_hip_batch_last_error(self._b).decode()

    # ---------------------------------------------------------------- continuous mode
    def submit(self, ci: CoreInputs, thr: float, ax: float, nroots: int, tol: float = 1e-8,
               velcorr: bool = True, thrp: float = 1e-4, fthr: float | None = None,
               triplet: bool = False) -> int:
        """Queue one molecule for the continuous mode (stda_hip_batch_submit); it joins
        the running lock-step set as soon as its setup is done.  Returns its ticket."""
        clow = np.asfortranarray(ci.clow, dtype=np.float64)
        eps = np.ascontiguousarray(ci.epsi, dtype=np.float64)
        iaoat = np.ascontiguousarray(ci.iaoat, dtype=np.int32)
        xyz = np.asfortranarray(ci.xyz, dtype=np.float64)
        mom = (np.ascontiguousarray(ci.moments, dtype=np.float64)
               if ci.moments is not None else None)
        sysd = _System(ci.ncent, ci.nao, ci.no, ci.nv, _ptr(clow), _ptr(eps), _ptr(iaoat),
                       _ptr(xyz), _ptr(mom))
        par = _Params(thr, thrp, -1.0 if fthr is None else fthr, ax, -100.0, -100.0,
                      1 if triplet else 0, 0, 0.5, 0.1, 4.0)
        t = ctypes.c_int64()
        st = self._L.stda_hip_batch_submit(self._b, ctypes.byref(sysd), ctypes.byref(par), nroots,
                                           tol, 1 if velcorr else 0, ctypes.byref(t))
        if st != 0:
            raise StdaHipError(st, self._L.stda_hip_error_string(st).decode())
        # the library reads these arrays until the molecule is finished
        self._pending[t.value] = (nroots, (clow, eps, iaoat, xyz, mom))
        return t.value

    def result(self, ticket: int, wait: bool = True):
        """``(omega, props, status)`` of a submitted molecule as :meth:`solve_tda` gives
        them, or None when it is not finished and ``wait`` is false."""
        nroots = self._pending[ticket][0]
        done = ctypes.c_int32()
        nf = ctypes.c_int32()
        om = np.zeros(nroots)
        pr = np.zeros((nroots, 5))
        stt = ctypes.c_int32()
        st = self._L.stda_hip_batch_result(self._b, ticket, 1 if wait else 0, ctypes.byref(done),
                                           ctypes.byref(nf), _ptr(om), _ptr(pr),
                                           ctypes.byref(stt))
        if st != 0:
            raise StdaHipError(st, self._L.stda_hip_error_string(st).decode())
        if not done.value:
            return None
        del self._pending[ticket]
        return om[:nf.value].copy(), pr[:nf.value, :4].copy(), int(stt.value)

    def stop(self):
        """Finish every submitted molecule and leave the continuous mode."""
        self._L.stda_hip_batch_stop(self._b)

    def configure(self, nmax_batched: int = 1024, poll: int = 4):
        """CSF limit of the lock-step solver and iterations between polls."""
        st = self._L.stda_hip_batch_configure(self._b, nmax_batched, poll)
        if st != 0:
            raise StdaHipError(st, self._L.stda_hip_error_string(st).decode())

    def configure_wide(self, enable: bool = True, two_phase: bool = False, nmax_wide: int = 0):
        """Molecules above the explicit-operator CSF limit (and at most
        nmax_wide CSFs; 0: no limit) in the matrix-free lock step (True) or on
        the single-molecule path of their worker (False); two_phase: the lock
        step of a call starts once every molecule is prepared."""
        st = self._L.stda_hip_batch_configure_wide(self._b, (2 if two_phase else 1) if enable else 0,
                                                   int(nmax_wide))
        if st != 0:
            raise StdaHipError(st, self._L.stda_hip_error_string(st).decode())

    def set_precision(self, precision: str):
        """Operator products of the single-path solves: "mixed" (FP32, the
        default) or "bf16" (split-bfloat16 phase, then FP32, then FP64)."""
        code = {"mixed": 1, "bf16": 2}[precision]
        st = self._L.stda_hip_batch_set_precision(self._b, code)
        if st != 0:
            raise StdaHipError(st, self._L.stda_hip_error_string(st).decode())

    def wide_count(self) -> int:
        """Molecules of the last call solved in the matrix-free lock step."""
        v = ctypes.c_int64(0)
        self._L.stda_hip_batch_wide_count(self._b, ctypes.byref(v))
        return int(v.value)

    def phase_times(self):
        """Batched-kernel phase times summed over the molecules (us) and the
        total Jacobi sweeps of the last call."""
        t = np.zeros(10)
        self._L.stda_hip_batch_phase_times(self._b, _ptr(t))
        names = ["gram", "jacobi", "ritz", "expand", "final", "unused", "sweeps",
                 "worker_setup", "worker_matrices", "worker_single"]
        return dict(zip(names, t.tolist()))

    def kernel_stats(self, name: str):
        """Per-kernel statistics summed over the workers and molecules of the
        last call: HIP-event ms, launches, algorithmic work."""
        ms, n, w = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        self._L.stda_hip_batch_kernel_stats(self._b, name.encode(), ctypes.byref(ms),
                                            ctypes.byref(n), ctypes.byref(w))
        return {"ms": ms.value, "launches": n.value, "work": w.value}

    def kernel_bytes(self) -> float:
        """Algorithmic bytes of the last batched Davidson launch."""
        v = ctypes.c_double()
        self._L.stda_hip_batch_kernel_bytes(self._b, ctypes.byref(v))
        return v.value

    def timings(self):
        """Host times (ms) of the last call's phases, the batched kernel's
        HIP-event time, and how many molecules took which path
        ("nonzero_status": no CSF in the window, or not converged)."""
        t = np.zeros(5)
        cnt = np.zeros(3, dtype=np.int64)
        self._L.stda_hip_batch_timings(self._b, _ptr(t), _ptr(cnt))
        return ({"solve": t[0], "driver": t[1], "properties": t[2], "total": t[3],
                 "batch_kernels": t[4]},
                {"batched": int(cnt[0]), "single": int(cnt[1]), "nonzero_status": int(cnt[2])})
