"""Spectral output: the tda.dat file and the g_spec broadening program.

* :func:`tda_dat` writes the file the reference's driver leaves for spectrum
  plotting, in either of its two layouts: the velocity-corrected table of
  ``apbtrafo`` (apbtrafo.f:115-128, 259-260; default sTDA) or
  ``print_tdadat`` (stda.f:2235-2265; sTD-DFT and uncorrected sTDA).
* :func:`tda_dat_xyz` writes the axis-resolved files of ``-aniso`` (tdax/y/z.dat);
  :func:`beta_hrs_file` and :func:`tpa_abs_file` the beta_HRS and 2PA-abs files of
  ``-resp`` and ``-2PA``; :func:`sos_rotation` is the sum-over-states optical rotation
  table of every sTDA run.
* :func:`gspec` / :func:`gspec_many` are the reference's spectrum program
  ``g_spec`` (g_spec/g_spec.f): they read the text of a tda.dat-style input and
  return the contents of ``rots.dat`` (line spectrum) and ``spec.dat``
  (Gaussian-broadened curve).  The keyword parsing, sorting, grid and line
  spectrum are host work done here; the broadening, nspec x npts x lines
  exp terms, runs in the HIP kernel behind ``stda_hip_gspec``, all spectra
  of a call in one launch.  ``python -m stda_amd.spectrum < tda.dat`` writes
  rots.dat and spec.dat in the working directory like the reference program.

Numbers are read with the reference's own reader (READL/READAA,
g_spec.f:689-790: at most ten digits before and after the point, the
fraction built by repeated division by ten) and written with Fortran's
edit descriptors, so that outputs match the reference program's files.
"""

from __future__ import annotations

import math
import re
import sys
from dataclasses import dataclass, field

import numpy as np

_F = np.float32
EV_PRINT = float(_F(27.21139))        # apbtrafo.f:259 (real*4 literal, widened)
EV_CM = float(_F(8065.54093))         # g_spec.f:325
MARGIN = float(_F(0.40))              # g_spec.f:258
ETHR_MIN = float(_F(0.0001))          # g_spec.f:335
FAU = 27.21139570                     # g_spec.f:76

FLAG_UV, FLAG_NM, FLAG_CM, FLAG_LOG = 1, 2, 4, 8   # STDA_HIP_GSPEC_*


# ---------------------------------------------------------------- Fortran edit descriptors
def _nonfinite(v: float, w: int):
    if v != v:
        return "NaN".rjust(w)
    if v in (math.inf, -math.inf):
        return ("Inf" if v > 0 else "-Inf").rjust(w)
    return None


def fmt_f(v: float, w: int = 16, d: int = 6) -> str:
    """Fw.d; asterisks when the value does not fit."""
    t = _nonfinite(v, w)
    if t is not None:
        return t
    s = "%.*f" % (d, v)
    return s.rjust(w) if len(s) <= w else "*" * w


def fmt_e(v: float, w: int = 14, d: int = 4) -> str:
    """Ew.d: 0.d...d mantissa, E+xx exponent, or +xxx without the E from 100 on."""
    t = _nonfinite(v, w)
    if t is not None:
        return t
    sign = "-" if math.copysign(1.0, v) < 0 else ""
    if v == 0.0:
        mant, ex = "0" * d, 0
    else:
        m, e = ("%.*e" % (d - 1, abs(v))).split("e")
        mant, ex = m.replace(".", ""), int(e) + 1
    tail = ("E%+03d" % ex) if -99 <= ex <= 99 else ("%+04d" % ex)
    return (sign + "0." + mant + tail).rjust(w)


def fmt_list_real8(v: float) -> str:
    """List-directed output of one real*8 as gfortran writes it (blank, then G25.17 field)."""
    a = abs(v)
    if a != 0.0 and (a < 0.1 or a >= 1e17):
        m, e = ("%.16E" % v).split("E")
        return " " + (m + "E%+04d" % int(e)).rjust(25)
    k = 1 if a == 0.0 else len(str(int(a))) if a >= 1 else 0
    return " " + ("%.*f" % (17 - k, v)).rjust(20) + " " * 5


def fmt_list_real4(v: float) -> str:
    """List-directed output of one real*4 as gfortran writes it (blank, then a G16.9 field:
    9 significant digits, F form for 0.1 <= |v| < 1e9, zero as 0.00000000)."""
    x = float(np.float32(v))
    a = abs(x)
    if a != 0.0 and (a < 0.1 or a >= 1e9):
        m, e = ("%.8E" % x).split("E")
        return " " + (m + "E%+03d" % int(e)).rjust(16)
    k = 1 if a == 0.0 else (len(str(int(a))) if a >= 1 else 0)
    return " " + ("%.*f" % (9 - k, x)).rjust(12) + " " * 4


def beta_hrs_file(omegas, beta_hrs) -> str:
    """The reference's beta_HRS file (linear_response.f:695, list-directed
    ``freq*27.21139, beta_HRS``: the frequency in eV formed in single precision)."""
    out = []
    for w, b in zip(omegas, beta_hrs):
        ev = np.float32(np.float32(w) * np.float32(27.21139))
        out.append(fmt_list_real4(float(ev)) + fmt_list_real8(float(b)))
    return "\n".join(out) + "\n"


def tpa_abs_file(eci, deltas) -> str:
    """The reference's 2PA-abs file (linear_response.f:4905-4909, format (I3,F7.3,4F20.3)):
    state, energy in eV (single precision) and delta parallel, perpendicular, circular, rho."""
    out = []
    for i, (e, d) in enumerate(zip(eci, deltas)):
        ev = float(np.float32(np.float32(e) * np.float32(27.21139)))
        out.append(("%3d" % (i + 1))[-3:] + fmt_f(ev, 7, 3) +
                   "".join(fmt_f(float(np.float32(x)), 20, 3) for x in d))
    return "\n".join(out) + "\n"


# ---------------------------------------------------------------- tda.dat
_HEADER_TAIL = [" MMASS", None, " LFAKTOR", "  0.5", " RFAKTOR", "  1.0", " WIDTH", "  0.20",
                " SHIFT", "  0.00", " DATXY"]


def tda_dat(omega, props, xmolw: float, velcorr: bool = True, thr: float | None = None) -> str:
    """Text of tda.dat for roots omega (Hartree) with props (k, >=4): fL, fV, RL, RV or Rv(corr).

    velcorr=True: the apbtrafo layout ('VELO', every root, energies
    omega * 27.21139 in double precision).  velcorr=False: the print_tdadat
    layout ('UV', roots above thr (Hartree) left out, energies formed in
    single precision as the reference's real*4 eigenvalues are).
    Rows: (i4, F10.4, 4f13.6)."""
    omega = np.asarray(omega, dtype=np.float64)
    props = np.asarray(props, dtype=np.float64)
    out = [" NM", " VELO" if velcorr else " UV"]
    for h in _HEADER_TAIL:
        out.append(fmt_list_real8(xmolw) if h is None else h)
    for i in range(len(omega)):
        if velcorr:
            ev = float(omega[i]) * EV_PRINT
        else:
            e32 = _F(omega[i])
            if thr is not None and float(e32) > thr:
                continue
            ev = float(_F(e32 * _F(27.21139)))
        out.append("%4d" % (i + 1) + fmt_f(ev, 10, 4) +
                   "".join(fmt_f(float(props[i, c]), 13, 6) for c in range(4)))
    return "\n".join(out) + "\n"


def tda_dat_xyz(omega, props_xyz, xmolw: float, velcorr: bool = True,
                thr: float | None = None):
    """Texts of tdax.dat, tday.dat and tdaz.dat (-aniso; apbtrafo.f:130-174, stda.f:1414-1425)
    from the axis-resolved properties (k, 3, 4) of ``StdaHip.props_xyz``, in the layout of
    :func:`tda_dat`."""
    p = np.asarray(props_xyz, dtype=np.float64)
    return tuple(tda_dat(omega, p[:, q, :], xmolw, velcorr, thr) for q in range(3))


# ---------------------------------------------------------------- SOS optical rotation
_SOS_LAMBDA = [float(_F(v)) for v in (632.8, 589.3, 579.0, 546.0, 436.0, 365.0)]   # sosor.f:58-63


def sos_rotation(omega, rv, xmolw: float, refindex: float = 1.4):
    """Sum-over-states specific optical rotation at six wavelengths, as ``sosor`` prints it
    after every sTDA run (sosor.f:19-112, called at stda.f:1471 with the roots and the
    velocity-corrected rotatory strengths): [(lambda nm, eV, [alpha] in
    grad cm^3 g^-1 dm^-1)], with the Lorentz factor of a solvent of refractive index 1.4.
    The roots enter in single precision, as the reference's real*4 eigenvalues do."""
    c_beta = float(_F(_F(2.0) * _F(137.036)) / _F(3.0))      # (2.*137.036/3.), real*4
    r_unit = float(_F(64604.8))
    reffaktor = (refindex ** 2 + 2.0) / 3.0
    x = [float(_F(w)) for w in np.asarray(omega, dtype=np.float64)]
    y = [float(v) for v in np.asarray(rv, dtype=np.float64)]
    out = []
    for lam in _SOS_LAMBDA:
        eau = 1.0e7 / lam / 2.19474625e5
        r1 = 0.0
        for xau, yi in zip(x, y):
            r1 = r1 + c_beta * (yi / r_unit) / (xau ** 2 - eau ** 2)
        vorfaktor = (38652.0 / xmolw) * (_SOS_LAMBDA[1] / lam) ** 2
        out.append((lam, 1.0e7 / (EV_CM * lam), r1 * vorfaktor * reffaktor))
    return out


def sos_rotation_text(rows, refval: float = 0.0) -> str:
    """The table sosor writes (formats 142/143: f6.1, f6.2, f12.2; the 589.3 nm line also
    carries the reference value read from a '.ref' file, 0 without one)."""
    lines = []
    for j, (lam, ev, a) in enumerate(rows):
        t = fmt_f(lam, 6, 1) + fmt_f(ev, 6, 2) + fmt_f(a, 12, 2)
        lines.append(t + fmt_f(refval, 12, 2) + " ##" if j == 1 else t)
    return "\n".join(lines) + "\n"


# ---------------------------------------------------------------- g_spec input
_SIMPLE = re.compile(r"-?(?:\d{1,10}(?:\.\d{0,10})?|\.\d{0,10})\Z")
_SCALE = [1.0]
for _ in range(10):
    _SCALE.append(_SCALE[-1] / 10)       # READAA's X = X/10, repeated


def read_numbers(line: str) -> list:
    """The numbers of one input line as READL reads them (g_spec.f:689-790).

    Lines of printable characters whose blank-separated tokens are all plain decimals (at
    most ten digits on either side of the point) take a direct path with READAA's
    arithmetic; anything else goes through the character scan."""
    a = line[:120].rstrip(" ")
    if len(a) < 120 and a.isprintable():
        vals = []
        for t in a.split():
            if not _SIMPLE.match(t):
                break
            neg = t[0] == "-"
            ip, _, fp = (t[1:] if neg else t).partition(".")
            v = float(int(ip)) if ip else 0.0
            if fp:
                v = v + float(int(fp)) * _SCALE[len(fp)]
            vals.append(-1.0 * v if neg else v)
        else:
            return vals
    return _read_numbers_scan(line)


def _read_numbers_scan(line: str) -> list:
    a = line[:120].ljust(120)
    nl = len(a)
    vals = []
    p = 0                                   # 0-based scan start
    while True:
        q = p
        while q < nl - 1:
            ch, nx = a[q], a[q + 1]
            if ch.isdigit() or ch == "." or (ch == "-" and (nx.isdigit() or nx == ".")):
                break
            q += 1
        else:
            return vals
        start = q
        ip = fp = 0.0
        sign = 1.0
        scale = 1.0
        nd = 0
        u = q
        if a[u] == "-":
            sign = -1.0
            u += 1
        while u < nl and a[u].isdigit() and nd < 10:
            ip = ip * 10 + int(a[u])
            nd += 1
            u += 1
        if u < nl and a[u] == ".":
            nd = 0
            u += 1
            while u < nl and a[u].isdigit() and nd < 10:
                fp = fp * 10 + int(a[u])
                scale = scale / 10
                nd += 1
                u += 1
        val = sign * (ip + fp * scale)
        # the token runs to the next blank; a D or E on the way starts an exponent
        u = start
        end = nl
        while u < nl:
            if a[u] == " ":
                end = u
                break
            if a[u] in "DE":
                ex, es = 0.0, 1.0
                v = u + 1
                end = nl
                while v < nl:
                    if a[v] == " ":
                        end = v
                        break
                    if a[v].isdigit():
                        ex = ex * 10.0 + int(a[v])
                    if a[v] == "-":
                        es = -1.0
                    v += 1
                val = val * 10.0 ** (es * ex)
                break
            u += 1
        vals.append(val)
        if end >= nl:
            return vals
        p = end


@dataclass
class GSpecInput:
    """One g_spec input after parsing, sorting and choosing the grid (g_spec.f:60-270)."""
    x: np.ndarray                      # line energies (eV), summation order
    y: np.ndarray
    wei: np.ndarray
    npts: int
    emin: float
    step: float
    flags: int
    rfaktor: float
    rots: str = ""
    opts: dict = field(default_factory=dict)


def _selection_order(key: np.ndarray) -> np.ndarray:
    """Permutation of dsort (g_spec.f:571-609): selection sort ascending with swaps, the last
    of equal minima taken first.  Distinct keys reduce to an argsort."""
    n = len(key)
    if n == 0 or len(np.unique(key)) == n:
        return np.argsort(key, kind="stable")
    ew = key.copy()
    perm = np.arange(n)
    for i in range(n - 1):
        seg = ew[i:]
        k = i + (len(seg) - 1 - int(np.argmin(seg[::-1])))
        if k != i:
            ew[i], ew[k] = ew[k], ew[i]
            perm[i], perm[k] = perm[k], perm[i]
    return perm


def parse_gspec(text: str) -> GSpecInput:
    """Keyword input of g_spec (g_spec.f:46-218), sort (:222-247), grid (:249-262) and the
    line spectrum written to rots.dat (:296-350)."""
    lines = text.split("\n")
    if lines and lines[-1] == "":
        lines.pop()
    imax = len(lines)
    o = dict(itype=2, uv=0, cm=0, nm=0, log=0, inv=0, fau=1.0, rfaktor=1.0, fktl=1.0, sh=0.0,
             ethr=0.0, forb=0.0, nol=False)
    xs = np.zeros(imax + 1)
    ys = np.zeros(imax + 1)
    ws = np.zeros(imax + 1)
    cols = [0.0] * 52                     # READL's array keeps values of earlier lines
    it = iter(lines)

    def line():
        return next(it)[:120].ljust(120)

    def nums(s):
        v = read_numbers(s)[:50]          # xx(50) in g_spec.f:14
        for q, val in enumerate(v):
            cols[q + 1] = val
        cols[len(v) + 1] = 0.0
        return len(v)

    k = 0
    try:
        while True:
            s = line()
            if "AU " in s:
                o["fau"] = FAU
            if "ETHR " in s:
                s = line()
                nums(s)
                o["ethr"] = cols[1]
            if "FORB " in s:
                s = line()
                nums(s)
                o["forb"] = cols[1]
            for key, name in (("UV", "uv"), ("NM", "nm"), ("CM", "cm"), ("LOG", "log"),
                              ("INV", "inv")):
                if key in s:
                    o[name] = 1
            if "VELO" in s:
                o["itype"] = 3
            if "MIX" in s:
                o["itype"] = 4
            for key, name in (("MMASS", None), ("RFAKTOR", "rfaktor"), ("LFAKTOR", "fktl"),
                              ("SHIFT", "sh"), ("WIDTH", "width")):
                if key in s:
                    s = line()
                    nums(s)
                    if name == "width":
                        ws[:] = cols[1]
                    elif name:
                        o[name] = cols[1]
                    if name == "fktl" and cols[1] < 1.0e-5:
                        o["nol"] = True
            if "DATAX" in s:
                k = 0
                while True:
                    s = line()
                    if "DATAY" in s:
                        break
                    n = nums(s)
                    if n > 0:
                        k += 1
                        kk = int(cols[1]) if n > 1 else k
                        xs[kk] = cols[2] * o["fau"] + o["sh"]
            if "DATAY" in s:
                k = 0
                while True:
                    s = line()
                    if "DATAW" in s or "END" in s:
                        break
                    n = nums(s)
                    if n > 0:
                        k += 1
                        kk = int(cols[1]) if n > 1 else k
                        ys[kk] = _strength(cols[o["itype"]], o)
            if "DATAW" in s:
                while True:
                    s = line()
                    if nums(s) == 2:
                        ws[int(cols[1])] = cols[2]
            if "DATXY" in s:
                k = 0
                col = (5 if o["uv"] == 0 else 3) + o["itype"] - 2
                while True:
                    s = line()
                    if nums(s) > 0:
                        k += 1
                        xs[k] = cols[2] * o["fau"] + o["sh"]
                        ys[k] = _strength(cols[col], o)
    except StopIteration:
        pass
    x, y, w = xs[1:k + 1].copy(), ys[1:k + 1].copy(), ws[1:k + 1].copy()
    au = o["fau"] >= 1.10
    perm = _selection_order(x if au else -1 * x)
    x, y, w = x[perm], y[perm], w[perm]
    if au and k:
        rel = np.abs(x[1:]) > 1.0e-6
        x[1:][rel] = x[1:][rel] - x[0]
        x[0] = 0.0
    if o["uv"] == 1 and o["forb"] > 1.0e-8:
        y[y < 1.0e-5] = o["forb"]
    xma = max(-1.0e10, float(x.max())) if k else -1.0e10
    xmi = min(1.0e10, float(x.min())) if k else 1.0e10
    dx = xma - xmi
    xma = xma + MARGIN * dx
    xmi = xmi - MARGIN * dx
    if xmi < 0:
        xmi = 0.5
    npts = max(2000, math.ceil(imax / 50.0))
    flags = ((FLAG_UV if o["uv"] else 0) | (FLAG_NM if o["nm"] else 0) |
             (FLAG_CM if o["cm"] else 0) | (FLAG_LOG if o["log"] else 0))
    return GSpecInput(x=x, y=y, wei=w, npts=npts, emin=xmi, step=(xma - xmi) / npts, flags=flags,
                      rfaktor=o["rfaktor"], rots=_rots(x, y, o), opts=o)


def _strength(v: float, o: dict) -> float:
    if o["inv"] == 1:
        v = -1.0 * v
    if o["log"] == 1 and v < 1.0e-6:
        v = 1.0e-4
    return v


def _abscissae(v: float, o: dict) -> list:
    out = []
    if o["nm"] == 0 and o["cm"] == 0:
        out.append(v)
    if o["nm"] == 1:
        out.append(1.0e7 / (EV_CM * v))
    if o["cm"] == 1:
        out.append(EV_CM * v)
    return out


def _rots(x, y, o) -> str:
    """rots.dat (g_spec.f:296-350): scaled lines, or lines merged within ETHR."""
    out = []
    rf = o["rfaktor"]
    if not (o["ethr"] > ETHR_MIN or o["nol"]):
        for xi, yi in zip(x.tolist(), y.tolist()):
            if xi == 0.0:
                continue
            yy = yi * o["fktl"]
            if o["uv"] == 1:
                yy = yy * 1.0e5
            if o["log"] == 1:
                if yy < 0:
                    yy = 1.0e-6
                yy = math.log10(yy) if yy > 0 else -math.inf
            out += [fmt_f(a) + fmt_e(yy * rf) for a in _abscissae(xi, o)]
    if o["ethr"] >= ETHR_MIN:
        n = len(x)
        alive = np.ones(n, dtype=bool)
        for i in range(n):
            if not alive[i]:
                continue
            tot = float(y[i])
            for j in range(i + 1, n):
                if abs(x[i] - x[j]) < o["ethr"]:
                    alive[j] = False
                    tot = tot + float(y[j])
            out += [fmt_f(a) + fmt_e(tot) for a in _abscissae(float(x[i]), o)]
    return "".join(s + "\n" for s in out)


def _spec_text(e: np.ndarray, r: np.ndarray) -> str:
    """spec.dat lines, formatted by the library (stda_hip_format_curve)."""
    from .core import _ptr, load_library
    e = np.ascontiguousarray(e, dtype=np.float64)
    r = np.ascontiguousarray(r, dtype=np.float64)
    buf = np.zeros(31 * len(e), dtype=np.uint8)
    st = load_library().stda_hip_format_curve(len(e), _ptr(e), _ptr(r), _ptr(buf))
    if st != 0:
        raise ValueError("stda_hip_format_curve failed")
    return buf.tobytes().decode("ascii")


# ---------------------------------------------------------------- broadening on the device
def broaden(h, inputs: list) -> list:
    """Curves (e, r) of parsed inputs; one kernel launch per distinct grid size."""
    out = [None] * len(inputs)
    groups: dict = {}
    for q, g in enumerate(inputs):
        groups.setdefault((g.npts, g.flags, g.rfaktor), []).append(q)
    for (npts, flags, rf), idx in groups.items():
        off = np.zeros(len(idx) + 1, dtype=np.int64)
        off[1:] = np.cumsum([len(inputs[q].x) for q in idx])
        cat = lambda attr: (np.concatenate([getattr(inputs[q], attr) for q in idx])
                            if off[-1] else np.zeros(0))
        e, r = h.gspec_curves(off, cat("x"), cat("y"), cat("wei"),
                              np.array([inputs[q].emin for q in idx]),
                              np.array([inputs[q].step for q in idx]), npts, flags, rf)
        for t, q in enumerate(idx):
            out[q] = (e[t], r[t])
    return out


def gspec_many(h, texts: list) -> list:
    """g_spec on many inputs: [(rots.dat text, spec.dat text)], broadening in one launch and
    the spec.dat lines of all inputs formatted in one library call."""
    inputs = [parse_gspec(t) for t in texts]
    curves = broaden(h, inputs)
    if not curves:
        return []
    text = _spec_text(np.concatenate([c[0] for c in curves]),
                      np.concatenate([c[1] for c in curves]))
    out, at = [], 0
    for g, c in zip(inputs, curves):
        n = 31 * len(c[0])
        out.append((g.rots, text[at:at + n]))
        at += n
    return out


def gspec(h, text: str):
    """g_spec on one input: (rots.dat text, spec.dat text)."""
    return gspec_many(h, [text])[0]


def main(argv=None) -> int:
    from .core import StdaHip
    text = sys.stdin.read()
    with StdaHip(0) as h:
        rots, spec = gspec(h, text)
    with open("rots.dat", "w") as f:
        f.write(rots)
    with open("spec.dat", "w") as f:
        f.write(spec)
    return 0


if __name__ == "__main__":
    sys.exit(main())
