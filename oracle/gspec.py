"""CPU restatement of the reference's spectrum program g_spec (TEST INFRASTRUCTURE ONLY).

Imported only by tests/: it is the checker that the spectrum path of the
package (HIP broadening kernel + host front end) is compared against.

g_spec (g_spec/g_spec.f) reads a tda.dat-style file from standard input and
writes two files: rots.dat (the line spectrum) and spec.dat (the Gaussian
broadened curve on a grid of max(2000, ceil(lines/50)) points).  This module
restates the whole program in pure Python loops on IEEE doubles, with the
default-real (single precision) literals of the Fortran source kept as the
float32 values the compiler widens them to.  Pinned by the outputs of the
reference program itself, built from its own source by ``make ref`` in this
directory (``_ref/g_spec``), and by the fixtures under tests/golden/gspec
that were written with it.
"""

from __future__ import annotations

import math

import numpy as np


def _f32(v: float) -> float:
    return float(np.float32(v))


# default-real literals of g_spec.f, as the compiler widens them to real*8
SQRT_PI = _f32(1.7724539)          # g_spec.f:384
F1_CD = _f32(22.97)                # g_spec.f:376
F1_UV = 9.186e-3                   # g_spec.f:377
EV_CM = _f32(8065.54093)           # g_spec.f:325-326, 419-420
MARGIN = _f32(0.40)                # g_spec.f:258-259
DIP = float(np.float32(np.float32(6.4607) * np.float32(1.5)))   # g_spec.f:395 (folded in real*4)
FAU = 27.21139570                  # g_spec.f:76
ETHR_MIN = _f32(0.0001)            # g_spec.f:335, 344


def readaa(a: str, istart: int):
    """One number of a line as READAA parses it (g_spec.f:707-790).

    1-based positions as in the source; returns (value, iend, iend2)."""
    nl = len(a)
    j = istart
    found = False
    while j <= nl - 1:
        n, m = a[j - 1], a[j]
        if n.isdigit() or n == ".":
            found = True
            break
        if n == "-" and (m.isdigit() or m == "."):
            found = True
            break
        j += 1
    if not found:
        return 0.0, 0, 0
    iend = j
    c1 = c2 = 0.0
    one = 1.0
    x = 1.0
    idig = 0
    i = j
    frac = False
    done = False
    while i <= nl:
        n = a[i - 1]
        if n.isdigit():
            idig += 1
            if idig > 10:
                done = True
                break
            c1 = c1 * 10 + int(n)
        elif n == "-" and i == j:
            one = -1.0
        elif n == ".":
            frac = True
            break
        else:
            done = True
            break
        i += 1
    if frac and not done:
        idig = 0
        ii = i + 1
        while ii <= nl:
            n = a[ii - 1]
            if n.isdigit():
                idig += 1
                if idig > 10:
                    break
                c2 = c2 * 10 + int(n)
                x = x / 10
            else:
                break
            ii += 1
    val = one * (c1 + c2 * x)
    iend2 = 0
    jj = iend
    while jj <= nl:
        n = a[jj - 1]
        iend2 = jj
        if n == " ":
            return val, iend, iend2
        if n in "DE":
            c1 = 0.0
            one = 1.0
            q = jj + 1
            while q <= nl:
                n2 = a[q - 1]
                iend2 = q
                if n2 == " ":
                    break
                if n2.isdigit():
                    c1 = c1 * 10.0 + int(n2)
                if n2 == "-":
                    one = -1.0
                q += 1
            return val * 10.0 ** (one * c1), iend, iend2
        jj += 1
    return val, iend, iend2


def readl(line: str, xx: list) -> int:
    """READL (g_spec.f:689-704): numbers of a 120-character line into xx[1..]."""
    a = line[:120].ljust(120)
    i = 0
    start = 1
    while True:
        i += 1
        v, ib, ie = readaa(a, start)
        xx[i] = v
        if ib > 0 and ie > 0:
            start = ie
            continue
        return i - 1


def _special(v: float, w: int):
    """NaN and infinities as the flang runtime writes them."""
    if math.isnan(v):
        return "NaN".rjust(w)
    if math.isinf(v):
        return ("-Inf" if v < 0 else "Inf").rjust(w)
    return None


def fmt_f(v: float, w: int = 16, d: int = 6) -> str:
    """Fortran Fw.d output (asterisks when the value does not fit)."""
    sp = _special(v, w)
    if sp is not None:
        return sp
    t = "%.*f" % (d, v)
    return t.rjust(w) if len(t) <= w else "*" * w


def fmt_e(v: float, w: int = 14, d: int = 4) -> str:
    """Fortran Ew.d output (0.dddd mantissa; three-digit exponents drop the E)."""
    sp = _special(v, w)
    if sp is not None:
        return sp
    neg = math.copysign(1.0, v) < 0
    if v == 0.0:
        digits, ex = "0" * d, 0
    else:
        m, e = ("%.*e" % (d - 1, abs(v))).split("e")
        digits, ex = m.replace(".", ""), int(e) + 1
    tail = "E%+03d" % ex if abs(ex) <= 99 else "%+04d" % ex
    return (("-" if neg else "") + "0." + digits + tail).rjust(w)


def fmt_line(a: float, b: float) -> str:
    """Format 1000 of g_spec.f: (f16.6, e14.4)."""
    return fmt_f(a) + fmt_e(b)


def _dsort(n, lab, y, wei, ew):
    """dsort (g_spec.f:571-609): selection sort on ew ascending, the last of equal minima first."""
    for i in range(n - 1):
        k = i
        pp = ew[i]
        for j in range(i + 1, n):
            if ew[j] > pp:
                continue
            k = j
            pp = ew[j]
        if k == i:
            continue
        ew[k] = ew[i]
        ew[i] = pp
        y[i], y[k] = y[k], y[i]
        wei[i], wei[k] = wei[k], wei[i]
        lab[i], lab[k] = lab[k], lab[i]


def run(text: str):
    """The whole program on the text of its standard input.

    Returns (rots, spec): the contents of rots.dat and spec.dat."""
    lines = text.split("\n")
    if lines and lines[-1] == "":
        lines.pop()
    imax = len(lines)
    jmax = max(2000, math.ceil(imax / 50.0))
    opt = dict(itype=2, iuv=0, icm=0, inm=0, ln=0, ienan=0, fau=1.0, rfaktor=1.0, fktl=1.0,
               sh=0.0, ethr=0.0, forb=0.0, nol=False)
    x = [0.0] * (imax + 1)
    y = [0.0] * (imax + 1)
    wei = [0.0] * (imax + 1)
    lab = [0] * (imax + 1)
    xx = [0.0] * 52
    pos = [0]

    def nxt():
        if pos[0] >= len(lines):
            raise EOFError
        s = lines[pos[0]]
        pos[0] += 1
        return s[:120].ljust(120)          # character*120 a8, read with (a)

    k = 0
    try:
        while True:                                    # label 10, g_spec.f:76-218
            a8 = nxt()
            if "AU " in a8:
                opt["fau"] = FAU
            if "ETHR " in a8:
                a8 = nxt()
                readl(a8, xx)
                opt["ethr"] = xx[1]
            if "FORB " in a8:
                a8 = nxt()
                readl(a8, xx)
                opt["forb"] = xx[1]
            if "UV" in a8:
                opt["iuv"] = 1
            if "ORD" in a8:
                pass                                   # ord.dat is never written (g_spec.f:438)
            if "NM" in a8:
                opt["inm"] = 1
            if "CM" in a8:
                opt["icm"] = 1
            if "VELO" in a8:
                opt["itype"] = 3
            if "MIX" in a8:
                opt["itype"] = 4
            if "LOG" in a8:
                opt["ln"] = 1
            if "INV" in a8:
                opt["ienan"] = 1
            if "MMASS" in a8:
                a8 = nxt()
                readl(a8, xx)
            if "RFAKTOR" in a8:
                a8 = nxt()
                readl(a8, xx)
                opt["rfaktor"] = xx[1]
            if "LFAKTOR" in a8:
                a8 = nxt()
                readl(a8, xx)
                opt["fktl"] = xx[1]
                if opt["fktl"] < 1.0e-5:
                    opt["nol"] = True
            if "SHIFT" in a8:
                a8 = nxt()
                readl(a8, xx)
                opt["sh"] = xx[1]
            if "WIDTH" in a8:
                a8 = nxt()
                readl(a8, xx)
                for q in range(1, imax + 1):
                    wei[q] = xx[1]
            if "DATAX" in a8:
                k = 0
                while True:
                    a8 = nxt()
                    if "DATAY" in a8:
                        break
                    n = readl(a8, xx)
                    if n > 0 and k <= imax:
                        k += 1
                        kk = int(xx[1]) if n > 1 else k
                        lab[kk] = kk
                        x[kk] = xx[2] * opt["fau"] + opt["sh"]
            if "DATAY" in a8:
                k = 0
                while True:
                    a8 = nxt()
                    if "DATAW" in a8 or "END" in a8:
                        break
                    n = readl(a8, xx)
                    if n > 0 and k <= imax:
                        k += 1
                        kk = int(xx[1]) if n > 1 else k
                        y[kk] = xx[opt["itype"]]
                        if opt["ienan"] == 1:
                            y[kk] = -1.0 * y[kk]
                        if opt["ln"] == 1 and y[kk] < 1.0e-6:
                            y[kk] = 1.0e-4
            if "DATAW" in a8:
                while True:
                    a8 = nxt()
                    n = readl(a8, xx)
                    if n == 2 and k <= imax:
                        wei[int(xx[1])] = xx[2]
            if "DATXY" in a8:
                k = 0
                while True:
                    a8 = nxt()
                    n = readl(a8, xx)
                    if n > 0 and k <= imax:
                        k += 1
                        lab[k] = k
                        x[k] = xx[2] * opt["fau"] + opt["sh"]
                        col = (5 + opt["itype"] - 2) if opt["iuv"] == 0 else (3 + opt["itype"] - 2)
                        y[k] = xx[col]
                        if opt["ienan"] == 1:
                            y[k] = -1.0 * y[k]
                        if opt["ln"] == 1 and y[k] < 1.0e-6:
                            y[k] = 1.0e-4
    except EOFError:
        pass
    nroots = k
    fau = opt["fau"]
    # sort (g_spec.f:222-247), 0-based from here on
    X = x[1:nroots + 1]
    Y = y[1:nroots + 1]
    W = wei[1:nroots + 1]
    L = lab[1:nroots + 1]
    if fau < 1.10:
        X = [-1 * v for v in X]
    _dsort(nroots, L, Y, W, X)
    if fau < 1.10:
        X = [-1 * v for v in X]
    if fau > 1.10:
        for i in range(1, nroots):
            if abs(X[i]) > 1.0e-6:
                X[i] = X[i] - X[0]
        if nroots:
            X[0] = 0.0
    xma, xmi = -1.0e10, 1.0e10
    for i in range(nroots):
        if X[i] > xma:
            xma = X[i]
        if X[i] < xmi:
            xmi = X[i]
        if opt["iuv"] == 1 and opt["forb"] > 1.0e-8 and Y[i] < 1.0e-5:
            Y[i] = opt["forb"]
    dx = xma - xmi
    xma = xma + MARGIN * dx
    xmi = xmi - MARGIN * dx
    if xmi < 0:
        xmi = 0.5
    # line spectrum (g_spec.f:296-350)
    rots = []
    inm, icm, rf = opt["inm"], opt["icm"], opt["rfaktor"]
    for i in range(nroots):
        yy = Y[i]
        xnm = 1.0e7 / (EV_CM * X[i]) if X[i] != 0.0 else math.inf
        xcm = EV_CM * X[i]
        if X[i] != 0.0:
            yy = yy * opt["fktl"]
            if opt["iuv"] == 1:
                yy = yy * 1.0e5
            if opt["ethr"] > ETHR_MIN or opt["nol"]:
                continue
            if opt["ln"] == 1:
                if yy < 0:
                    yy = 1.0e-6
                yy = math.log10(yy) if yy > 0 else -math.inf
            if inm == 0 and icm == 0:
                rots.append(fmt_line(X[i], yy * rf))
            if inm == 1:
                rots.append(fmt_line(xnm, yy * rf))
            if icm == 1:
                rots.append(fmt_line(xcm, yy * rf))
    if opt["ethr"] >= ETHR_MIN:
        iff = [1] * nroots
        for i in range(nroots):
            s = Y[i]
            if iff[i] > 0:
                for j in range(i + 1, nroots):
                    if abs(X[i] - X[j]) < opt["ethr"]:
                        iff[j] = 0
                        s = s + Y[j]
                if inm == 0 and icm == 0:
                    rots.append(fmt_line(X[i], s))
                if inm == 1:
                    rots.append(fmt_line(1.0e7 / (EV_CM * X[i]), s))
                if icm == 1:
                    rots.append(fmt_line(EV_CM * X[i], s))
    # Gaussian curves (g_spec.f:371-441)
    f1 = F1_UV if opt["iuv"] == 1 else F1_CD
    npts = jmax
    st = (xma - xmi) / npts
    spec = []
    for j in range(1, npts + 1):
        ee = xmi + j * st
        r = 0.0
        for i in range(nroots):
            if abs(X[i]) > 1.0e-6:
                f3 = 1.0 / (f1 * SQRT_PI * W[i])
                width = 1.0 / (W[i] * W[i])
                edif = ee - X[i]
                if opt["iuv"] == 0:
                    f2 = X[i] * Y[i]
                else:
                    dip2 = DIP * Y[i] / (3.67493e-2 * X[i])
                    f2 = X[i] * dip2
                r = r + f3 * f2 * math.exp(-width * (edif * edif))
        r = r * rf
        e = ee
        if inm == 1:
            with np.errstate(divide="ignore"):
                e = float(np.float64(1.0e7) / np.float64(EV_CM * e))
        if icm == 1:
            e = EV_CM * e
        if opt["ln"] == 1:
            with np.errstate(all="ignore"):
                v = float(np.log10(r))
            if v < 0:
                v = 0.0
            spec.append(fmt_line(e, v))
        else:
            spec.append(fmt_line(e, r))
    return "".join(s + "\n" for s in rots), "".join(s + "\n" for s in spec)
