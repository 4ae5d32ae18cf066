"""numpy restatement of the library's response-solver loop (test infrastructure).

Mirrors ``stda_hip_lresp`` in ``stda_amd/csrc/core.cpp`` step for step, with
dense operators in place of the device's matrix-free ones, so that its
convergence behaviour (subspace size, restarts, exits) can be checked on the
CPU against the oracle's dense solve of the reference's equations
((A+B) - w^2 (A-B)^-1)(X+Y) = -2 mu (linear_response.f:90-131).

The model is the FP64 loop; the library's mixed-precision mode runs the same loop with
the FP32 operator first and switches to this FP64 loop (the space's operator images
recomputed in FP64) once its residuals stop falling.

``lresp_model(apply, ed, mu, omegas, maxdim)`` returns ``(alpha, status, info)``
with ``alpha`` (nfreq, 3, 3) as alpha[f, r, c] = -2 mu_r^T (X+Y)_c, ``status``
"ok" or "not converged", ``info`` a dict (iterations, restarts, final size).
"""

import numpy as np


def _orthonormalize(V, m, T):
    """Project T off V[:, :m] twice, symmetric orthonormalisation, again, and drop the
    directions that lost more than ~30 % of their norm (Gram eigenvalue below 0.5)."""
    T = T / np.linalg.norm(T, axis=0)
    for _ in range(2):
        if m:
            T = T - V[:, :m] @ (V[:, :m].T @ T)
    w, U = np.linalg.eigh(T.T @ T)
    keep = w > 1e-14 * max(1.0, w.max())
    T2 = T @ (U[:, keep] / np.sqrt(w[keep]))
    for _ in range(2):
        if m:
            T2 = T2 - V[:, :m] @ (V[:, :m].T @ T2)
    w, U = np.linalg.eigh(T2.T @ T2)
    keep = w > 0.5
    return T2 @ (U[:, keep] / np.sqrt(w[keep]))


def lresp_model(apply, ed, mu, omegas, maxdim, tol=1e-10, maxit=500):
    n = len(ed)
    nfreq = len(omegas)
    nrhs = 3 * nfreq
    rhs = -2.0 * mu
    wmax = max(abs(w) for w in omegas)
    maxdim = min(n, max(maxdim, 3 + 4 * nrhs))
    V = np.zeros((n, maxdim))
    P = np.zeros((n, maxdim))
    M = np.zeros((n, maxdim))
    T = _orthonormalize(V, 0, rhs / ed[:, None])
    m0 = mnew = T.shape[1]
    V[:, :mnew] = T
    m = 0
    rn0 = np.linalg.norm(rhs, axis=0)
    wc = np.repeat(np.asarray(omegas, dtype=float), 3)
    info = dict(iterations=0, restarts=0, size=0)
    A = None
    for it in range(maxit):
        P[:, m:m + mnew], M[:, m:m + mnew] = apply(V[:, m:m + mnew])
        m += mnew
        ap = V[:, :m].T @ P[:, :m]
        ap = 0.5 * (ap + ap.T)
        am = V[:, :m].T @ M[:, :m]
        am = 0.5 * (am + am.T)
        br = V[:, :m].T @ rhs
        L = np.linalg.cholesky(am)
        S = L.T @ ap @ L
        lam, Z = np.linalg.eigh(0.5 * (S + S.T))
        W = L @ Z
        Yz = np.linalg.solve(L.T, Z)
        G = W.T @ br
        A = np.zeros((m, nrhs))
        B = np.zeros((m, nrhs))
        for f, w in enumerate(omegas):
            c = G / (lam - w * w)[:, None]
            A[:, 3 * f:3 * f + 3] = W @ c
            B[:, 3 * f:3 * f + 3] = w * (Yz @ c)
        R1 = P[:, :m] @ A - (V[:, :m] @ B) * wc - np.tile(rhs, nfreq)
        R2 = M[:, :m] @ B - (V[:, :m] @ A) * wc
        rn = np.r_[np.linalg.norm(R1, axis=0), np.linalg.norm(R2, axis=0)]
        lim = tol * np.maximum(1.0, np.tile(rn0, 2 * nfreq))
        un = np.nonzero(~(rn < lim))[0]
        info.update(iterations=it + 1, size=m)
        if len(un) == 0 or m == n:
            return _alpha(mu, V[:, :m] @ A, nfreq), "ok", info
        dd = ed[:, None]
        ww = wc[None, :]
        det = dd * dd - ww * ww
        det = np.where(np.abs(det) < 1e-6, np.where(det < 0, -1e-6, 1e-6), det)
        R = np.hstack([(dd * R1 + ww * R2) / det, (ww * R1 + dd * R2) / det])[:, un]
        if m + R.shape[1] > maxdim:
            cols = [np.eye(m)[:, :m0], A, B]
            room = max(0, (maxdim // 2 - m0 - 2 * nrhs) // 2)
            low = [q for q in range(m) if 0 < lam[q] <= 1.5625 * wmax * wmax][:room] if wmax > 0 else []
            for q in low:
                cols += [W[:, q:q + 1], Yz[:, q:q + 1]]
            Q = np.zeros((m, 0))
            for x in np.hstack(cols).T:
                nx0 = np.linalg.norm(x)
                if not nx0 > 0:
                    continue
                x = x / nx0
                for _ in range(2):
                    x = x - Q @ (Q.T @ x)
                if np.linalg.norm(x) > 1e-8:
                    Q = np.c_[Q, x / np.linalg.norm(x)]
            r = Q.shape[1]
            V[:, :r], P[:, :r], M[:, :r] = V[:, :m] @ Q, P[:, :m] @ Q, M[:, :m] @ Q
            m, m0 = r, min(m0, r)
            info["restarts"] += 1
            R = R[:, :max(0, maxdim - m)]
            if R.shape[1] == 0:
                return None, "not converged", info
        T = _orthonormalize(V, m, R)
        mnew = T.shape[1]
        if mnew == 0:
            return None, "not converged", info
        V[:, m:m + mnew] = T
    return None, "not converged", info


def _alpha(mu, xpy, nfreq):
    return np.stack([-2.0 * mu.T @ xpy[:, 3 * f:3 * f + 3] for f in range(nfreq)])
