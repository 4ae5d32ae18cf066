"""Continuous-batching service for many excited-state jobs on one GPU.

Jobs are sTDA spectrum requests (lowest ``nroots`` roots with transition
properties) or sTD-DFT polarizability requests (alpha(w) at a list of
frequencies).  A :class:`Server` holds

* one batch engine (:class:`stda_amd.core.StdaHipBatch`, ``batch_workers``
  device contexts) in its continuous mode: a spectrum job is submitted the
  moment it arrives (and fits, below), its setup runs on the next free worker
  context, and it joins the running lock-step solver at the solver's next pass
  -- requests are admitted while others are on the GPU, not at batch
  boundaries -- or, when larger, runs the single-molecule mixed path on that
  worker; a collector thread returns each result as soon as it is done;
* ``workers`` contexts for polarizability jobs, each job on its own context and
  host thread, overlapping with everything else on the device.

Admission is by HBM footprint: a job enters only when its upper-bound device
footprint fits in what the running jobs leave of ``hbm_budget``; it waits in
arrival order otherwise.  All device state of a job (charges, gamma, CSF
space, matrices, Davidson subspaces) stays resident from setup to result.  The
C calls release the interpreter lock, so the engine, the polarizability
workers and the device overlap.

Every kernel of a job touches only that job's data and reduces in a fixed
order, so a job's result does not depend on what else is running; the GPU
tests check this (alone vs. in the mix).
"""

from __future__ import annotations

import math
import threading
import time
from collections import deque
from dataclasses import dataclass

import numpy as np

from .core import StdaHip, StdaHipBatch
from .prep import CoreInputs


@dataclass
class Job:
    jid: int
    kind: str                     # "spectrum" or "polar"
    ci: CoreInputs
    thr: float = 7.0              # Ethr (eV)
    ax: float = 0.25
    nroots: int = 20
    omegas: tuple = (0.0,)        # polarizability frequencies (Hartree)
    arrival: float = 0.0          # seconds after the start of the run
    start: float = math.nan
    end: float = math.nan
    result: object = None         # spectrum: (omega, props); polar: [(w, alpha 3x3)]
    error: str = ""

    @property
    def latency(self) -> float:
        return self.end - self.arrival


def footprint_bytes(job: Job) -> int:
    """Upper bound of the device memory one job needs (bytes).

    Charges and contracted charges over all pairs (q_ia, p_ia, packed
    occupied/virtual pair charges), the explicit FP64/FP32 A and 0.5 B of the
    batched solver when the CSF space can be small enough for it, the Davidson
    or response subspace for the largest possible CSF space (every occupied x
    virtual pair) and the work buffers of one block of operator applications.
    """
    ci = job.ci
    na, nao, no, nv = ci.ncent, ci.nao, ci.no, ci.nv
    P = no * nv
    charges = 8 * na * (2 * P + no * (no + 1) + nv * (nv + 1) // 2 + no + nv)
    if job.kind == "spectrum":
        k = job.nroots
        maxdim = max(6 * k, k + 64)
    else:
        k = 3 * len(job.omegas)
        maxdim = min(P, 2400)    # the response subspace (at most 2400 vectors)
    dav = 8 * P * 3 * min(maxdim, P)
    nsmall = min(P, 4096)
    explicit = 20 * nsmall * nsmall
    per_vec = 8 * (2 * P + 2 * nao * no + 2 * nao * nao + na)
    work = per_vec * min(64, max(1, min(maxdim, P)))
    return int(charges + dav + explicit + work + 64 * 2 ** 20)


@dataclass
class RunStats:
    n: int
    ok: int
    wall: float
    solves_per_s: float
    p50_ms: float
    p99_ms: float
    mean_ms: float
    submitted: int = 0            # spectrum jobs submitted to the continuous engine


class Server:
    """Runs jobs with continuous batching (see the module documentation)."""

    def __init__(self, workers: int = 4, hbm_budget: float = 160e9, precision: str = "mixed",
                 device: int = 0, tol: float = 1e-8, batch_workers: int = 8):
        self.ctx = [StdaHip(device) for _ in range(workers)]
        for h in self.ctx:
            h.set_precision(precision)
        self.engine = StdaHipBatch(device, batch_workers)
        self.budget = float(hbm_budget)
        self.tol = tol
        self._cv = threading.Condition()
        self._used = 0.0

    def close(self):
        for h in self.ctx:
            h.close()
        self.engine.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # admission by footprint (called with self._cv held)
    def _fits(self, fp: float) -> bool:
        return self._used + fp <= self.budget or self._used == 0.0

    def run(self, jobs: list, realtime: bool = True, progress=None) -> RunStats:
        """Serve ``jobs``.  With ``realtime`` the arrival times are honoured
        (open loop); otherwise every job is queued at t = 0 (closed batch).
        ``progress(done, total)`` is called about every 10 s while jobs run."""
        jobs = sorted(jobs, key=lambda j: j.arrival)
        if not realtime:
            for j in jobs:
                j.arrival = 0.0
        t0 = time.perf_counter()
        now = lambda: time.perf_counter() - t0
        spec_wait: deque = deque()       # arrived spectrum jobs that do not fit yet
        polar_q: deque = deque()
        free_ctx = list(self.ctx)
        outstanding = {}                 # ticket -> (job, footprint)
        done_feeding = threading.Event()
        threads = []
        stats = {"submitted": 0}

        def fp_of(j):
            return min(float(footprint_bytes(j)), self.budget)

        def admit_spectra():             # with self._cv held
            while spec_wait and self._fits(fp_of(spec_wait[0])):
                j = spec_wait.popleft()
                fp = fp_of(j)
                self._used += fp
                j.start = now()
                t = self.engine.submit(j.ci, j.thr, j.ax, j.nroots, tol=self.tol)
                outstanding[t] = (j, fp)
                stats["submitted"] += 1

        def collector():
            while True:
                with self._cv:
                    tickets = list(outstanding)
                    if not tickets and not spec_wait and done_feeding.is_set():
                        return
                finished = []
                for t in tickets:
                    r = self.engine.result(t, wait=False)
                    if r is not None:
                        finished.append((t, r))
                with self._cv:
                    for t, (om, pr, st) in finished:
                        j, fp = outstanding.pop(t)
                        j.end = now()
                        if st == 0:
                            j.result = (om, pr)
                        else:
                            j.error = f"status {st}"
                        self._used -= fp
                    if finished:
                        admit_spectra()
                        self._cv.notify_all()
                    else:
                        self._cv.wait(0.0005)

        def polar_job(h: StdaHip, j: Job, fp: float):
            j.start = now()
            try:
                h.setup(j.ci, j.thr, j.ax)
                j.result = h.lresp(list(j.omegas), tol=self.tol)
            except Exception as e:
                j.error = str(e)
            j.end = now()
            with self._cv:
                self._used -= fp
                free_ctx.append(h)
                admit_spectra()
                self._cv.notify_all()

        def polar_loop():
            while True:
                with self._cv:
                    while True:
                        if polar_q and free_ctx and self._fits(fp_of(polar_q[0])):
                            break
                        if done_feeding.is_set() and not polar_q:
                            return
                        self._cv.wait(0.05)
                    j = polar_q.popleft()
                    fp = fp_of(j)
                    self._used += fp
                    h = free_ctx.pop()
                t = threading.Thread(target=polar_job, args=(h, j, fp), daemon=True)
                t.start()
                threads.append(t)

        loops = [threading.Thread(target=collector, daemon=True),
                 threading.Thread(target=polar_loop, daemon=True)]
        for t in loops:
            t.start()
        for j in jobs:                      # the arrival process
            if realtime:
                delay = j.arrival - now()
                if delay > 0:
                    time.sleep(delay)
            with self._cv:
                if j.kind == "spectrum":
                    spec_wait.append(j)
                    admit_spectra()
                else:
                    polar_q.append(j)
                self._cv.notify_all()
        done_feeding.set()
        with self._cv:
            self._cv.notify_all()
        last = [now()]

        def wait(t):
            while t.is_alive():
                t.join(timeout=1.0)
                if progress is not None and now() - last[0] >= 10.0:
                    last[0] = now()
                    progress(sum(1 for j in jobs if not math.isnan(j.end)), len(jobs))

        for t in loops:
            wait(t)
        for t in list(threads):     # every polarizability job has been started
            wait(t)
        self.engine.stop()
        wall = now()
        lat = np.array([j.latency for j in jobs]) * 1e3 if jobs else np.zeros(1)
        ok = sum(1 for j in jobs if not j.error)
        return RunStats(n=len(jobs), ok=ok, wall=wall, solves_per_s=ok / wall if wall > 0 else 0.0,
                        p50_ms=float(np.percentile(lat, 50)), p99_ms=float(np.percentile(lat, 99)),
                        mean_ms=float(lat.mean()), submitted=stats["submitted"])


def request_mix(n: int, seed: int = 2000, lo: int = 10, hi: int = 400,
                polar_fraction: float = 0.3):
    """Atom counts (log-uniform in [lo, hi], :func:`stda_amd.synth.mixed_sizes`) and
    job kinds ("spectrum" / "polar") of a serving stream of ``n`` requests; request i
    is the molecule ``synth.cluster(sizes[i], seed=seed + i, basis="minimal")``.  The
    kinds come from a stream of their own, independent of the sizes."""
    from .synth import mixed_sizes
    sizes = mixed_sizes(n, lo, hi, seed)
    u = np.random.Generator(np.random.PCG64([seed, 1])).uniform(size=n)
    kinds = np.where(u < 1.0 - polar_fraction, "spectrum", "polar")
    return sizes, kinds


def poisson_arrivals(n: int, rate: float, seed: int = 0) -> np.ndarray:
    """Arrival times (s) of an open-loop Poisson stream of ``rate`` requests/s."""
    rng = np.random.Generator(np.random.PCG64(seed))
    return np.cumsum(rng.exponential(1.0 / rate, size=n))
