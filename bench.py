"""Benchmark of the MI355X excited-state core.

Default workload ``mixed512``: 512 independent synthetic molecules, atom
counts log-uniform in [10, 200] (C:H:N:O = 45:45:5:5 clusters, minimal s/p
basis, half the electrons occupied, seeds 1000 + i), Ethr = 7 eV, a_x = 0.25,
the 20 lowest sTDA roots of each with transition properties and the
velocity correction.  One step = one batched mixed-precision solve of all
512 molecules (``StdaHipBatch``: setup of every molecule on 16 concurrent
device contexts, the molecules with at most 4096 CSFs solved in lock step
by the batched kernels, the larger ones by the single-molecule mixed path
of their worker).  The
plain FP64 GPU path (one molecule per solve, sequentially, as the drop-in
replaces the reference routine) is timed once over the same 512 molecules;
every batched root is compared with it (<= 1 meV, |df| <= 1e-4).  The
molecules' boundary arrays are generated on the host before timing; the
timed region starts from them (host arrays, copied by each solve).

Other workloads: ``cluster400`` (400-atom split-valence cluster, 800
occupied, Ethr = 10 eV, 50 roots, plain FP64 single-molecule drop-in, step =
one solve with the inputs already in HBM), ``cluster100``, ``polyene``
(C14H16, the CPU-runnable case) and ``serve`` (open-loop Poisson stream of
70% spectrum / 30% polarizability jobs, sizes log-uniform in [10, 400]).

Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
# the batch workers drive one stream each (plus one for their batched
# launches); give the process enough hardware queues for them to overlap
os.environ.setdefault("GPU_MAX_HW_QUEUES", "32")

FP64_PEAK_TFLOPS = 78.6      # MI355X dense FP64 (vector and matrix)
FP32_PEAK_TFLOPS = 157.3     # MI355X dense FP32 matrix
# HBM bytes per launch of PT2 pass 1 over the mixed512 molecules in the plain
# pass: FETCH_SIZE (x2) + WRITE_SIZE (x1) counters, scaled by a 1 GiB copy
# (profiles/pt2_pipe_mixed512_pmc.md); the writes are the pass-1 terms kept
# for pass 2
PT2_MIXED_TRAFFIC = 9.34e8
HBM_PEAK_GBS = 8000.0
# HIP-event timed statistics of the library: single kernels (one launch per
# record: the PT2 row pass (FP64; FP32 with the FP64 recheck of row tiles
# near the threshold in mixed precision), the lock-step operator kernel, the slice partials
# of the tall-skinny V^T W products and the V C products) and composite spans
# (the AO operator products, several GEMMs and element-wise launches; PT2
# pass 2 with its finishing kernel)
SINGLE_KERNELS = ("pt2_rows", "pt2_rows_f32", "pt2_rows_recheck", "batch_lockstep", "tsgemm_tn",
                  "tsgemm_nn")
HBM_KERNELS = ("batch_lockstep", "tsgemm_tn", "tsgemm_nn")
COMPOSITE_SPANS = ("pt2_cols", "ao_matvec", "ao_matvec_f32")
EV = 27.211385050
METRIC = "excited-state solves/s (batched mixed vs plain FP64) at <=1 meV; p50/p99 latency"


def single_workload(name: str):
    from stda_amd import synth
    from stda_amd.prep import prepare
    if name == "cluster400":
        mol, ints = synth.cluster(400, seed=1, basis="split", nocc=800)
        p = dict(thr=10.0, ax=0.25, nroots=50)
    elif name == "cluster100":
        mol, ints = synth.cluster(100, seed=1, basis="split", nocc=200)
        p = dict(thr=10.0, ax=0.25, nroots=50)
    elif name == "polyene":
        mol, ints = synth.polyene()
        p = dict(thr=7.0, ax=0.25, nroots=10)
    else:
        raise SystemExit(f"unknown workload {name}")
    ci = prepare(mol, p["thr"], p["ax"], ints)
    return ci, p


GEMM_PROBE_SHAPES = [  # (M, N, K, transA, transB): the operator GEMMs of the 400-atom cluster
    (2340, 12800, 1540, 0, 0),   # T = Cv X^T for 16 vectors (nao x no*16, K = nv)
    (2340, 2340, 800, 0, 1),     # D = T Co^T per vector
    (2340, 800, 2340, 0, 0),     # W = G Co per vector
    (1540, 12800, 2340, 1, 0),   # R = Cv^T W for 16 vectors
    (640, 6400, 250, 0, 0),      # the same transforms of a 200-atom minimal-basis molecule
    (640, 640, 390, 0, 1),
]


def gemm_probe():
    """The operator-shape GEMMs through the library's matrix-core kernel (FP64, FP32,
    split-bfloat16; three calls each, host arrays) and, for comparison, through torch's
    vendor GEMMs on device tensors (FP64, FP32).  Run under rocprofv3 to read each
    kernel's duration; prints the shapes in order."""
    import torch
    from stda_amd.core import StdaHip
    rng = np.random.default_rng(0)
    with StdaHip(0) as h:
        for (m, n, k, ta, tb) in GEMM_PROBE_SHAPES:
            A = rng.standard_normal((k, m) if ta else (m, k))
            B = rng.standard_normal((n, k) if tb else (k, n))
            for mode in ("fp64", "fp32", "bf16"):
                for _ in range(3):
                    h.gemm(A, B, ta=bool(ta), tb=bool(tb), mode=mode)
            tA, tB = torch.tensor(A, device="cuda"), torch.tensor(B, device="cuda")
            oA, oB = (tA.t() if ta else tA), (tB.t() if tb else tB)
            for dt in (torch.float64, torch.float32):
                a, b = oA.to(dt), oB.to(dt)
                for _ in range(3):
                    a @ b
                torch.cuda.synchronize()
            print(f"gemm probe {m}x{n}x{k} ta={ta} tb={tb}", flush=True)


def copy_probe(nbytes: int):
    """One device-to-device copy of nbytes (a kernel of known HBM traffic:
    nbytes read + nbytes written), used to scale the TCC byte counters."""
    import torch
    n = nbytes // 8
    a = torch.ones(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes": n * 8}), flush=True)


def host_info():
    """CPU model and the CPUs this process may use (the batched line depends on them)."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        ncpu = len(os.sched_getaffinity(0))
    except AttributeError:
        ncpu = os.cpu_count()
    return {"cpu_model": model, "usable_cpus": ncpu}


def gen_workers():
    return max(1, min(16, os.cpu_count() or 1))


def log(msg):
    print(msg, file=sys.stderr, flush=True)


# --------------------------------------------------------------------------- single molecule
def run_single(args):
    t0 = time.perf_counter()
    ci, p = single_workload(args.workload)
    log(f"# workload {args.workload}: natoms={ci.ncent} nao={ci.nao} no={ci.no} nv={ci.nv} "
        f"(inputs {time.perf_counter() - t0:.1f}s)")
    from stda_amd.core import StdaHip
    h = StdaHip(0)
    h.load(ci, p["thr"], p["ax"])
    h.set_precision(args.precision)
    for w in range(args.warmup):
        h.select()
        h.solve_tda(p["nroots"], velcorr=True, vectors=False)
        log(f"# warmup {w + 1}/{args.warmup} done")
    lat, stats, hbm = [], [], []
    tstart = time.perf_counter()
    for s in range(args.steps):
        ts = time.perf_counter()
        nci, nprim, nsec = h.select()
        sol = h.solve_tda(p["nroots"], velcorr=True, vectors=False)
        lat.append(time.perf_counter() - ts)
        stats.append({k: h.kernel_stats(k) for k in ("pt2_rows", "pt2_rows_f32",
                                                       "pt2_rows_recheck", "pt2_cols",
                                                       "ao_matvec", "ao_matvec_f32")})
        hbm.append({k: h.kernel_stats(k) for k in ("tsgemm_tn", "tsgemm_nn")})
        log(f"# step {s + 1}/{args.steps}: {lat[-1]:.3f}s nci={nci} timings={sol.timings_ms}")
    total = time.perf_counter() - tstart
    root_err = None
    if args.precision != "fp64":
        h.set_precision("fp64")
        ref = h.solve_tda(p["nroots"], velcorr=True, vectors=False)
        root_err = {"max_abs_dE_eV": float(np.abs(sol.omega - ref.omega).max() * EV),
                    "max_abs_dfL": float(np.abs(sol.props[:, 0] - ref.props[:, 0]).max())}
    h.close()

    last = stats[-1]
    dom = max(last, key=lambda k: last[k]["ms"])
    st = last[dom]
    achieved = st["work"] / (st["ms"] * 1e-3) / 1e12 if st["ms"] > 0 else 0.0
    # the FP32 kernels against the FP32 matrix peak, the others against FP64
    peak = FP32_PEAK_TFLOPS if dom in ("pt2_rows_f32", "ao_matvec_f32") else FP64_PEAK_TFLOPS
    roofline = {"kernel": dom, "bound": "mfma", "achieved": round(achieved, 3),
                "peak": peak, "unit": "TFLOP/s",
                "frac": round(achieved / peak, 4), "traffic": None,
                "avg_launch_ms": round(st["ms"] / max(st["launches"], 1), 3),
                "work_per_launch": st["work"] / max(st["launches"], 1)}
    prof = os.path.join(ROOT, "profiles", "pt2_pipe_cluster400_pmc.md")
    if dom == "pt2_rows" and args.workload == "cluster400" and os.path.exists(prof):
        for ln in open(prof):
            if ln.startswith("traffic_bytes_per_launch:"):
                roofline["traffic"] = float(ln.split(":")[1])
    cpu = None
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle
        secs, desc, threads = oracle.cpu_baseline_sample(ci, p["thr"], p["ax"], p["nroots"],
                                                         budget_s=args.cpu_budget)
        cpu = {"value": 1.0 / secs, "unit": "solves/s", "cores": threads, "kind": "port",
               "sample": desc}
    lat_ms = np.array(lat) * 1e3
    return {
        "metric": METRIC,
        "value": args.steps / total,
        "unit": "solves/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": total / args.steps * 1e3,
        "p50_ms": float(np.percentile(lat_ms, 50)), "p99_ms": float(np.percentile(lat_ms, 99)),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32 contractions + f64 Davidson",
        "precision": args.precision, "root_error_vs_fp64": root_err, "data": "synthetic",
        "config": {"workload": f"{args.workload}: single-molecule sTDA drop-in, "
                               f"{p['nroots']} lowest roots",
                   "natoms": ci.ncent, "nao": ci.nao, "no": ci.no, "nv": ci.nv,
                   "nci": int(nci), "nprim": int(nprim), "nsec": int(nsec),
                   "ethr_ev": p["thr"], "ax": p["ax"], "nroots": p["nroots"],
                   "davidson_tol": 1e-8},
        "roofline": roofline,
        "kernels": {k: {"ms": round(v["ms"], 3), "launches": v["launches"],
                        "tflops": round(v["work"] / (v["ms"] * 1e-3) / 1e12, 3) if v["ms"] else 0}
                    for k, v in last.items() if v["launches"]},
        # the solver's tall-skinny products (HBM-bound; one launch in eight timed)
        "hbm_kernels": plain_rates(hbm[-1]),
        "cpu_baseline": cpu,
    }


# --------------------------------------------------------------------------- batched
PLAIN_KERNELS = ("tsgemm_tn", "tsgemm_nn", "pt2_rows")


def plain_fp64(cis, thr, ax, nroots, kst=None):
    """Sequential plain FP64 solves, one molecule per solve (the drop-in path).
    kst (dict): HIP-event statistics of the tall-skinny products summed over
    the molecules (one stream, so an event pair brackets the kernel alone)."""
    from stda_amd.core import StdaHip, StdaHipError
    out, lat = [], []
    with StdaHip(0) as h:
        h.set_precision("fp64")
        t0 = time.perf_counter()
        for ci in cis:
            ts = time.perf_counter()
            try:
                h.setup(ci, thr, ax)
                s = h.solve_tda(nroots, velcorr=True, tol=1e-8, vectors=False)
                out.append((s.omega.copy(), s.props.copy()))
                if kst is not None:
                    for k in PLAIN_KERNELS:
                        st = h.kernel_stats(k)
                        acc = kst.setdefault(k, {"ms": 0.0, "launches": 0.0, "work": 0.0})
                        for f in acc:
                            acc[f] += st[f]
            except StdaHipError as e:
                out.append(e.status)
            lat.append(time.perf_counter() - ts)
        wall = time.perf_counter() - t0
    return out, wall, np.array(lat)


def plain_rates(kst):
    """Per-launch HIP-event time and achieved rate of the kernels timed in the
    sequential plain pass: the tall-skinny products (HBM, algorithmic bytes
    8 n (m + p)) and PT2 pass 1 (FP64 MFMA, 4 Na Nsec Nprim flops)."""
    out = {}
    for k, st in kst.items():
        if st["ms"] <= 0:
            continue
        d = {"ms": round(st["ms"], 2), "launches": int(st["launches"]),
             "avg_launch_ms": round(st["ms"] / max(st["launches"], 1.0), 4)}
        if k in HBM_KERNELS:
            gbs = st["work"] / (st["ms"] * 1e-3) / 1e9
            d.update(achieved_GBps=round(gbs, 1), frac=round(gbs / HBM_PEAK_GBS, 4))
        else:
            tf = st["work"] / (st["ms"] * 1e-3) / 1e12
            d.update(achieved_TFLOPs=round(tf, 3), frac=round(tf / FP64_PEAK_TFLOPS, 4))
        out[k] = d
    return out


def compare(res, ref):
    dE = dF = 0.0
    mism = 0
    for (om, pr, st), r in zip(res, ref):
        if isinstance(r, int):
            mism += st != r
            continue
        if st != 0 or len(om) != len(r[0]):
            mism += 1
            continue
        dE = max(dE, float(np.abs(om - r[0]).max()) * EV)
        dF = max(dF, float(np.abs(pr[:, :2] - r[1][:, :2]).max()))
    return {"max_abs_dE_eV": dE, "max_abs_df": dF, "mismatched_molecules": mism}


def cpu_baseline_batch(cis, thr, ax, nroots, budget_s):
    """Oracle restatement on a seeded sample of 8 molecules (with CSFs),
    extrapolated to the mean seconds per solve of the whole set."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    rng = np.random.default_rng(7)
    idx = rng.choice(len(cis), size=min(8, len(cis)), replace=False)
    secs, descs, threads = [], [], 1
    for i in idx:
        try:
            s, d, threads = oracle.cpu_baseline_sample(cis[i], thr, ax, nroots,
                                                       budget_s=budget_s / len(idx))
        except Exception:
            continue          # no CSF: nothing to time
        secs.append(s)
        descs.append(f"mol {int(i)} (natoms {cis[i].ncent}): {s:.3g} s")
    mean = float(np.mean(secs)) if secs else float("nan")
    return {"value": 1.0 / mean if secs else None, "unit": "solves/s", "cores": threads,
            "kind": "port",
            "sample": (f"oracle restatement (phases timed in full or on row samples and "
                       f"extrapolated, as for one molecule) on {len(secs)} seeded molecules of "
                       f"the set; mean seconds per solve {mean:.3g}; " + "; ".join(descs))}


def run_mixed(args):
    from stda_amd import synth
    from stda_amd.core import StdaHipBatch
    n, thr, ax, nroots = args.nmol, 7.0, 0.25, 20
    t0 = time.perf_counter()
    cis = None
    cache = args.inputs_cache
    if cache and os.path.exists(cache):
        # written by this script (below); the inputs are plain arrays
        import pickle
        with open(cache, "rb") as f:
            cis = pickle.load(f)
    if cis is None or len(cis) != n:
        cis = synth.mixed_batch(n, lo=10, hi=200, base_seed=1000, basis="minimal", thr=thr, ax=ax,
                                workers=gen_workers())
        if cache:
            import pickle
            with open(cache, "wb") as f:
                pickle.dump(cis, f)
    log(f"# workload mixed{n}: {n} molecules, natoms {min(c.ncent for c in cis)}-"
        f"{max(c.ncent for c in cis)} (inputs {time.perf_counter() - t0:.1f}s)")
    if args.inputs_only:
        return None
    plain_kst = {}
    ref, plain_wall, plain_lat = plain_fp64(cis, thr, ax, nroots, plain_kst)
    nok = sum(1 for r in ref if not isinstance(r, int))
    log(f"# plain FP64 sequential: {plain_wall:.3f}s, {nok}/{n} molecules with CSFs")
    if args.plain_only:   # for counter profiles of the plain pass (one thread, one stream)
        return {"plain_fp64": {"value": nok / plain_wall, "unit": "solves/s",
                               "kernels": plain_rates(plain_kst)}}
    with StdaHipBatch(0, workers=args.workers) as b:
        b.configure(args.nmax_batched, args.poll)
        b.configure_wide(not args.no_wide, two_phase=args.two_phase, nmax_wide=args.nmax_wide)
        b.set_precision(args.batch_precision)
        for w in range(args.warmup):
            b.solve_tda(cis, thr, ax, nroots)
            log(f"# warmup {w + 1}/{args.warmup} done")
        walls, phases = [], []
        tstart = time.perf_counter()
        for s in range(args.steps):
            ts = time.perf_counter()
            res = b.solve_tda(cis, thr, ax, nroots)
            walls.append(time.perf_counter() - ts)
            phases.append(b.timings())
            log(f"# step {s + 1}/{args.steps}: {walls[-1]:.3f}s {phases[-1]} wide {b.wide_count()}")
        nwide = b.wide_count()
        total = time.perf_counter() - tstart
        kst = {k: b.kernel_stats(k) for k in SINGLE_KERNELS + COMPOSITE_SPANS}
        phase_us = b.phase_times()
    err = compare(res, ref)
    try:
        spectra = spectra_leg(cis, res)
    except Exception as exc:   # a side measurement never takes the main line down
        spectra = {"error": repr(exc)}
    try:
        prep = prep_leg()
    except Exception as exc:
        prep = {"error": repr(exc)}
    t, cnt = phases[-1]
    # Roofline: PT2 pass 1 (pt2_pipe_kernel), the kernel with the largest
    # summed device time in the batched step (rocprof, profiles/).  Its rate is
    # taken from the sequential plain FP64 pass over the same molecules, where
    # one stream runs at a time and an event pair brackets the kernel alone;
    # in the batched step 16 worker streams and the lock-step graph share the
    # device, so the workers' event times there (listed under "batched") also
    # hold the waits for it.  Algorithmic work: 4 Na Nsec Nprim flops.
    pst = plain_kst.get("pt2_rows", {"ms": 0.0, "launches": 0.0, "work": 0.0})
    achieved = pst["work"] / (pst["ms"] * 1e-3) / 1e12 if pst["ms"] > 0 else 0.0
    roofline = {"kernel": "pt2_rows (pt2_pipe_kernel, PT2 pass 1)", "bound": "mfma",
                "achieved": round(achieved, 3), "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": round(achieved / FP64_PEAK_TFLOPS, 4), "traffic": PT2_MIXED_TRAFFIC,
                "avg_launch_ms": round(pst["ms"] / max(pst["launches"], 1.0), 4),
                "work_per_launch": pst["work"] / max(pst["launches"], 1.0),
                "measured_in": "sequential plain FP64 pass (one stream)",
                "batched": {k: {"ms": round(v["ms"], 2), "launches": int(v["launches"])}
                            for k, v in kst.items()}}
    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline_batch(cis, thr, ax, nroots, args.cpu_budget)
    value = nok * args.steps / total
    plain = nok / plain_wall
    return {
        "metric": METRIC,
        "value": value,
        "unit": "solves/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": total / args.steps * 1e3,
        "p50_ms": float(np.percentile(np.array(walls) * 1e3, 50)),
        "p99_ms": float(np.percentile(np.array(walls) * 1e3, 99)),
        "latency_note": "closed batch: every request of a step completes with the step",
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": ("f32 operator products + f64 Davidson numerics and refinement"
                  if args.batch_precision == "mixed" else
                  "split-bf16 then f32 operator products on the single path, f32 in the lock "
                  "steps + f64 Davidson numerics and refinement"),
        "plain_fp64": {"value": plain, "unit": "solves/s", "wall_s": plain_wall,
                       "kernels": plain_rates(plain_kst),
                       "p50_ms": float(np.percentile(plain_lat * 1e3, 50)),
                       "p99_ms": float(np.percentile(plain_lat * 1e3, 99))},
        "speedup_vs_plain_fp64": value / plain,
        "root_error_vs_fp64": err,
        "data": "synthetic",
        "config": {"workload": f"mixed{n}: {n} molecules, natoms log-uniform [10,200], minimal "
                               f"basis, Ethr 7 eV, ax 0.25, 20 roots each, batched mixed precision",
                   "molecules": n, "molecules_with_csf": nok, "workers": args.workers,
                   "nmax_batched": args.nmax_batched, "nmax_wide": args.nmax_wide,
                   "batch_precision": args.batch_precision,
                   "wide": not args.no_wide, "poll": args.poll, "host": host_info(),
                   "paths": dict(cnt, wide_of_batched=nwide), "phases_ms": t,
                   "phase_ms_summed": {k: round(v / 1e3, 2) for k, v in phase_us.items()
                                       if k not in ("sweeps", "unused")}},
        "roofline": roofline,
        "spectra": spectra,
        "preparation": prep,
        "cpu_baseline": cpu,
    }


def prep_leg(natoms: int = 400):
    """The preparation before the boundary for one 400-atom molecule (minimal basis): AO
    integrals, moment transforms and Loewdin on the device against the same steps with
    numpy on the host; max deviations of the Loewdin orbitals and moments."""
    from stda_amd import synth
    from stda_amd.aoints import one_electron
    from stda_amd.core import StdaHip
    from stda_amd.prep import prepare
    mol = synth.cluster(natoms, seed=7, basis="minimal")[0]
    t0 = time.perf_counter()
    ints = one_electron(mol)
    t1 = time.perf_counter()
    host = prepare(mol, 10.0, 0.25, ints)
    t2 = time.perf_counter()
    with StdaHip(0) as h:
        prepare(mol, 10.0, 0.25, device=h)          # warm-up (library kernels, solver)
        t3 = time.perf_counter()
        dev = prepare(mol, 10.0, 0.25, device=h)
        t4 = time.perf_counter()
    log(f"# preparation {natoms} atoms: host {t2 - t0:.2f}s, device {t4 - t3:.3f}s")
    return {"workload": f"cluster of {natoms} atoms, minimal basis, nao {mol.nao}",
            "host_s": {"integrals": t1 - t0, "transforms": t2 - t1},
            "device_s": t4 - t3,
            "max_abs_dclow": float(np.abs(dev.clow - host.clow).max()),
            "max_abs_dmoments": float(np.abs(dev.moments - host.moments).max())}


def spectra_leg(cis, res, reps: int = 20):
    """The batch's roots through tda.dat and g_spec: every molecule's spectrum (the CD curve
    g_spec draws from a velocity-corrected tda.dat) broadened in one launch of its kernel.  Kernel time from HIP events on the context's stream;
    host parsing/formatting timed separately (outside the kernel figure)."""
    from stda_amd import spectrum as S
    from stda_amd.core import StdaHip
    t0 = time.perf_counter()
    texts = [S.tda_dat(om, pr, ci.xmolw, velcorr=True)
             for ci, (om, pr, st) in zip(cis, res) if st == 0 and len(om)]
    inputs = [S.parse_gspec(t) for t in texts]
    parse_ms = (time.perf_counter() - t0) * 1e3
    lines = sum(len(g.x) for g in inputs)
    with StdaHip(0) as h:
        S.broaden(h, inputs)                      # warm-up
        k0 = h.kernel_stats("gspec")
        t1 = time.perf_counter()
        for _ in range(reps):
            curves = S.broaden(h, inputs)
        call_ms = (time.perf_counter() - t1) * 1e3 / reps
        k1 = h.kernel_stats("gspec")
    t2 = time.perf_counter()
    S._spec_text(np.concatenate([c[0] for c in curves]), np.concatenate([c[1] for c in curves]))
    fmt_ms = (time.perf_counter() - t2) * 1e3
    kms = (k1["ms"] - k0["ms"]) / max(k1["launches"] - k0["launches"], 1.0)
    # the reference program itself (built from its source into oracle/_ref) on the same files
    ref_prog = os.path.join(ROOT, "oracle", "_ref", "g_spec")
    ref = None
    if os.path.exists(ref_prog):
        import subprocess
        import tempfile
        tr = time.perf_counter()
        done = 0
        with tempfile.TemporaryDirectory() as d:
            for text in texts:
                subprocess.run([ref_prog], input=text.encode(), stdout=subprocess.DEVNULL,
                               cwd=d, check=True)
                done += 1
                if time.perf_counter() - tr > 5.0:
                    break
        el = time.perf_counter() - tr
        ref = {"value": done / el, "unit": "spectra/s", "cores": 1, "kind": "reference",
               "sample": f"reference g_spec (flang build of g_spec/g_spec.f) run once per "
                         f"tda.dat on the first {done} files, one process each"}
    terms = (k1["work"] - k0["work"]) / max(k1["launches"] - k0["launches"], 1.0)
    log(f"# spectra: {len(inputs)} spectra, {lines} lines, kernel {kms:.4f} ms, call "
        f"{call_ms:.3f} ms, parse {parse_ms:.0f} ms, format {fmt_ms:.0f} ms")
    return {"spectra": len(inputs), "lines": lines, "points": inputs[0].npts if inputs else 0,
            "kernel_ms": round(kms, 4), "terms_per_launch": terms,
            "terms_per_s": terms / (kms * 1e-3) if kms > 0 else None,
            "call_ms": round(call_ms, 3), "spectra_per_s_call": len(inputs) / (call_ms * 1e-3),
            "host_parse_ms": round(parse_ms, 1), "host_format_ms": round(fmt_ms, 1),
            "end_to_end_spectra_per_s": len(inputs) / ((parse_ms + call_ms + fmt_ms) * 1e-3),
            "cpu_baseline": ref}


# --------------------------------------------------------------------------- serving
def run_serve(args):
    from stda_amd import synth
    from stda_amd.serve import Server, Job, poisson_arrivals
    n = args.nmol
    t0 = time.perf_counter()
    cis = synth.mixed_batch(n, lo=10, hi=400, base_seed=2000, basis="minimal", thr=7.0, ax=0.25,
                            workers=gen_workers())
    log(f"# serve workload: {n} molecules (inputs {time.perf_counter() - t0:.1f}s)")
    from stda_amd.serve import request_mix
    kinds = request_mix(n, seed=2000)[1]
    rates = [float(r) for r in args.rates.split(",")]
    sweep = []
    with Server(workers=4, batch_workers=args.workers) as srv:
        warm = [Job(jid=i, kind=str(kinds[i]), ci=cis[i], omegas=(0.0, 0.05, 0.1, 0.15))
                for i in range(min(16, n))]
        srv.run(warm, realtime=False)
        for rate in rates:
            arr = poisson_arrivals(n, rate, seed=int(rate))
            jobs = [Job(jid=i, kind=str(kinds[i]), ci=cis[i], arrival=float(arr[i]),
                        omegas=(0.0, 0.05, 0.1, 0.15)) for i in range(n)]
            st = srv.run(jobs, realtime=True,
                         progress=lambda d, t: log(f"#   rate {rate}: {d}/{t} jobs done"))
            errs = {}
            for j in jobs:
                if j.error:
                    key = f"{j.kind}: {j.error[:110]}"
                    errs[key] = errs.get(key, 0) + 1
            slow = sorted(jobs, key=lambda j: -j.latency)[:3]
            sweep.append({"rate": rate, "solves_per_s": st.solves_per_s, "p50_ms": st.p50_ms,
                          "p99_ms": st.p99_ms, "ok": st.ok, "n": st.n,
                          "spectra_submitted": st.submitted, "errors": errs,
                          "slowest": [{"kind": j.kind, "natoms": int(j.ci.ncent),
                                       "latency_s": round(j.latency, 2)} for j in slow]})
            log(f"# rate {rate}: {sweep[-1]}")
    best = max(sweep, key=lambda r: r["solves_per_s"])
    return {
        "metric": METRIC, "value": best["solves_per_s"], "unit": "solves/s",
        "n_gpus": 1, "steps": len(rates), "warmup": 1,
        "ms_per_step": None, "p50_ms": best["p50_ms"], "p99_ms": best["p99_ms"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 operator products + f64 Davidson numerics and refinement",
        "data": "synthetic",
        "config": {"workload": f"serve: open-loop Poisson, {n} requests per rate, 70% spectrum "
                               f"(20 roots) / 30% polarizability (4 frequencies), natoms "
                               f"log-uniform [10,400]", "sweep": sweep},
        "roofline": None, "cpu_baseline": None,
    }


def rocprof_summary(db_path: str, top: int = 25):
    """Per-kernel statistics of a `rocprofv3 --kernel-trace` database
    (rocpd SQLite) of a bench run, as text: the whole process, the plain FP64
    pass (the stream of the first PT2 pass-1 launch, up to its last launch)
    and the rest (the batched steps)."""
    import sqlite3
    con = sqlite3.connect(db_path)
    cur = con.cursor()
    rows = cur.execute("select name, stream_id, start, end from kernels order by start").fetchall()
    if not rows:
        return "no kernels"
    first = next((r for r in rows if "pt2_pipe_kernel" in r[0]), None)
    plain_sid = first[1] if first else None
    plain_end = max((r[3] for r in rows if r[1] == plain_sid), default=0) if first else 0

    def table(sel, title):
        agg = {}
        for n, _, s0, e0 in sel:
            k = n.replace("(anonymous namespace)::", "").replace("void ", "").split("(")[0]
            a = agg.setdefault(k, [0, 0.0, float("inf"), 0.0])
            d = (e0 - s0) / 1e3
            a[0] += 1
            a[1] += d
            a[2] = min(a[2], d)
            a[3] = max(a[3], d)
        out = [f"## {title}", "", "| kernel | launches | total ms | avg us | min us | max us |",
               "|---|---|---|---|---|---|"]
        for k, a in sorted(agg.items(), key=lambda kv: -kv[1][1])[:top]:
            out.append(f"| `{k}` | {a[0]} | {a[1] / 1e3:.2f} | {a[1] / a[0]:.1f} | {a[2]:.1f} | {a[3]:.1f} |")
        return "\n".join(out)

    plain = [r for r in rows if r[1] == plain_sid and r[3] <= plain_end]
    rest = [r for r in rows if not (r[1] == plain_sid and r[3] <= plain_end)]
    return "\n\n".join([table(rows, "whole process"),
                          table(plain, "plain FP64 pass (one stream)"),
                          table(rest, "batched steps (worker streams and the lock-step graph)")])


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="mixed512",
                    help="mixed512 (default), cluster400, cluster100, polyene, serve")
    ap.add_argument("--nmol", type=int, default=512)
    ap.add_argument("--workers", type=int, default=16)
    ap.add_argument("--nmax-batched", type=int, default=1024,
                    help="CSF limit of the lock-step solver (0: every molecule on the worker path)")
    ap.add_argument("--no-wide", action="store_true",
                    help="molecules above --nmax-batched on the single-molecule path instead of "
                         "the matrix-free lock step")
    ap.add_argument("--nmax-wide", type=int, default=16384,
                    help="CSF limit of the matrix-free lock step (0: none; larger molecules take "
                         "the single-molecule path)")
    ap.add_argument("--batch-precision", choices=["mixed", "bf16"], default="mixed",
                    help="mixed512: operator products of the single-path solves (FP32, or a "
                         "split-bfloat16 phase before the FP32 one)")
    ap.add_argument("--two-phase", action="store_true",
                    help="lock-step iterations start once every molecule is prepared")
    ap.add_argument("--poll", type=int, default=4,
                    help="lock-step iterations between reads of the state table")
    ap.add_argument("--rates", default="50,200,800,2000")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    ap.add_argument("--precision", choices=["fp64", "mixed", "bf16"], default="fp64",
                    help="single-molecule workloads only")
    ap.add_argument("--copy-probe", type=int, default=0,
                    help="only copy this many bytes device-to-device (calibrates HBM counters)")
    ap.add_argument("--inputs-cache", default="",
                    help="mixed workload: read the generated molecules from this file, or write "
                         "them there (a run under a counter profile cannot fork the generator)")
    ap.add_argument("--plain-only", action="store_true",
                    help="mixed workload: only the sequential plain FP64 pass (for counter profiles)")
    ap.add_argument("--inputs-only", action="store_true",
                    help="with --inputs-cache: generate and write the molecules, then exit")
    ap.add_argument("--rocprof-summary", default="",
                    help="print per-kernel statistics of a rocprofv3 kernel-trace database and exit")
    ap.add_argument("--gemm-probe", action="store_true",
                    help="only run the operator-shape GEMMs (library kernel and vendor GEMM)")
    args = ap.parse_args()
    if args.rocprof_summary:
        print(rocprof_summary(args.rocprof_summary))
        return
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return
    if args.gemm_probe:
        gemm_probe()
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for {args.gpus}); the multi-GPU "
              "decomposition is a design, not a measured path in this build.", file=sys.stderr)
        sys.exit(2)
    if args.workload.startswith("mixed"):
        if args.workload != "mixed512":
            args.nmol = int(args.workload[5:])
        line = run_mixed(args)
        if line is None:   # --inputs-only
            return
    elif args.workload == "serve":
        line = run_serve(args)
    else:
        line = run_single(args)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
