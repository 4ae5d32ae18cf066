"""Benchmark of the MI355X excited-state core.

One step = one full excited-state solve of the workload on the GPU: the
device pipeline from the FP64 Loewdin orbitals (gamma, transition charges,
screening, PT2 CSF selection) through the FP64 block Davidson to the
transition properties with the velocity correction.  The system arrays are
copied to HBM before the timed region (``StdaHip.load``); the host-side
generation of the synthetic molecule and its AO integrals is outside it.

Default workload: a 400-atom organic-like cluster (C:H:N:O = 45:45:5:5,
random positions >= 1.0 Angstrom apart, seed 1) in a 3-21G-type
split-valence s/p basis with 800 occupied orbitals, Ethr = 10 eV,
a_x = 0.25, the 50 lowest sTDA roots, Davidson residual 1e-8.

Prints one JSON line (see README of the metric in the output keys).
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

FP64_PEAK_TFLOPS = 78.6      # MI355X dense FP64 (vector and matrix)
HBM_PEAK_GBS = 8000.0


def workload(name: str):
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
    return mol, ci, p


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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="cluster400")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="fp64")
    ap.add_argument("--copy-probe", type=int, default=0,
                    help="only copy this many bytes device-to-device (calibrates HBM counters)")
    args = ap.parse_args()
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for {args.gpus}); the multi-GPU "
              "decomposition is a design, not a measured path in this build.", file=sys.stderr)
        sys.exit(2)

    t0 = time.perf_counter()
    mol, ci, p = workload(args.workload)
    print(f"# workload {args.workload}: natoms={ci.ncent} nao={ci.nao} no={ci.no} nv={ci.nv} "
          f"(inputs {time.perf_counter() - t0:.1f}s)", file=sys.stderr, flush=True)

    from stda_amd.core import StdaHip
    h = StdaHip(0)
    h.load(ci, p["thr"], p["ax"])
    h.set_precision(args.precision)
    for w in range(args.warmup):
        h.select()
        h.solve_tda(p["nroots"], velcorr=True, vectors=False)
        print(f"# warmup {w + 1}/{args.warmup} done", file=sys.stderr, flush=True)
    lat = []
    stats = []
    tstart = time.perf_counter()
    for s in range(args.steps):
        ts = time.perf_counter()
        nci, nprim, nsec = h.select()
        sol = h.solve_tda(p["nroots"], velcorr=True, vectors=False)
        lat.append(time.perf_counter() - ts)
        stats.append({k: h.kernel_stats(k) for k in ("pt2_rows", "pt2_cols", "ao_matvec",
                                                       "ao_matvec_f32")})
        host = {k: h.kernel_stats(k) for k in ("dav_matvec", "dav_iter_other", "dav_rr_eig",
                                                "dav_ortho", "mixed_fp32_phase",
                                                "mixed_fp64_refine")}
        print(f"# host spans: {host}", file=sys.stderr, flush=True)
        print(f"# step {s + 1}/{args.steps}: {lat[-1]:.3f}s nci={nci} "
              f"timings={sol.timings_ms}", file=sys.stderr, flush=True)
    total = time.perf_counter() - tstart
    root_err = None
    if args.precision != "fp64":
        # per-root error of the mixed-precision roots against the plain FP64 GPU
        # path (itself checked against the FP64 oracle by the parity tests)
        h.set_precision("fp64")
        ref = h.solve_tda(p["nroots"], velcorr=True, vectors=False)
        root_err = {"max_abs_dE_eV": float(np.abs(sol.omega - ref.omega).max() * 27.211385050),
                    "max_abs_dfL": float(np.abs(sol.props[:, 0] - ref.props[:, 0]).max())}
    h.close()

    # dominant kernel of the last step
    last = stats[-1]
    dom = max(last, key=lambda k: last[k]["ms"])
    st = last[dom]
    achieved = st["work"] / (st["ms"] * 1e-3) / 1e12 if st["ms"] > 0 else 0.0
    roofline = {"kernel": dom, "bound": "mfma", "achieved": round(achieved, 3),
                "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": round(achieved / FP64_PEAK_TFLOPS, 4), "traffic": None,
                "avg_launch_ms": round(st["ms"] / max(st["launches"], 1), 3),
                "work_per_launch": st["work"] / max(st["launches"], 1)}

    # HBM bytes per launch of the dominant kernel, measured with rocprofv3
    # counters for this workload (FETCH_SIZE + WRITE_SIZE, copy-calibrated)
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
    line = {
        "metric": "excited-state solves/s (plain FP64 sTDA, 1 molecule per solve); p50/p99 latency",
        "value": args.steps / total,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": total / args.steps * 1e3,
        "p50_ms": float(np.percentile(lat_ms, 50)),
        "p99_ms": float(np.percentile(lat_ms, 99)),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32 contractions + f64 Davidson",
        "precision": args.precision,
        "root_error_vs_fp64": root_err,
        "data": "synthetic",
        "config": {"workload": f"{args.workload}: plain FP64 sTDA drop-in, "
                               f"{p['nroots']} lowest roots",
                   "natoms": ci.ncent, "nao": ci.nao, "no": ci.no, "nv": ci.nv,
                   "nci": int(nci), "nprim": int(nprim), "nsec": int(nsec),
                   "ethr_ev": p["thr"], "ax": p["ax"], "nroots": p["nroots"],
                   "davidson_tol": 1e-8},
        "roofline": roofline,
        "kernels": {k: {"ms": round(v["ms"], 3), "launches": v["launches"],
                        "tflops": round(v["work"] / (v["ms"] * 1e-3) / 1e12, 3) if v["ms"] else 0}
                    for k, v in last.items() if v["launches"]},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
