"""Continuum-LDOS throughput on one MI355X.

A step is one whole pass of the device chain over the workload: per-k
eigensolve of H(k), then for every energy the resolvent, the k->R transform,
the T-matrix and the continuum contraction into the W x W map (maps stay in
HBM).  Default workload: the 3-orbital square lattice (K=512, 51 energies,
128x128 cells, os=4, w=2 -> 512x512 pixels), synthetic model drawn from a
fixed seed.  Prints one JSON line.

  python bench.py [--steps K] [--warmup W] [--gpus 1] [--precision P] [--no-cpu]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "calcqpi-release_amd"))

import numpy as np  # noqa: E402

METRIC = "LDOS pixel-energy points/s, mixed vs FP64 on 1 GPU at <=1e-4 rel err; grad time"
HBM_PEAK_GBS = 8000.0
PEAK_TFLOPS = {"fp64": 78.6, "fp32": 157.3, "mixed": 2500.0}


def flops_per_point(wl):
    """Algorithmic flops of the contraction per pixel-energy point:
    the two complex-by-real gathers a and b over (2w+1)^2 * |orbital set| terms
    for each of the b columns, plus the T sandwich per spin block."""
    side2 = (2 * wl.window + 1) ** 2
    nsets = 2 if wl.spin else 1
    per = (wl.maxband >> 1) if wl.spin else wl.maxband
    b = wl.bands
    return nsets * (8 * side2 * per * b + 8 * b * b + 8 * b)


def stage_bytes(wl, nw, E):
    """Algorithmic HBM bytes of the transform stages for E energies."""
    K, b = wl.K, wl.bands
    eig = K * K * (16 * b * b + 8 * b)
    f1 = 16 * b * b * K * nw
    win = 16 * b * b * nw * nw
    return {1: E * (eig + f1), 2: E * (f1 + win)}


def cpu_baseline(wl, nsample):
    """Oracle (FP64 C restatement of the reference algorithm, reference loop
    nest for the contraction) timed on this host: eigensolve once plus
    nsample energies, extrapolated linearly to all energies."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle

    t0 = time.perf_counter()
    ev, V = oracle.calchamiltonian(wl.hops, wl.bands, wl.efermi, wl.K)
    t_eig = time.perf_counter() - t0
    t0 = time.perf_counter()
    for om in wl.energies[:nsample]:
        gR = oracle.realspace(oracle.greens_k(ev, V, om, wl.eta))
        T = oracle.tmatrix(wl.V, gR[0, 0])
        oracle.wannierldos(gR, T, wl.tiles, wl.n, wl.window, wl.oversamp, wl.maxband, wl.spin)
    t_e = (time.perf_counter() - t0) / nsample
    E = len(wl.energies)
    total = t_eig + E * t_e
    return {"value": wl.points / total, "unit": "points/s", "cores": oracle.num_threads(),
            "kind": "port",
            "sample": "eigensolve + %d of %d energies at full size, extrapolated "
                      "(%.2f s eigensolve, %.2f s/energy)" % (nsample, E, t_eig, t_e)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--precision", default="fp64", choices=["fp64", "fp32", "mixed"])
    ap.add_argument("--config", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-energies", type=int, default=2)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU only (--gpus 1); "
              "multi-GPU sharding is described in DESIGN.md, not run", file=sys.stderr)
        sys.exit(2)

    import torch

    import cqpi
    from cqpi import synth

    if cqpi.device_count() < 1:
        raise SystemExit("no HIP device")
    wl = synth.by_id(args.config)
    q = cqpi.engine_for(wl, args.precision)
    q.set_timing(False)

    def step():
        q.compute_eigen()
        q.run(wl.energies, wl.eta, host=False)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    value = wl.points / dt

    # per-stage device time from HIP events on the library stream (one extra step)
    q.set_timing(True)
    step()
    st = q.stage_times()
    q.set_timing(False)
    names = list(st)
    dom = max(range(len(names)), key=lambda i: st[names[i]][0])
    ms, nl = st[names[dom]]
    avg_s = ms / max(nl, 1) / 1e3
    E = len(wl.energies)
    h = q.window_half()
    nw = 2 * h + 1
    if dom == 4:
        per_launch = wl.W * wl.W * E / max(nl, 1) * flops_per_point(wl)
        achieved = per_launch / avg_s / 1e12
        roof = {"kernel": names[dom], "bound": "mfma" if args.precision == "mixed" else "valu",
                "achieved": achieved, "peak": PEAK_TFLOPS[args.precision], "unit": "TFLOP/s",
                "frac": achieved / PEAK_TFLOPS[args.precision], "traffic": None}
    elif dom in (1, 2):
        per_launch = stage_bytes(wl, nw, E)[dom] / max(nl, 1)
        achieved = per_launch / avg_s / 1e9
        roof = {"kernel": names[dom], "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": None}
    else:
        roof = {"kernel": names[dom], "bound": "latency", "achieved": None, "peak": None,
                "unit": None, "frac": None, "traffic": None}
    roof["avg_launch_ms"] = avg_s * 1e3
    roof["launches_per_step"] = nl

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "points/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else args.precision,
        "data": "synthetic (seeded model, Gaussian x angular tiles)",
        "config": {"workload": "square lattice 3 orbitals, K=512, 51 energies in [-0.1,0.1], "
                               "eta=0.002, 128x128 cells, os=4, w=2 (512x512 pixels)"
                   if args.config == 2 else "config %d" % args.config,
                   "kpoints": wl.K, "energies": E, "cells": wl.n, "oversamp": wl.oversamp,
                   "window": wl.window, "bands": wl.bands, "precision": args.precision},
        "stages_ms": {k: v[0] for k, v in st.items()},
        "roofline": roof,
    }
    q.close()
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(wl, args.cpu_energies)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
