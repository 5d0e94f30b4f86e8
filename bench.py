"""Continuum-LDOS throughput on one MI355X: mixed precision vs FP64.

A step is one whole pass of the device chain over the workload: per-k
eigensolve of H(k), then for every energy the resolvent, the k->R transform,
the T-matrix and the continuum contraction into the W x W map (maps stay in
HBM).  Default workload: the 3-orbital square lattice (K=512, 51 energies,
128x128 cells, os=4, w=2 -> 512x512 pixels), synthetic model drawn from a
fixed seed.

`value` is the mixed-precision path (split-fp16 MFMA contraction, FP64
eigensolve / resolvent / transform / T-matrix / periodic term).  The same
line reports the FP64 path, their ratio, the accuracy of both against each
other and against the FP64 CPU oracle, the roofline of the dominant kernel
and the CPU baseline.  Prints one JSON line.

  python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu]
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
PEAK_TFLOPS = {"fp64": 78.6, "mixed": 2500.0}   # FP64 vector/matrix; dense fp16 MFMA


def flops_per_point(wl):
    """Algorithmic flops of the contraction per pixel-energy point: the two
    complex-by-real gathers a and b over (2w+1)^2 x |orbital block| terms for
    each of the b columns, plus the T sandwich, per spin block."""
    side2 = (2 * wl.window + 1) ** 2
    nsets = 2 if wl.spin else 1
    per = (wl.maxband >> 1) if wl.spin else wl.maxband
    b = wl.bands
    return nsets * (8 * side2 * per * b + 8 * b * b + 8 * b)


def stage_bytes(wl, nw, E):
    """Algorithmic HBM bytes of the transform stages for E energies:
    stage 1 reads the eigen data and writes the ky-transformed rows of the
    window, stage 2 reads those rows and writes the G0(R) window."""
    K, b = wl.K, wl.bands
    eig = K * K * (16 * b * b + 8 * b)
    f1 = 16 * b * b * K * nw
    win = 16 * b * b * nw * nw
    return {1: E * (eig + f1), 2: E * (f1 + win)}


def rel_errors(x, ref, floor=1e-3):
    m = np.max(np.abs(ref))
    den = np.maximum(np.abs(ref), floor * m)
    return float(np.max(np.abs(x - ref) / den)), float(np.linalg.norm(x - ref) / np.linalg.norm(ref))


def cpu_baseline(wl, nsample):
    """The FP64 oracle (C restatement of the reference algorithm with the
    reference loop nest for the contraction, OpenMP over cells) timed on this
    host: eigensolve once plus nsample energies, extrapolated linearly to all
    energies.  Also returns the oracle maps of the sampled energies."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle

    t0 = time.perf_counter()
    ev, V = oracle.calchamiltonian(wl.hops, wl.bands, wl.efermi, wl.K)
    t_eig = time.perf_counter() - t0
    maps = []
    t0 = time.perf_counter()
    for om in wl.energies[:nsample]:
        gR = oracle.realspace(oracle.greens_k(ev, V, om, wl.eta))
        T = oracle.tmatrix(wl.V, gR[0, 0])
        maps.append(oracle.wannierldos(gR, T, wl.tiles, wl.n, wl.window, wl.oversamp, wl.maxband, wl.spin))
    t_e = (time.perf_counter() - t0) / nsample
    E = len(wl.energies)
    total = t_eig + E * t_e
    info = {"value": wl.points / total, "unit": "points/s", "cores": oracle.num_threads(),
            "kind": "port",
            "sample": "oracle eigensolve + %d of %d energies at full size, extrapolated linearly "
                      "(%.2f s eigensolve, %.2f s/energy, 1 process, %d OpenMP threads)"
                      % (nsample, E, t_eig, t_e, oracle.num_threads())}
    return info, np.array(maps)


def time_steps(q, wl, steps, warmup, torch):
    def step():
        q.compute_eigen()
        q.run(wl.energies, wl.eta, host=False)

    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    q.set_timing(True)
    step()
    st = q.stage_times()
    q.set_timing(False)
    return dt, st


def gradient_timing(cqpi, synth, torch, reps=3):
    """Gradient workload (3-orbital model, K=512, 21 energies, 64x64 cells):
    FP64 forward (keeping the per-energy windows) and the hand-written
    adjoint of an L2 map loss, both timed on the device."""
    wl = synth.workload("square_3orb_grad")
    tgt = synth.perturbed(wl)
    with cqpi.engine_for(tgt) as qt:
        qt.run(tgt.energies, tgt.eta, host=False)
        target = qt.ldos_tensor().clone()
    q = cqpi.engine_for(wl, "fp64")
    q.set_keep_windows(True)
    fw, bw = [], []
    for it in range(reps + 1):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        q.compute_eigen()
        q.run(wl.energies, wl.eta, host=False)
        rho = q.ldos_tensor()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        rhobar = 2.0 * (rho - target)          # dL/drho of L = sum (rho - target)^2, on the device
        torch.cuda.synchronize()
        q.gradient_device(rhobar.data_ptr())
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        if it:
            fw.append(t1 - t0)
            bw.append(t2 - t1)
    q.close()
    f, b = float(np.median(fw)), float(np.median(bw))
    return {"workload": "square lattice 3 orbitals, K=512, 21 energies, 64x64 cells, os=4, w=2; "
                        "L2 map loss; gradients w.r.t. %d hopping amplitudes (re, im), mu and V" % len(wl.hops),
            "forward_ms": f * 1e3, "backward_ms": b * 1e3, "backward_over_forward": b / f,
            "gradient_over_forward": (f + b) / f,
            "note": "FP64; forward = eigensolve + all energies; backward = dL/drho on the device + "
                    "adjoint; maps and dL/drho stay in HBM"}


def roofline(wl, q, st, precision):
    names = list(st)
    dom = max(range(len(names)), key=lambda i: st[names[i]][0])
    ms, nl = st[names[dom]]
    avg_s = ms / max(nl, 1) / 1e3
    E = len(wl.energies)
    nw = 2 * q.window_half() + 1
    if dom == 4:
        per_launch = wl.W * wl.W * E / max(nl, 1) * flops_per_point(wl)
        achieved = per_launch / avg_s / 1e12
        peak = PEAK_TFLOPS[precision]
        roof = {"kernel": names[dom], "bound": "mfma", "achieved": achieved, "peak": peak,
                "unit": "TFLOP/s", "frac": achieved / peak, "traffic": None}
    elif dom in (1, 2):
        per_launch = stage_bytes(wl, nw, E)[dom] / max(nl, 1)
        achieved = per_launch / avg_s / 1e9
        roof = {"kernel": names[dom], "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": None}
    else:
        roof = {"kernel": names[dom], "bound": "hbm", "achieved": None, "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": None, "traffic": None}
    roof["avg_launch_ms"] = avg_s * 1e3
    roof["launches_per_step"] = nl
    return roof


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="square_3orb")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-energies", type=int, default=2)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU (--gpus 1); the multi-GPU energy sharding is a design, "
              "not run in this build", file=sys.stderr)
        sys.exit(2)

    import torch

    import cqpi
    from cqpi import synth

    if cqpi.device_count() < 1:
        raise SystemExit("no HIP device")
    wl = synth.workload(args.workload)
    E = len(wl.energies)

    qm = cqpi.engine_for(wl, "mixed")
    dt_m, st_m = time_steps(qm, wl, args.steps, args.warmup, torch)
    q6 = cqpi.engine_for(wl, "fp64")
    dt_6, st_6 = time_steps(q6, wl, args.steps, args.warmup, torch)

    # accuracy: every energy of the mixed path against the FP64 device path
    rm_ = qm.run(wl.energies, wl.eta, host=True)
    r6 = q6.run(wl.energies, wl.eta, host=True)
    errs = [rel_errors(rm_[e], r6[e]) for e in range(E)]
    acc = {"mixed_vs_fp64_gpu_max_rel": max(e[0] for e in errs),
           "mixed_vs_fp64_gpu_rms_rel": max(e[1] for e in errs),
           "floor": "max|d| / max(|ref|, 1e-3 max|ref|) per energy, max over energies"}

    out = {
        "metric": METRIC,
        "value": wl.points / dt_m,
        "unit": "points/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt_m * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "mixed: fp16x2 split MFMA contraction (f32 accumulate), f64 elsewhere",
        "data": "synthetic (seeded 3-orbital model, Gaussian x {xz,yz,xy} tiles)",
        "config": {"workload": "square lattice 3 orbitals, K=512, 51 energies in [-0.1,0.1], "
                               "eta=0.002, 128x128 cells, os=4, w=2 (512x512 pixels)"
                   if args.workload == "square_3orb" else "%s: %s" % (args.workload, wl.name),
                   "kpoints": wl.K, "energies": E, "cells": wl.n, "oversamp": wl.oversamp,
                   "window": wl.window, "bands": wl.bands},
        "fp64": {"value": wl.points / dt_6, "ms_per_step": dt_6 * 1e3,
                 "stages_ms": {k: v[0] for k, v in st_6.items()},
                 "roofline": roofline(wl, q6, st_6, "fp64")},
        "speedup_mixed_vs_fp64": dt_6 / dt_m,
        "stages_ms": {k: v[0] for k, v in st_m.items()},
        "accuracy": acc,
        "roofline": roofline(wl, qm, st_m, "mixed"),
    }
    qm.close()
    q6.close()
    out["gradient"] = gradient_timing(cqpi, synth, torch)
    if not args.no_cpu:
        info, maps = cpu_baseline(wl, args.cpu_energies)
        out["cpu_baseline"] = info
        e6 = [rel_errors(r6[e], maps[e]) for e in range(len(maps))]
        em = [rel_errors(rm_[e], maps[e]) for e in range(len(maps))]
        acc["fp64_gpu_vs_oracle_max_rel"] = max(e[0] for e in e6)
        acc["mixed_vs_oracle_max_rel"] = max(e[0] for e in em)
        acc["mixed_vs_oracle_rms_rel"] = max(e[1] for e in em)
        acc["oracle_energies"] = len(maps)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
