"""Continuum-LDOS throughput on one MI355X: mixed precision vs FP32 vs FP64.

A step is one whole pass of the device chain over the workload: per-k
eigensolve of H(k), then for every energy the resolvent, the k->R transform,
the T-matrix and the continuum contraction into the W x W map (maps stay in
HBM).  Default workload: the 3-orbital square lattice (K=512, 51 energies,
128x128 cells, os=4, w=2 -> 512x512 pixels), synthetic model drawn from a
fixed seed.

`value` is the mixed-precision path (split-fp16 MFMA contraction, FP64
eigensolve / resolvent / transform / T-matrix / periodic term).  The same
line reports the FP32 and FP64 paths, their ratios, the accuracy against the
FP64 path and the FP64 CPU oracle, the roofline of the dominant kernel, the
batched workload (128 uneven instances in one HIP graph), the gradient
workload and the CPU baseline.  Prints one JSON line.

  python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu] [--no-batch]
                  [--workload NAME] [--energies N]
  python bench.py --probe        one mixed and one FP64 step after a 256 MiB
                                 device copy (the traffic calibration), for
                                 rocprofv3 --pmc runs
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
PEAK_TFLOPS = {"fp64": 78.6, "fp32": 157.3, "mixed": 2500.0}   # FP64, FP32 vector; dense fp16 MFMA


def flops_per_point(wl):
    """Algorithmic flops of the contraction per pixel-energy point: the two
    complex-by-real gathers a and b over (2w+1)^2 x |orbital block| terms for
    each of the b columns, plus the T sandwich, per spin block."""
    side2 = (2 * wl.window + 1) ** 2
    nsets = 2 if wl.spin else 1
    per = (wl.maxband >> 1) if wl.spin else wl.maxband
    b = wl.bands
    return nsets * (8 * side2 * per * b + 8 * b * b + 8 * b)


def stage_bytes(wl, nw, E, csize=16):
    """Algorithmic HBM bytes of the transform stages for E energies, at the
    fused minimum: stage 1 (resolvent fused with the ky transform) reads the
    eigen data once and writes the ky-transformed rows of the window, stage 2
    reads those rows and writes the G0(R) window (FP64) or, in mixed
    precision, the fp16 hi/lo planes (the same 16 bytes per complex point).
    csize: bytes per complex point of the rows (8: FP32 lines)."""
    K, b = wl.K, wl.bands
    eig = K * K * (16 * b * b + 8 * b)
    f1 = csize * b * b * K * nw
    win = 16 * b * b * nw * nw
    return {1: eig + E * f1, 2: E * (f1 + win)}


def fft_flops(wl, nw, E):
    """FP32/FP64 arithmetic of the fused resolvent + ky transform for E
    energies: the radix-8 line transforms (about 5 N log2 N flops per line of
    N = K points, b^2 lines per kx row) and the resolvent sandwich (8 flops
    per pair and band per k-point)."""
    K, b = wl.K, wl.bands
    lg = int(np.log2(K))
    return E * (K * b * b * 5 * K * lg + K * K * b * b * b * 8)


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


def time_steps(q, wl, steps, warmup, torch, stage_events=False, eigen=True):
    """Wall time per step (eigensolve + all energies) over `steps` steps after
    `warmup`, then the stage times of one more step.  stage_events: every
    step records stage events, which keeps each run on one stream (one launch
    per stage: what a kernel-trace profile of the stages should see).
    eigen=False: no eigensolve in the step (the surface Green's function does
    not use it)."""
    def step():
        if eigen:
            q.compute_eigen()
        q.run(wl.energies, wl.eta, host=False)

    q.set_timing(stage_events)
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
    q.set_timing(stage_events)   # (with stage_events the later accuracy runs stay on one stream too)
    return dt, st


def variants_timing(cqpi, synth, torch, steps=5, warmup=3, threshold=0.01):
    """The reference's other two contraction kinds at the 3-orbital geometry
    (K=512, 51 energies, 128x128 cells, os=4, w=2), whole chain on the device:
    the Josephson map of a superconducting model (3 orbitals x particle-hole,
    6 bands; FP64 and FP32) and the threshold-list map (FP64) of the 3-orbital
    workload with the driver's list rule at the given threshold; and the LDOS
    map from the surface Green's function of a layered 3-orbital model."""
    out = {}
    wlj = synth.bdg_square(3, K=512, n=128, E=51, seed=1006, w=2, os_=4, eta=0.002, elo=-0.1, ehi=0.1)
    jos = {"workload": "superconducting square lattice, 3 orbitals x particle-hole (6 bands), K=512, "
                       "51 energies in [-0.1,0.1], eta=0.002, 128x128 cells, os=4, w=2; tip deltat=0.001, "
                       "etat=eta"}
    maps = {}
    for prec in ("fp64", "fp32"):
        q = cqpi.engine_for(wlj, prec)
        q.set_josephson(True, deltat=0.001)
        dt, st = time_steps(q, wlj, steps, warmup, torch)
        jos[prec] = {"ms_per_step": dt * 1e3, "points_per_s": wlj.points / dt,
                     "stages_ms": {k: v[0] for k, v in st.items()}}
        maps[prec] = q.run(wlj.energies, wlj.eta, host=True)
        q.close()
    jos["fp32_vs_fp64_gpu_max_rel"] = max(rel_errors(maps["fp32"][e], maps["fp64"][e])[0]
                                          for e in range(len(wlj.energies)))
    out["josephson"] = jos
    wl = synth.workload("square_3orb")
    flags, counts = cqpi.make_threshold_list(wl.tiles, wl.window, wl.oversamp, threshold)
    dense = (2 * wl.window + 1) ** 4 * wl.oversamp ** 2 * wl.bands ** 2
    q = cqpi.AmdQPI.from_list(wl.W, wl.K, wl.n, wl.window, wl.oversamp, wl.bands, flags, counts)
    q.set_model(wl.hops, wl.efermi, wl.scmodel)
    q.set_impurity(wl.V)
    dt, st = time_steps(q, wl, steps, warmup, torch)
    q.close()
    out["threshold_list"] = {
        "workload": "square lattice 3 orbitals, K=512, 51 energies, 128x128 cells, os=4, w=2 (the headline "
                    "workload) with the threshold list",
        "threshold": threshold, "entries": int(len(flags)), "entries_dense": dense,
        "fp64": {"ms_per_step": dt * 1e3, "points_per_s": wl.points / dt,
                 "stages_ms": {k: v[0] for k, v in st.items()}}}
    # surface Green's function of a layered model (green=surface, epserr 1e-5)
    wls = synth.layered(3, K=512, n=128, E=51, seed=1007, w=2, os_=4, eta=0.002, elo=-0.1, ehi=0.1)
    surf = {"workload": "3-orbital square-lattice layers coupled along z (blocks at R=(0,0,-1) and (1,0,-1)), "
                        "K=512, 51 energies in [-0.1,0.1], eta=0.002, 128x128 cells, os=4, w=2; green=surface, "
                        "epserr=1e-5"}
    for prec in ("fp64", "mixed"):
        q = cqpi.engine_for(wls, prec)
        q.set_surface("surface", 1e-5)
        dt, st = time_steps(q, wls, steps, warmup, torch, eigen=False)
        surf[prec] = {"ms_per_step": dt * 1e3, "points_per_s": wls.points / dt,
                      "stages_ms": {k: v[0] for k, v in st.items()}}
        q.close()
    out["surface_green"] = surf
    return out


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


TRAFFIC_MD = os.path.join(ROOT, "profiles", "pmc_traffic_square3orb_K512_E51.md")


def measured_traffic(kernel, precision):
    """HBM bytes per launch of `kernel` (name prefix) in the `precision` step,
    from the committed rocprofv3 --pmc summary of this workload (FETCH_SIZE +
    WRITE_SIZE, copy-calibrated), or None."""
    try:
        txt = open(TRAFFIC_MD).read()
    except OSError:
        return None
    if "```json" not in txt:
        return None
    table = json.loads(txt.split("```json", 1)[1].split("```", 1)[0])
    line = "f64" if precision == "fp64" else "f32"   # line type of the transform kernels
    for name, row in table.get("kernels", {}).items():
        label = name.rsplit(" (", 1)[1].rstrip(")") if " (" in name else None
        typed = "f32" in name or "f64" in name
        if name.startswith(kernel) and label in (None, precision) and (not typed or line in name):
            return row.get("bytes_per_launch")
    return None


def roofline(wl, q, st, precision, stage=None):
    """Roofline of the dominant stage's kernel (or of `stage`, an index into
    the stage list)."""
    names = list(st)
    dom = max(range(len(names)), key=lambda i: st[names[i]][0]) if stage is None else stage
    ms, nl = st[names[dom]]
    avg_s = ms / max(nl, 1) / 1e3
    E = len(wl.energies)
    nw = 2 * q.window_half() + 1
    kern = {1: "k_res_ky_s", 2: "k_kx_s", 4: "k_cmix_areg" if precision == "mixed" else "k_contract"}.get(dom, "")
    if dom == 4:
        per_launch = wl.W * wl.W * E / max(nl, 1) * flops_per_point(wl)
        achieved = per_launch / avg_s / 1e12
        peak = PEAK_TFLOPS[precision]
        roof = {"kernel": names[dom], "bound": "mfma", "achieved": achieved, "peak": peak,
                "unit": "TFLOP/s", "frac": achieved / peak, "traffic": measured_traffic(kern, precision)}
    elif dom in (1, 2):
        csize = 16 if precision == "fp64" else 8
        per_launch = stage_bytes(wl, nw, E, csize)[dom] / max(nl, 1)
        achieved = per_launch / avg_s / 1e9
        roof = {"kernel": names[dom], "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(kern, precision),
                "algorithmic_bytes_per_launch": per_launch,
                "note": "algorithmic bytes = fused minimum: eigen data read once and the ky-transformed "
                        "window rows written once, at the line precision (FP32 in the reduced-precision "
                        "modes).  The kernel is not HBM-bound: its limit is the vector ALU and LDS work "
                        "of the line transforms (valu_* fields)"}
        if dom == 1:
            fl = fft_flops(wl, nw, E) / max(nl, 1)
            roof["valu_tflops"] = fl / avg_s / 1e12
            roof["valu_peak_tflops"] = 78.6   # FP32 or FP64 vector FMA without packed FP32 (not used)
            roof["valu_frac"] = roof["valu_tflops"] / 78.6
    else:
        roof = {"kernel": names[dom], "bound": "hbm", "achieved": None, "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": None, "traffic": None}
    roof["avg_launch_ms"] = avg_s * 1e3
    roof["launches_per_step"] = nl
    roof["timing"] = "HIP events on the library stream around the launch, inside the timed run"
    return roof


def batch_timing(cqpi, synth, torch, count=128, reps=5):
    """Batched workload: `count` independent instances with uneven shapes
    (orbitals {1,3,5}, k-grids {128,256,512}, 16-64 energies, maps
    {32,64,128} cells), autotuned, all run as one captured HIP graph.  Error:
    every map of every instance, mixed vs the FP64 batch, on the device."""
    wls = synth.batch_uneven(count)
    pts = sum(wl.W * wl.W * len(wl.energies) for wl in wls)
    out = {}
    maps = {}
    for prec in ("mixed", "fp64"):
        b = cqpi.Batch(wls, prec)
        t0 = time.perf_counter()
        if prec == "mixed":
            b.autotune(reps=2)
        t_tune = time.perf_counter() - t0
        b.run()  # eager pass, capture, first launch
        ms = min(b.run() for _ in range(reps))
        info = b.info()
        out[prec] = {"graph_ms": ms, "points_per_s": pts / (ms / 1e3), "graph_nodes": info["graph_nodes"],
                     "device_memory_GB": info["gpumem"] / 1e9}
        if prec == "mixed":
            out[prec]["autotune_s"] = t_tune
        maps[prec] = b
    worst = 0.0
    for i in range(count):
        a = maps["mixed"].instance(i).ldos_tensor()
        r = maps["fp64"].instance(i).ldos_tensor()
        m = torch.amax(torch.abs(r), dim=1, keepdim=True)
        den = torch.maximum(torch.abs(r), 1e-3 * m)
        worst = max(worst, float(torch.max(torch.abs(a - r) / den)))
    for b in maps.values():
        b.close()
    return {"workload": "%d instances, seed 1004: orbitals {1,3,5}, k-grid {128,256,512}^2, 16-64 energies "
                        "in [-0.3,0.3], maps {32,64,128}^2 cells, w=2, os=4, V = U(0.1,2) x Hermitian unit" % count,
            "points": pts, "mixed": out["mixed"], "fp64": out["fp64"],
            "speedup_mixed_vs_fp64": out["fp64"]["graph_ms"] / out["mixed"]["graph_ms"],
            "max_rel_mixed_vs_fp64": worst,
            "note": "one hipGraphLaunch per run (graph captured once; 4 parallel branches); tiles autotuned "
                    "per shape before capture"}


def probe(cqpi, synth, torch, workload):
    """One mixed and one FP64 step after a 256 MiB device copy: the run that
    the rocprofv3 --pmc traffic passes profile.  Stage timing is on, which
    keeps each run on one stream, so every stage is one launch over all
    energies (the timed steps split the energies over two streams)."""
    src = torch.ones(32 << 20, dtype=torch.float64, device="cuda")
    dst = src.clone()
    torch.cuda.synchronize()
    del src, dst
    wl = synth.workload(workload)
    for prec in ("mixed", "fp64"):
        with cqpi.engine_for(wl, prec) as q:
            q.set_timing(True)
            q.compute_eigen()
            q.run(wl.energies, wl.eta, host=False)
    torch.cuda.synchronize()
    print("probe done: copy 256 MiB, %s one mixed and one fp64 step" % workload, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="square_3orb")
    ap.add_argument("--energies", type=int, default=0, help="use only the first N energies of the workload")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-grad", action="store_true")
    ap.add_argument("--no-variants", action="store_true", help="skip the Josephson, threshold-list and surface Green's function maps")
    ap.add_argument("--probe", action="store_true")
    ap.add_argument("--stage-events", action="store_true",
                    help="record stage events in every step (one stream, one launch per stage)")
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
    if args.probe:
        probe(cqpi, synth, torch, args.workload)
        return
    wl = synth.workload(args.workload)
    if args.energies:
        wl.energies = wl.energies[:args.energies]
    E = len(wl.energies)

    timed = {}
    for prec in ("fp64", "fp32", "mixed"):
        q = cqpi.engine_for(wl, prec)
        dt, st = time_steps(q, wl, args.steps, args.warmup, torch, args.stage_events)
        timed[prec] = (q, dt, st)
    qm, dt_m, st_m = timed["mixed"]
    q6, dt_6, st_6 = timed["fp64"]
    q3, dt_3, st_3 = timed["fp32"]

    # accuracy: every energy of the reduced-precision paths against the FP64 device path
    r6 = q6.run(wl.energies, wl.eta, host=True)
    rm_ = qm.run(wl.energies, wl.eta, host=True)
    r3 = q3.run(wl.energies, wl.eta, host=True)
    errs = [rel_errors(rm_[e], r6[e]) for e in range(E)]
    errs3 = [rel_errors(r3[e], r6[e]) for e in range(E)]
    acc = {"mixed_vs_fp64_gpu_max_rel": max(e[0] for e in errs),
           "mixed_vs_fp64_gpu_rms_rel": max(e[1] for e in errs),
           "fp32_vs_fp64_gpu_max_rel": max(e[0] for e in errs3),
           "floor": "max|d| / max(|ref|, 1e-3 max|ref|) per energy, max over energies"}

    default = args.workload == "square_3orb" and not args.energies
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
        "data": "synthetic (seeded model, Gaussian x angular-factor tiles)",
        "config": {"workload": "square lattice 3 orbitals, K=512, 51 energies in [-0.1,0.1], "
                               "eta=0.002, 128x128 cells, os=4, w=2 (512x512 pixels)"
                   if default else "%s: %s, %d energies" % (args.workload, wl.name, E),
                   "kpoints": wl.K, "energies": E, "cells": wl.n, "oversamp": wl.oversamp,
                   "window": wl.window, "bands": wl.bands},
        "fp64": {"value": wl.W * wl.W * E / dt_6, "ms_per_step": dt_6 * 1e3,
                 "stages_ms": {k: v[0] for k, v in st_6.items()},
                 "roofline": roofline(wl, q6, st_6, "fp64")},
        "fp32": {"value": wl.W * wl.W * E / dt_3, "ms_per_step": dt_3 * 1e3,
                 "stages_ms": {k: v[0] for k, v in st_3.items()}},
        "speedup_mixed_vs_fp64": dt_6 / dt_m,
        "speedup_fp32_vs_fp64": dt_6 / dt_3,
        "stages_ms": {k: v[0] for k, v in st_m.items()},
        "accuracy": acc,
        "roofline": roofline(wl, qm, st_m, "mixed"),
        "contraction_roofline": roofline(wl, qm, st_m, "mixed", 4),
    }
    for q, _, _ in timed.values():
        q.close()
    if not args.no_batch:
        out["batch"] = batch_timing(cqpi, synth, torch)
    if not args.no_grad:
        out["gradient"] = gradient_timing(cqpi, synth, torch)
    if not args.no_variants:
        out["variants"] = variants_timing(cqpi, synth, torch)
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
