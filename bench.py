"""Continuum-LDOS throughput on one MI355X: mixed precision vs FP32 vs FP64.

A step is one whole pass of the device chain over the workload: per-k
eigensolve of H(k), then for every energy the resolvent, the k->R transform,
the T-matrix and the continuum contraction into the W x W map (maps stay in
HBM).  Default workload: the 3-orbital square lattice (K=512, 51 energies,
128x128 cells, os=4, w=2 -> 512x512 pixels), synthetic model drawn from a
fixed seed.

`value` is the mixed-precision path: split-fp16 MFMA contraction with FP32
accumulation, FP32 resolvent and transform lines, FP64 eigensolve, T-matrix,
periodic term and G0 window.  The same line reports the FP32 and FP64 paths,
their ratios, the accuracy per energy against the FP64 path and the FP64 CPU
oracle, the roofline of the dominant kernels, an end-to-end rate that
includes moving the maps to the host, the single-orbital workload (K=256, 11
energies, 64x64 cells, os=8, w=1), the 10-band hexagonal workload (K=1024,
101 energies, 256x256 cells, os=6, w=3), the batched workload (128 uneven
instances in one HIP graph), the gradient workload, the Josephson,
threshold-list and surface variants and the CPU baseline.  Prints one JSON
line.

  python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu] [--no-batch]
                  [--no-grad] [--no-variants] [--no-hex] [--no-config1] [--workload NAME] [--energies N]
  python bench.py --probe [--workload NAME] [--energies N]
                                 one mixed and one FP64 step after a 256 MiB
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


def fused_transform_bytes(wl, E):
    """Algorithmic HBM bytes of the resolvent + k->R transform for E energies
    at the fused minimum: the eigen data read once, K^2 (16 b^2 + 8 b), and
    the G0(R) window written once per energy, 16 b^2 nk^2 with nk = n + 2w + 2
    the window side of the reference's copy (src/qpihipkernel.cc:383-393).
    The intermediate rows between the two passes are not algorithmic bytes:
    they are overhead traffic."""
    K, b = wl.K, wl.bands
    nk = wl.n + 2 * wl.window + 2
    return K * K * (16 * b * b + 8 * b) + E * 16 * b * b * nk * nk


def transform_valu_flops(wl, nw, E, pairs=None):
    """FP32/FP64 arithmetic of the two transform passes for E energies: the
    radix-8 line transforms (about 5 N log2 N flops per line of N = K points;
    one line per transformed orbital pair per kx row in the first pass and per
    window column in the second) and the resolvent (8 flops per transformed
    pair and band per k-point).  pairs: the transformed pairs, b^2, or
    b (b + 1) / 2 where the run used the pair symmetry."""
    K, b = wl.K, wl.bands
    npair = b * b if pairs is None else pairs
    lg = int(np.log2(K))
    line = 5 * K * lg
    return E * (K * npair * line + nw * npair * line + K * K * npair * b * 8)


def rel_errors(x, ref, floor=1e-3):
    m = np.max(np.abs(ref))
    den = np.maximum(np.abs(ref), floor * m)
    return float(np.max(np.abs(x - ref) / den)), float(np.linalg.norm(x - ref) / np.linalg.norm(ref))


def host_cpu():
    """Model, sockets, cores and threads of this host (lscpu, else /proc/cpuinfo)."""
    import subprocess
    info = {}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            k, _, v = line.partition(":")
            info[k.strip()] = v.strip()
    except (OSError, subprocess.SubprocessError):
        pass
    model = info.get("Model name")
    if not model:
        try:
            for line in open("/proc/cpuinfo"):
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
        except OSError:
            model = None
    def num(key):
        try:
            return int(info[key])
        except (KeyError, ValueError):
            return None
    return {"model": model, "sockets": num("Socket(s)"), "cores_per_socket": num("Core(s) per socket"),
            "threads_per_core": num("Thread(s) per core"), "cpus": num("CPU(s)"),
            "affinity_cpus": len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else None}


def load_oracle():
    """The FP64 CPU oracle, its OpenMP threads bound close (OMP_PROC_BIND is
    read when the runtime starts)."""
    os.environ.setdefault("OMP_PROC_BIND", "close")
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    return oracle


def cpu_baseline(wl, nsample):
    """The FP64 oracle (C restatement of the reference algorithm with the
    reference loop nest for the contraction, OpenMP over cells) timed on this
    host: eigensolve once plus nsample energies, extrapolated linearly to all
    energies.  Also returns the oracle maps of the sampled energies."""
    oracle = load_oracle()

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
    nt = oracle.num_threads()
    cpu = host_cpu()
    info = {"value": wl.points / total, "unit": "points/s", "cores": nt, "kind": "port",
            "sample": "oracle eigensolve + %d of %d energies at full size, extrapolated linearly "
                      "(%.2f s eigensolve, %.2f s/energy, 1 process, %d OpenMP threads, OMP_PROC_BIND=%s)"
                      % (nsample, E, t_eig, t_e, nt, os.environ.get("OMP_PROC_BIND")),
            "host": cpu, "mpi_ranks": 1, "omp_proc_bind": os.environ.get("OMP_PROC_BIND")}
    return info, np.array(maps)


def time_steps(q, wl, steps, warmup, torch, stage_events=False, eigen=True, stage_reps=5):
    """Wall time per step (eigensolve + all energies) over `steps` steps after
    `warmup`, then the stage times of `stage_reps` more steps (median per
    stage).  stage_events: every step records stage events, which keeps each
    run on one stream (one launch per stage: what a kernel-trace profile of
    the stages should see).  eigen=False: no eigensolve in the step (the
    surface Green's function does not use it)."""
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
    runs = []
    for _ in range(max(stage_reps, 1)):
        step()
        runs.append(q.stage_times())
    st = {k: (float(np.median([r[k][0] for r in runs])), runs[0][k][1]) for k in runs[0]}
    q.set_timing(stage_events)   # (with stage_events the later accuracy runs stay on one stream too)
    return dt, st


def variants_timing(cqpi, synth, torch, steps=5, warmup=3, threshold=0.01):
    """The reference's other two contraction kinds at the 3-orbital geometry
    (K=512, 51 energies, 128x128 cells, os=4, w=2), whole chain on the device:
    the Josephson map of a superconducting model (3 orbitals x particle-hole,
    6 bands; FP64, FP32 and mixed) and the threshold-list map (FP64) of the 3-orbital
    workload with the driver's list rule at the given threshold; and the LDOS
    map from the surface Green's function of a layered 3-orbital model."""
    out = {}
    wlj = synth.bdg_square(3, K=512, n=128, E=51, seed=1006, w=2, os_=4, eta=0.002, elo=-0.1, ehi=0.1)
    jos = {"workload": "superconducting square lattice, 3 orbitals x particle-hole (6 bands), K=512, "
                       "51 energies in [-0.1,0.1], eta=0.002, 128x128 cells, os=4, w=2; tip deltat=0.001, "
                       "etat=eta"}
    maps = {}
    for prec in ("fp64", "fp32", "mixed"):
        q = cqpi.engine_for(wlj, prec)
        q.set_josephson(True, deltat=0.001)
        dt, st = time_steps(q, wlj, steps, warmup, torch)
        jos[prec] = {"ms_per_step": dt * 1e3, "points_per_s": wlj.points / dt,
                     "stages_ms": {k: v[0] for k, v in st.items()}}
        maps[prec] = q.run(wlj.energies, wlj.eta, host=True)
        q.close()
    for prec in ("fp32", "mixed"):
        jos[prec + "_vs_fp64_gpu_max_rel"] = max(rel_errors(maps[prec][e], maps["fp64"][e])[0]
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
    # the same forward with every orbital pair transformed (the adjoint's
    # transforms always carry every pair)
    q.set_pair_symmetry(False)
    fa = []
    for it in range(reps + 1):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        q.compute_eigen()
        q.run(wl.energies, wl.eta, host=False)
        torch.cuda.synchronize()
        if it:
            fa.append(time.perf_counter() - t0)
    q.close()
    f, b, f_all = float(np.median(fw)), float(np.median(bw)), float(np.median(fa))
    return {"workload": "square lattice 3 orbitals, K=512, 21 energies, 64x64 cells, os=4, w=2; "
                        "L2 map loss; gradients w.r.t. %d hopping amplitudes (re, im), mu and V" % len(wl.hops),
            "forward_ms": f * 1e3, "backward_ms": b * 1e3, "backward_over_forward": b / f,
            "gradient_over_forward": (f + b) / f,
            "forward_every_pair_ms": f_all * 1e3, "gradient_over_forward_every_pair": (f_all + b) / f_all,
            "note": "FP64; forward = eigensolve + all energies (the transform with the pair symmetry of "
                    "the real hoppings; forward_every_pair_ms: with it switched off); backward = dL/drho on "
                    "the device + adjoint; maps and dL/drho stay in HBM"}


TRAFFIC_MD = {"square_3orb": os.path.join(ROOT, "profiles", "pmc_traffic_square3orb_K512_E51.md"),
              "hex_5d_spin": os.path.join(ROOT, "profiles", "pmc_traffic_hex10band_K1024_E4.md")}


def measured_traffic(kernel, precision, workload="square_3orb"):
    """HBM bytes per launch of `kernel` (name prefix) in the `precision` step
    of the workload's probe, from the committed rocprofv3 --pmc summary
    (FETCH_SIZE + WRITE_SIZE, copy-calibrated), divided by the probe's
    energies (the summary records them) so that it can be scaled to another
    energy count; returns (bytes per launch, energies of that launch) or None."""
    try:
        txt = open(TRAFFIC_MD[workload]).read()
    except (OSError, KeyError):
        return None
    if "```json" not in txt:
        return None
    table = json.loads(txt.split("```json", 1)[1].split("```", 1)[0])
    line = "f64" if precision == "fp64" else "f32"   # line type of the transform kernels
    for name, row in table.get("kernels", {}).items():
        label = name.rsplit(" (", 1)[1].rstrip(")") if " (" in name else None
        typed = "f32" in name or "f64" in name
        if name.startswith(kernel) and label in (None, precision) and (not typed or line in name):
            return row.get("bytes_per_launch"), table.get("energies")
    return None


def stage_launch_s(st, name):
    ms, nl = st[name]
    return ms / max(nl, 1) / 1e3, nl


def transform_roofline(wl, q, st, precision, workload):
    """Roofline of the fused resolvent + k->R transform (the two passes,
    k_res_ky_s and k_kx_s or, with the pair symmetry, k_kx_sym; one launch
    each per run with stage events), against
    the fused minimum bytes; `traffic` is what the two kernels moved in the
    committed profile (the ky-transformed rows go out to HBM and back between
    the passes: overhead traffic).  Their limit is the vector ALU and LDS work
    of the line transforms, reported as valu_*."""
    E = len(wl.energies)
    t1, n1 = stage_launch_s(st, "resolvent+transform_ky")
    t2, n2 = stage_launch_s(st, "transform_kx")
    runs = max(n1, n2, 1)                    # launches of each pass per run (energy chunks)
    total = t1 * n1 + t2 * n2                # both passes over all E energies
    per_run = fused_transform_bytes(wl, E)
    achieved = per_run / total / 1e9
    t = total / runs
    traffic = None
    m1 = measured_traffic("k_res_ky_s", precision, workload)
    m2 = measured_traffic("k_kx_s", precision, workload)
    if m1 and m2 and m1[1] and m2[1]:
        traffic = (m1[0] / m1[1] + m2[0] / m2[1]) * E
    nw = 2 * q.window_half() + 1
    sym = q.pair_symmetry()[1]
    b = wl.bands
    fl = transform_valu_flops(wl, nw, E, b * (b + 1) // 2 if sym else b * b)
    line = "FP64" if precision == "fp64" else "FP32"
    return {"kernel": "resolvent+transform (k_res_ky_s + %s, %s lines)" % ("k_kx_sym" if sym else "k_kx_s", line),
            "pair_symmetry": bool(sym), "bound": "hbm",
            "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic / runs if traffic else None, "algorithmic_bytes_per_launch": per_run / runs,
            "avg_launch_ms": t * 1e3, "pass_ms": [t1 * 1e3, t2 * 1e3], "launches_per_step": runs,
            "valu_tflops": fl / total / 1e12, "valu_peak_tflops": 78.6, "valu_frac": fl / total / 1e12 / 78.6,
            "note": "achieved = fused minimum (eigen data read once + the G0 window, 16 b^2 nk^2 per energy, "
                    "written once) / the time of both passes; traffic (HBM bytes of both kernels, rocprof "
                    "FETCH_SIZE + WRITE_SIZE scaled to these energies) also moves the ky-transformed rows "
                    "between the passes.  The passes are bound by the vector ALU and LDS work of the line "
                    "transforms (valu_*; FP32 and FP64 both 78.6 TFLOP/s without packed FP32)",
            "timing": "HIP events on the library stream around each launch, inside the bench run"}


def contraction_roofline(wl, st, precision, workload, kernel):
    """MFMA (mixed) or vector-ALU (FP32, FP64) roofline of the contraction
    with the useful flops per point F of flops_per_point."""
    E = len(wl.energies)
    t, nl = stage_launch_s(st, "contraction")
    per_launch = wl.W * wl.W * E / max(nl, 1) * flops_per_point(wl)
    achieved = per_launch / t / 1e12
    peak = PEAK_TFLOPS[precision]
    m = measured_traffic(kernel, precision, workload)
    return {"kernel": kernel, "bound": "mfma" if precision == "mixed" else "valu", "achieved": achieved,
            "peak": peak, "unit": "TFLOP/s", "frac": achieved / peak,
            "traffic": m[0] / m[1] * E / max(nl, 1) if m and m[1] else None,
            "flops_per_point": flops_per_point(wl), "avg_launch_ms": t * 1e3, "launches_per_step": nl,
            "timing": "HIP events on the library stream around the launch, inside the bench run"}


def per_energy_errors(a, r, torch):
    """[max_rel, rms_rel] per energy of maps a against r (E x W*W, device or host)."""
    if not torch.is_tensor(a):
        return [list(rel_errors(a[e], r[e])) for e in range(r.shape[0])]
    m = torch.amax(torch.abs(r), dim=1, keepdim=True)
    den = torch.maximum(torch.abs(r), 1e-3 * m)
    mx = torch.amax(torch.abs(a - r) / den, dim=1)
    rms = torch.linalg.vector_norm(a - r, dim=1) / torch.linalg.vector_norm(r, dim=1)
    return [[float(x), float(y)] for x, y in zip(mx.cpu(), rms.cpu())]


def hex_timing(cqpi, synth, torch, steps, warmup, with_oracle, cells=16, oracle_energy=50):
    """The 10-band hexagonal workload (5 d orbitals x spin, SOC, full 10x10 V;
    K=1024, 101 energies, 256x256 cells, os=6, w=3): FP64 and mixed, each step
    the eigensolve and all 101 energies; errors of every mixed map against the
    FP64 one on the device; with_oracle: both against the FP64 oracle on a
    block of cells x cells cells at the centre of one energy's map (the whole
    map is beyond the reference loop nest on the host)."""
    wl = synth.workload("hex_5d_spin")
    E = len(wl.energies)
    out = {"workload": "hexagonal lattice, 5 d orbitals x spin (10 bands, SOC lambda=0.05), K=1024, 101 energies "
                       "in [-0.2,0.2], eta=0.001, full Hermitian 10x10 V, 256x256 cells, os=6, w=3 "
                       "(1536x1536 pixels), spin blocks", "points": wl.points}
    qs = {}
    for prec in ("fp64", "mixed"):
        q = cqpi.engine_for(wl, prec)
        dt, st = time_steps(q, wl, steps, warmup, torch, stage_reps=2)
        q.run(wl.energies, wl.eta, host=False)
        kern = "k_cmix_stq" if prec == "mixed" else "k_contract_st"
        t_eig, _ = stage_launch_s(st, "eigensolve")
        K, b = wl.K, wl.bands
        eig_bytes = K * K * (16 * b * b + 8 * b)
        out[prec] = {"ms_per_step": dt * 1e3, "points_per_s": wl.points / dt,
                     "stages_ms": {k: v[0] for k, v in st.items()},
                     "contraction_roofline": contraction_roofline(wl, st, prec, "hex_5d_spin", kern),
                     "transform_roofline": transform_roofline(wl, q, st, prec, "hex_5d_spin"),
                     "eigensolve_roofline": {"kernel": "k_eigen_g<10>", "bound": "hbm",
                                             "achieved": eig_bytes / t_eig / 1e9, "peak": HBM_PEAK_GBS,
                                             "unit": "GB/s", "frac": eig_bytes / t_eig / 1e9 / HBM_PEAK_GBS,
                                             "algorithmic_bytes_per_launch": eig_bytes,
                                             "note": "writes the eigen data once per model; bound by the FP64 "
                                                     "vector ALU work of the Jacobi sweeps"}}
        qs[prec] = q
    out["speedup_mixed_vs_fp64"] = out["fp64"]["ms_per_step"] / out["mixed"]["ms_per_step"]
    a, r = qs["mixed"].ldos_tensor(), qs["fp64"].ldos_tensor()
    pe = per_energy_errors(a, r, torch)
    out["accuracy"] = {"mixed_vs_fp64_gpu_max_rel": max(x[0] for x in pe),
                       "mixed_vs_fp64_gpu_rms_rel": max(x[1] for x in pe),
                       "mixed_vs_fp64_gpu_per_energy": pe}
    if with_oracle:
        oracle = load_oracle()
        t0 = time.perf_counter()
        ev, V = oracle.calchamiltonian(wl.hops, wl.bands, wl.efermi, wl.K)
        om = wl.energies[oracle_energy]
        gR = oracle.realspace(oracle.greens_k(ev, V, om, wl.eta))
        T = oracle.tmatrix(wl.V, gR[0, 0])
        c0 = wl.n // 2 - cells // 2
        ref = oracle.wannierldos_cells(gR, T, wl.tiles, wl.n, wl.window, wl.oversamp,
                                       (c0, c0 + cells, c0, c0 + cells), wl.maxband, wl.spin)
        del ev, V, gR
        os_ = wl.oversamp
        sl = slice(c0 * os_, (c0 + cells) * os_)
        blk = ref.reshape(wl.W, wl.W)[sl, sl]
        acc = out["accuracy"]
        for prec in ("fp64", "mixed"):
            got = qs[prec].ldos_tensor()[oracle_energy].cpu().numpy().reshape(wl.W, wl.W)[sl, sl]
            mx, rms = rel_errors(got.ravel(), blk.ravel())
            acc["%s_vs_oracle_max_rel" % prec] = mx
            acc["%s_vs_oracle_rms_rel" % prec] = rms
        acc["oracle_sample"] = ("energy index %d (omega=%.4f), cells [%d,%d)^2 of 256^2 (%d x %d pixels), "
                                "oracle eigensolve + transform + reference loop nest on those cells, %.1f s"
                                % (oracle_energy, om, c0, c0 + cells, cells * os_, cells * os_,
                                   time.perf_counter() - t0))
    for q in qs.values():
        q.close()
    return out


def config1_timing(cqpi, synth, torch, steps, warmup, with_oracle):
    """The single-orbital workload (K=256, 11 energies in [-0.5,0.5], eta=0.01,
    64x64 cells, os=8, w=1): FP64, FP32 and mixed steps (eigensolve + all
    energies) with stage times; its contraction (88 useful flops per point) is
    bound by HBM: the algorithmic bytes per energy are the map written once
    (8 W^2) plus the G0 window read once (16 b^2 nw^2).  with_oracle: every
    energy of every map against the FP64 oracle (the whole workload takes the
    oracle about a second)."""
    wl = synth.workload("square_1orb")
    E = len(wl.energies)
    out = {"workload": "square lattice 1 orbital (t=-1, t'=0.3, mu=0.5), K=256, 11 energies in [-0.5,0.5], "
                       "eta=0.01, 64x64 cells, os=8, w=1 (512x512 pixels), V=1.0", "points": wl.points}
    qs = {}
    for prec in ("fp64", "fp32", "mixed"):
        q = cqpi.engine_for(wl, prec)
        dt, st = time_steps(q, wl, steps, warmup, torch)
        q.run(wl.energies, wl.eta, host=False)
        t, nl = stage_launch_s(st, "contraction")
        nw = 2 * q.window_half() + 1
        per_launch = E * (8 * wl.W * wl.W + 16 * wl.bands ** 2 * nw * nw) / max(nl, 1)
        out[prec] = {"ms_per_step": dt * 1e3, "points_per_s": wl.points / dt,
                     "stages_ms": {k: v[0] for k, v in st.items()},
                     "contraction_roofline": {
                         "bound": "hbm", "achieved": per_launch / t / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": per_launch / t / 1e9 / HBM_PEAK_GBS, "algorithmic_bytes_per_launch": per_launch,
                         "avg_launch_ms": t * 1e3, "launches_per_step": nl,
                         "flops_per_point": flops_per_point(wl),
                         "note": "map written once (8 W^2 B) + G0 window read once (16 b^2 nw^2 B) per energy"},
                     "transform_roofline": transform_roofline(wl, q, st, prec, "square_1orb")}
        qs[prec] = q
    for prec in ("fp32", "mixed"):
        out["speedup_%s_vs_fp64" % prec] = out["fp64"]["ms_per_step"] / out[prec]["ms_per_step"]
    acc = {}
    r = qs["fp64"].ldos_tensor()
    for prec in ("fp32", "mixed"):
        pe = per_energy_errors(qs[prec].ldos_tensor(), r, torch)
        acc["%s_vs_fp64_gpu_max_rel" % prec] = max(x[0] for x in pe)
    if with_oracle:
        oracle = load_oracle()
        t0 = time.perf_counter()
        ref = oracle.cldos_run(wl)
        acc["oracle_s"] = time.perf_counter() - t0
        for prec in ("fp64", "fp32", "mixed"):
            pe = per_energy_errors(qs[prec].ldos_tensor().cpu().numpy(), ref, torch)
            acc["%s_vs_oracle_max_rel" % prec] = max(x[0] for x in pe)
            acc["%s_vs_oracle_rms_rel" % prec] = max(x[1] for x in pe)
            acc["%s_vs_oracle_per_energy" % prec] = pe
    out["accuracy"] = acc
    for q in qs.values():
        q.close()
    return out


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


def probe(cqpi, synth, torch, workload, energies=0):
    """One mixed and one FP64 step after a 256 MiB device copy: the run that
    the rocprofv3 --pmc traffic passes profile.  Stage timing is on, which
    keeps each run on one stream, so every stage is one launch over all
    energies (the timed steps split the energies over two streams)."""
    src = torch.ones(32 << 20, dtype=torch.float64, device="cuda")
    dst = src.clone()
    torch.cuda.synchronize()
    del src, dst
    wl = synth.workload(workload)
    if energies:
        wl.energies = wl.energies[:energies]
    for prec in ("mixed", "fp64"):
        with cqpi.engine_for(wl, prec) as q:
            q.set_timing(True)
            q.compute_eigen()
            q.run(wl.energies, wl.eta, host=False)
    torch.cuda.synchronize()
    print("probe done: copy 256 MiB, %s (%d energies) one mixed and one fp64 step"
          % (workload, len(wl.energies)), flush=True)


def end_to_end(q, wl, steps, torch):
    """The mixed step with the maps moved to the host as well (cqpi_run with
    a host buffer: the driver's path up to the idl write)."""
    host = np.empty((len(wl.energies), wl.W * wl.W))
    q.compute_eigen()
    q.run(wl.energies, wl.eta, host=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        q.compute_eigen()
        q.run(wl.energies, wl.eta, host=True)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    del host
    return {"ms_per_step": dt * 1e3, "points_per_s": wl.points / dt,
            "map_bytes": wl.points * 8,
            "note": "mixed step plus the copy of all E maps (FP64) to pageable host memory; value leaves them in HBM"}


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
    ap.add_argument("--no-hex", action="store_true", help="skip the 10-band hexagonal workload")
    ap.add_argument("--no-config1", action="store_true", help="skip the single-orbital workload")
    ap.add_argument("--no-variants", action="store_true", help="skip the Josephson, threshold-list and surface Green's function maps")
    ap.add_argument("--probe", action="store_true")
    ap.add_argument("--stage-events", action="store_true",
                    help="record stage events in every step (one stream, one launch per stage)")
    ap.add_argument("--cpu-energies", type=int, default=2)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU (--gpus 1): this build runs on one GPU; the energy sharding over "
              "several GPUs is a design only", file=sys.stderr)
        sys.exit(2)

    import torch

    import cqpi
    from cqpi import synth

    if cqpi.device_count() < 1:
        raise SystemExit("no HIP device")
    if args.probe:
        probe(cqpi, synth, torch, args.workload, args.energies)
        return
    wl = synth.workload(args.workload)
    if args.energies:
        wl.energies = wl.energies[:args.energies]
    E = len(wl.energies)
    print("bench: %s, %d energies, steps %d warmup %d" % (args.workload, E, args.steps, args.warmup),
          file=sys.stderr, flush=True)

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
    errs = per_energy_errors(rm_, r6, torch)
    errs3 = per_energy_errors(r3, r6, torch)
    acc = {"mixed_vs_fp64_gpu_max_rel": max(e[0] for e in errs),
           "mixed_vs_fp64_gpu_rms_rel": max(e[1] for e in errs),
           "fp32_vs_fp64_gpu_max_rel": max(e[0] for e in errs3),
           "mixed_vs_fp64_gpu_per_energy": errs,
           "floor": "per energy: max|d| / max(|ref|, 1e-3 max|ref|) and ||d|| / ||ref||; maxima over energies"}

    default = args.workload == "square_3orb" and not args.energies
    ws = args.workload if args.workload in TRAFFIC_MD else "square_3orb"
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
        "dtype": "mixed: split-fp16 (hi+lo) MFMA contraction with f32 accumulation; f32 resolvent and "
                 "transform lines; f64 eigensolve, T-matrix, periodic term and G0 window",
        "data": "synthetic (seeded model, Gaussian x angular-factor tiles)",
        "config": {"workload": "square lattice 3 orbitals, K=512, 51 energies in [-0.1,0.1], "
                               "eta=0.002, 128x128 cells, os=4, w=2 (512x512 pixels)"
                   if default else "%s: %s, %d energies" % (args.workload, wl.name, E),
                   "kpoints": wl.K, "energies": E, "cells": wl.n, "oversamp": wl.oversamp,
                   "window": wl.window, "bands": wl.bands},
        "fp64": {"value": wl.points / dt_6, "ms_per_step": dt_6 * 1e3,
                 "stages_ms": {k: v[0] for k, v in st_6.items()},
                 "roofline": transform_roofline(wl, q6, st_6, "fp64", ws),
                 "contraction_roofline": contraction_roofline(wl, st_6, "fp64", ws, "k_contract_st")},
        "fp32": {"value": wl.points / dt_3, "ms_per_step": dt_3 * 1e3,
                 "stages_ms": {k: v[0] for k, v in st_3.items()}},
        "speedup_mixed_vs_fp64": dt_6 / dt_m,
        "speedup_fp32_vs_fp64": dt_6 / dt_3,
        "stages_ms": {k: v[0] for k, v in st_m.items()},
        "accuracy": acc,
        "roofline": transform_roofline(wl, qm, st_m, "mixed", ws),
        "contraction_roofline": contraction_roofline(wl, st_m, "mixed", ws, "k_cmix_tf"),
        "end_to_end": end_to_end(qm, wl, max(3, args.steps // 4), torch),
    }
    for q, _, _ in timed.values():
        q.close()
    if not args.no_config1 and default:
        print("bench: single-orbital workload", file=sys.stderr, flush=True)
        out["config1_1orb"] = config1_timing(cqpi, synth, torch, max(5, args.steps // 2), 5, not args.no_cpu)
    if not args.no_hex and default:
        print("bench: 10-band hexagonal workload", file=sys.stderr, flush=True)
        out["hex_10band"] = hex_timing(cqpi, synth, torch, 2, 1, not args.no_cpu)
    if not args.no_batch:
        print("bench: batch of 128 instances", file=sys.stderr, flush=True)
        out["batch"] = batch_timing(cqpi, synth, torch)
    if not args.no_grad:
        print("bench: gradient", file=sys.stderr, flush=True)
        out["gradient"] = gradient_timing(cqpi, synth, torch)
    if not args.no_variants:
        print("bench: variants", file=sys.stderr, flush=True)
        out["variants"] = variants_timing(cqpi, synth, torch)
    if not args.no_cpu:
        print("bench: CPU baseline", file=sys.stderr, flush=True)
        info, maps = cpu_baseline(wl, args.cpu_energies)
        out["cpu_baseline"] = info
        e6 = per_energy_errors(r6[:len(maps)], maps, torch)
        em = per_energy_errors(rm_[:len(maps)], maps, torch)
        acc["fp64_gpu_vs_oracle_max_rel"] = max(e[0] for e in e6)
        acc["mixed_vs_oracle_max_rel"] = max(e[0] for e in em)
        acc["mixed_vs_oracle_rms_rel"] = max(e[1] for e in em)
        acc["vs_oracle_per_energy"] = {"energies": list(range(len(maps))), "fp64": e6, "mixed": em}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
