"""Summarise rocprofv3 --pmc runs of `bench.py --probe` into HBM traffic and
MFMA figures per kernel (markdown with a JSON block that bench.py reads).

    python -m cqpi.pmcsummary FETCH.csv WRITE.csv [MFMA.csv] [--energies N] > summary.md
    python -m cqpi.pmcsummary --stats KERNEL_STATS.csv > summary.md

The probe first copies 256 MiB with a plain device copy (the runtime's
copyBuffer kernel), then runs one mixed-precision and one FP64 step.
FETCH_SIZE and WRITE_SIZE are in KiB per dispatch; each is scaled by the
factor that makes the copy read and write exactly 256 MiB.  Kernels with two
dispatches are labelled by the probe's order (mixed first, then fp64).
MFMA busy is SQ_VALU_MFMA_BUSY_CYCLES / (GRBM_GUI_ACTIVE / XCDs * SIMDs).
"""
from __future__ import annotations

import csv
import json
import re
import sys
from collections import OrderedDict

COPY_BYTES = 256 << 20
XCDS, SIMDS = 8, 1024   # MI355X: 8 XCDs, 256 CUs x 4 SIMDs


def short_name(name: str) -> str:
    """k_name<args> from a demangled or mangled kernel name; the complex line
    type of the transform kernels is shown as f32 / f64."""
    name = name.replace("HIP_vector_type<float, 2u>", "f32").replace("HIP_vector_type<double, 2u>", "f64")
    i = name.find("::k_")
    if i >= 0:
        j = i + 2
        k = j
        while k < len(name) and (name[k].isalnum() or name[k] == "_"):
            k += 1
        out = name[j:k]
        if k < len(name) and name[k] == "<":
            depth, m = 0, k
            while m < len(name):
                depth += name[m] == "<"
                depth -= name[m] == ">"
                m += 1
                if depth == 0:
                    break
            out += name[k:m]
        return out
    m = re.search(r"_GLOBAL__N_1(\d+)", name)
    if m:
        # anonymous namespace "_GLOBAL__N_1", then <length><identifier>
        rest = name[m.end():]
        ln = int(m.group(1))
        ident, tail = rest[:ln], rest[ln:]
        if not ident.startswith("k_"):
            return name[:40]
        args = []
        if tail.startswith("I"):
            end = tail.find("EEv")  # template arguments end before the void return type
            for t in re.findall(r"HIP_vector_typeI([fd])Lj2EE|Li(\d+)E", tail[: end + 1 if end > 0 else None]):
                if t[0]:
                    args.append("f32" if t[0] == "f" else "f64")
                elif t[1]:
                    args.append(t[1])
        return ident + ("<%s>" % ", ".join(args) if args else "")
    return name[:40]


def read(path, counter):
    rows = []
    with open(path) as f:
        for r in csv.DictReader(f):
            if r["Counter_Name"] == counter:
                rows.append((r["Kernel_Name"], int(r["Grid_Size"]), float(r["Counter_Value"]),
                             int(r["Dispatch_Id"])))
    rows.sort(key=lambda x: x[3])
    return rows


def per_kernel(rows):
    """OrderedDict label -> value, labels 'name' or 'name (mixed|fp64)'."""
    groups = OrderedDict()
    for name, grid, val, _ in rows:
        if "rocclr" in name or "at::native" in name:
            continue
        groups.setdefault(short_name(name), []).append(val)
    out = OrderedDict()
    for k, v in groups.items():
        if len(v) == 2:
            out[k + " (mixed)"] = v[0]
            out[k + " (fp64)"] = v[1]
        else:
            out[k] = sum(v) / len(v)
    return out


def calibration(rows, expected_kib):
    copies = [val for name, grid, val, _ in rows if "copyBuffer" in name]
    big = max(copies) if copies else 0.0
    return expected_kib / big if big else None


def stats(path):
    """Per-kernel time table of a rocprofv3 --kernel-trace --stats run."""
    with open(path) as f:
        rows = list(csv.DictReader(f))
    total = sum(float(r["TotalDurationNs"]) for r in rows)
    print("| kernel | calls | avg us | min us | max us | share % |")
    print("|---|---|---|---|---|---|")
    for r in rows:
        print("| `%s` | %s | %.1f | %.1f | %.1f | %.1f |" % (
            short_name(r["Name"]) if "k_" in r["Name"] else r["Name"][:40], r["Calls"],
            float(r["AverageNs"]) / 1e3, float(r["MinNs"]) / 1e3, float(r["MaxNs"]) / 1e3,
            100 * float(r["TotalDurationNs"]) / total))


def main(argv):
    if argv and argv[0] == "--stats":
        stats(argv[1])
        return
    energies = None
    if "--energies" in argv:
        k = argv.index("--energies")
        energies = int(argv[k + 1])
        argv = argv[:k] + argv[k + 2:]
    fetch = read(argv[0], "FETCH_SIZE")
    write = read(argv[1], "WRITE_SIZE")
    ff = calibration(fetch, COPY_BYTES / 1024)
    fw = calibration(write, COPY_BYTES / 1024)
    F, Wr = per_kernel(fetch), per_kernel(write)
    mf = {}
    if len(argv) > 2:
        busy = per_kernel(read(argv[2], "SQ_VALU_MFMA_BUSY_CYCLES"))
        act = per_kernel(read(argv[2], "GRBM_GUI_ACTIVE"))
        mops = per_kernel(read(argv[2], "SQ_INSTS_VALU_MFMA_MOPS_F16"))
        for k in busy:
            if busy[k] > 0 and act.get(k):
                mf[k] = {"mfma_busy": busy[k] / (act[k] / XCDS * SIMDS),
                         "mfma_f16_flops": mops.get(k, 0.0) * 512}
    table = OrderedDict()
    for k in F:
        fb = F[k] * 1024 * ff
        wb = Wr.get(k, 0.0) * 1024 * fw
        table[k] = {"fetch_bytes": fb, "write_bytes": wb, "bytes_per_launch": fb + wb}
        table[k].update(mf.get(k, {}))
    print("| kernel | HBM read MB | HBM write MB | MFMA busy | MFMA f16 GFLOP |")
    print("|---|---|---|---|---|")
    for k, r in table.items():
        print("| `%s` | %.1f | %.1f | %s | %s |" % (
            k, r["fetch_bytes"] / 1e6, r["write_bytes"] / 1e6,
            "%.1f%%" % (100 * r["mfma_busy"]) if "mfma_busy" in r else "",
            "%.1f" % (r["mfma_f16_flops"] / 1e9) if "mfma_f16_flops" in r else ""))
    print()
    print("Copy calibration: FETCH_SIZE x %.3f, WRITE_SIZE x %.3f (256 MiB copy read and written)." % (ff, fw))
    print()
    print("```json")
    doc = {"calibration": {"fetch": ff, "write": fw}, "kernels": table}
    if energies:
        doc["energies"] = energies  # energies of each probe launch
    print(json.dumps(doc, indent=1))
    print("```")


if __name__ == "__main__":
    main(sys.argv[1:])
