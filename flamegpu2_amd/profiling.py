"""Per-kernel counter table from rocprofv3 runs of `bench.py --traffic-probe`.

    rocprofv3 --kernel-trace --stats --output-format csv -d <T> -o probe -- python -u bench.py --traffic-probe
    rocprofv3 --pmc FETCH_SIZE SQ_WAVES SQ_WAVE_CYCLES SQ_BUSY_CYCLES SQ_INSTS_LDS SQ_LDS_BANK_CONFLICT \\
              --output-format csv -d <A> -o probe -- python -u bench.py --traffic-probe
    rocprofv3 --pmc WRITE_SIZE SQ_INSTS_VALU SQ_INSTS_SALU SQ_INSTS_VMEM_RD SQ_INSTS_VMEM_WR \\
              --output-format csv -d <B> -o probe -- python -u bench.py --traffic-probe
    python -m flamegpu2_amd.profiling <T> <A> <B> <out.md> [file holding the commands, for the summary]

For every kernel of the boids step: mean duration (kernel trace), HBM bytes read and
written per dispatch (FETCH_SIZE / WRITE_SIZE in KiB, each divided by what the counter
reads for the probe's elementwise negation over 1 GiB, whose true traffic is 1 GiB each
way), the resulting GB/s against the 8 TB/s peak, instruction counts, LDS instructions and
bank-conflict cycles, and the mean number of resident waves per CU while the shader
engines are busy.
"""
import collections
import csv
import glob
import os
import sys

HBM_PEAK_GBS = 8000.0
CAL_BYTES = 1 << 30
CUS = 256
SQ_BLOCKS = 32   # shader engines (8 XCDs x 4), each with its own SQ counters


def _rows(d, pattern):
    files = glob.glob(os.path.join(d, "**", pattern), recursive=True)
    if not files:
        raise SystemExit(f"no {pattern} under {d}")
    out = []
    for f in files:
        out += list(csv.DictReader(open(f)))
    return out


def _short(name):
    n = name.split("(")[0]
    for pre in ("void ", "fgx::", "flamegpu::detail::"):
        n = n.replace(pre, "")
    return n


def counters(d, last):
    """{kernel: {counter: mean over the kernel's last `last` dispatches}}."""
    acc = collections.defaultdict(lambda: collections.defaultdict(dict))
    for r in _rows(d, "*counter_collection.csv"):
        per = acc[r["Kernel_Name"]][r["Counter_Name"]]
        i = int(r["Dispatch_Id"])
        per[i] = per.get(i, 0.0) + float(r["Counter_Value"])
    out = {}
    for k, cs in acc.items():
        out[k] = {}
        for c, v in cs.items():
            keep = [v[i] for i in sorted(v)[-last:]]
            out[k][c] = sum(keep) / len(keep)
    return out


def durations(d, last):
    """{kernel: mean duration in us over its last `last` dispatches} from the kernel trace."""
    per = collections.defaultdict(dict)
    for r in _rows(d, "*kernel_trace.csv"):
        per[r["Kernel_Name"]][int(r["Dispatch_Id"])] = (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3
    out = {}
    for k, v in per.items():
        keep = [v[i] for i in sorted(v)[-last:]]
        out[k] = sum(keep) / len(keep)
    return out


def report(trace_dir, dir_a, dir_b, out_md, command="", last=2):
    dur = durations(trace_dir, last)
    ca = counters(dir_a, last)
    cb = counters(dir_b, last)
    cal = [k for k in ca if "neg_kernel" in k]
    if not cal:
        raise SystemExit("calibration kernel (elementwise negation) not found")
    f_scale = ca[cal[0]]["FETCH_SIZE"] * 1024 / CAL_BYTES
    w_scale = cb[cal[0]]["WRITE_SIZE"] * 1024 / CAL_BYTES

    def waves(c):
        busy = c.get("SQ_BUSY_CYCLES", 0.0) / SQ_BLOCKS
        return 4.0 * c.get("SQ_WAVE_CYCLES", 0.0) / busy / CUS if busy else 0.0

    lines = ["# boids spatial-3D, 1,048,576 agents — per-kernel counters (rocprofv3)", ""]
    if command:
        lines += ["Commands:", "", "```", command.strip(), "```", ""]
    lines += [f"Each figure is the mean over a kernel's last {last} dispatches of the probe (steady-state steps).",
              f"Calibration: for the negation over 1 GiB, FETCH_SIZE reads {f_scale:.3f}x and WRITE_SIZE "
              f"{w_scale:.3f}x the true bytes; every kernel's counters are divided by these factors.",
              f"Resident waves per CU while busy: 4 x SQ_WAVE_CYCLES (counted per 4 cycles) / (SQ_BUSY_CYCLES / "
              f"{SQ_BLOCKS} shader engines) / {CUS} CUs, out of 32; the negation kernel reads "
              f"{waves(ca[cal[0]]):.1f}.", "",
              "| kernel | µs | read MB | write MB | HBM GB/s | % of 8 TB/s | VALU insts | SALU insts | LDS insts | "
              "LDS bank-conflict cycles | waves/CU |",
              "|---|---|---|---|---|---|---|---|---|---|---|"]
    names = sorted((k for k in ca if k in dur and k != cal[0] and "FillFunctor" not in k), key=lambda k: -dur[k])
    for k in names:
        a, b = ca[k], cb.get(k, {})
        us = dur[k]
        rd = a.get("FETCH_SIZE", 0.0) * 1024 / f_scale
        wr = b.get("WRITE_SIZE", 0.0) * 1024 / w_scale
        gbs = (rd + wr) / (us * 1e-6) / 1e9 if us else 0.0
        lines.append(f"| `{_short(k)}` | {us:.1f} | {rd / 1e6:.1f} | {wr / 1e6:.1f} | {gbs:.0f} | "
                     f"{100 * gbs / HBM_PEAK_GBS:.1f} | {b.get('SQ_INSTS_VALU', 0):.3g} | {b.get('SQ_INSTS_SALU', 0):.3g} | "
                     f"{a.get('SQ_INSTS_LDS', 0):.3g} | {a.get('SQ_LDS_BANK_CONFLICT', 0):.3g} | {waves(a):.1f} |")
    text = "\n".join(lines) + "\n"
    with open(out_md, "w") as f:
        f.write(text)
    return text


if __name__ == "__main__":
    if len(sys.argv) < 5:
        raise SystemExit(__doc__)
    cmd = open(sys.argv[5]).read() if len(sys.argv) > 5 else __doc__.split("\n\n")[1]
    print(report(*sys.argv[1:5], command=cmd))
