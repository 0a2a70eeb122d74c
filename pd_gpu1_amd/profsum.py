"""Summarise rocprofv3 CSV output of the force kernels into a short text report
and the HBM traffic per launch that bench.py reports as roofline.traffic.

    python -m pd_gpu1_amd.profsum PROFDIR OUT.txt KERNEL[,KERNEL...] \\
        --copy-bytes 1073741824 --command "..." [--command "..."]

PROFDIR holds the rocprofv3 output directories trace/ (--kernel-trace
--stats), fetch/ and write/ (--pmc FETCH_SIZE, --pmc WRITE_SIZE on the
workload), cfetch/ and cwrite/ (the same counters on a plain device copy of
--copy-bytes bytes read and written, which calibrates the counters: bytes per
launch = counter KiB * 1024 * copy bytes / copy counter bytes), and any other
counter directories (their per-kernel means are listed).  traffic.json is
written next to OUT.txt: the bytes per launch keyed by kernel name fragment,
with the commit (--commit) and commands it was measured with.
"""
import argparse
import collections
import csv
import json
import os


def _rows(path):
    with open(path) as fh:
        return list(csv.DictReader(fh))


def _counters(root, name):
    agg = collections.defaultdict(list)
    path = os.path.join(root, name, f"{name}_counter_collection.csv")
    for r in _rows(path):
        agg[(r["Kernel_Name"], r["Counter_Name"])].append(float(r["Counter_Value"]))
    return agg


def _mean_for(agg, fragment, counter=None):
    vals = [v for (k, c), v in agg.items() if fragment in k and (counter is None or c == counter)]
    if not vals:
        return None
    flat = [x for v in vals for x in v]
    return sum(flat) / len(flat)


def summarize(root, out_txt, kernels, copy_bytes, commands, commit=None):
    lines = ["rocprofv3 summary (generated by python -m pd_gpu1_amd.profsum)", ""]
    lines += ["Commands:"] + [f"  {c}" for c in commands] + [""]
    ks = _rows(os.path.join(root, "trace", "trace_kernel_stats.csv"))
    lines.append("Kernel trace (--kernel-trace --stats), top kernels:")
    lines.append(f"{'kernel':80s} {'calls':>6s} {'avg_us':>10s} {'pct':>6s}")
    for r in ks[:16]:
        lines.append(f"{r['Name'][:80]:80s} {r['Calls']:>6s} {float(r['AverageNs']) / 1e3:10.1f} "
                     f"{float(r['Percentage']):6.2f}")
    cf, cw = _counters(root, "cfetch"), _counters(root, "cwrite")
    copy_f = _mean_for(cf, "copy16")
    copy_w = _mean_for(cw, "copy16")
    ff = copy_bytes / (copy_f * 1024)
    fw = copy_bytes / (copy_w * 1024)
    lines += ["", f"Copy calibration (k_copy16 reads and writes {copy_bytes} B per launch): FETCH_SIZE "
                  f"{copy_f:.6g} KiB, WRITE_SIZE {copy_w:.6g} KiB per launch -> scale fetch x{ff:.3f}, "
                  f"write x{fw:.3f}"]
    fe, wr = _counters(root, "fetch"), _counters(root, "write")
    traffic = {}
    lines += ["", "HBM traffic per launch (counter KiB x 1024 x scale):"]
    for k in kernels:
        fv, wv = _mean_for(fe, k), _mean_for(wr, k)
        if fv is None or wv is None:
            continue
        fb, wb = fv * 1024 * ff, wv * 1024 * fw
        traffic[k] = fb + wb
        avg = [float(r["AverageNs"]) for r in ks if k in r["Name"]]
        rate = f", {(fb + wb) / (avg[0] * 1e-9) / 1e9:.0f} GB/s at the traced mean duration" if avg else ""
        lines.append(f"  {k:30s} fetch {fb / 1e9:8.3f} GB  write {wb / 1e9:8.3f} GB  total {(fb + wb) / 1e9:8.3f} GB"
                     + rate)
    for sub in sorted(os.listdir(root)):
        if sub in ("trace", "fetch", "write", "cfetch", "cwrite"):
            continue
        path = os.path.join(root, sub, f"{sub}_counter_collection.csv")
        if not os.path.exists(path):
            continue
        agg = _counters(root, sub)
        lines += ["", f"Counters ({sub}), mean per dispatch:"]
        for (kname, cname), v in sorted(agg.items()):
            if any(k in kname for k in kernels):
                lines.append(f"  {kname[:64]:64s} {cname:26s} {sum(v) / len(v):.5g}")
    with open(out_txt, "w") as fh:
        fh.write("\n".join(lines) + "\n")
    with open(os.path.join(os.path.dirname(os.path.abspath(out_txt)), "traffic.json"), "w") as fh:
        json.dump({"commit": commit, "summary": os.path.basename(out_txt), "commands": commands,
                   "kernels": traffic}, fh, indent=1)
    return lines


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("root")
    ap.add_argument("out")
    ap.add_argument("kernels")
    ap.add_argument("--copy-bytes", type=int, default=1 << 30)
    ap.add_argument("--command", action="append", default=[])
    ap.add_argument("--commit", default=None, help="commit of the measured build (recorded in traffic.json)")
    a = ap.parse_args(argv)
    print("\n".join(summarize(a.root, a.out, a.kernels.split(","), a.copy_bytes, a.command, a.commit)))


if __name__ == "__main__":
    main()
