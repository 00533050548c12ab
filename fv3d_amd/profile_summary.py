"""Summaries of rocprofv3 runs of bench.py's explicit leg.

    python -m fv3d_amd.profile_summary TRACE_DIR FETCH_DIR WRITE_DIR [STALL_DIR VALU_DIR]
        [--json OUT.json] [--n 96]

TRACE_DIR holds a ``--kernel-trace --stats`` run (``*_kernel_stats.csv``);
FETCH_DIR and WRITE_DIR hold ``--pmc FETCH_SIZE`` and ``--pmc WRITE_SIZE``
runs of ``bench.py --copy-calibration`` (``*_counter_collection.csv``).
The two counters are in KiB per dispatch; they are scaled to bytes by the
ratio that a device-to-device copy of 1 GiB (``__amd_rocclr_copyBuffer``,
1 GiB read and 1 GiB written) shows in the same runs.  Prints a markdown
summary; with --json writes the measured bytes per launch of the two stencil
kernels together with the SHA-256 of the kernel sources they were measured
on (bench.py reports them as ``roofline.traffic`` only for those sources).
"""
import argparse
import collections
import csv
import glob
import json
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GIB = float(1 << 30)
# the stage's two kernels (the residual kernel: face-centred, or lane per slot)
KERNELS = {"gradient": ["fv_gradient_tile_kernel<4, double>"],
           "residual": ["fv_residual_face_kernel<0, true, true>",
                        "fv_residual_group_kernel<0, true, true, 4, double>"]}


def _one(d, suffix):
    files = sorted(glob.glob(os.path.join(d, "**", "*" + suffix), recursive=True))
    if not files:
        raise SystemExit("no *%s under %s" % (suffix, d))
    return files[-1]


def kernel_stats(d):
    rows = list(csv.DictReader(open(_one(d, "_kernel_stats.csv"))))
    out = []
    for r in rows:
        out.append({"name": r["Name"], "calls": int(r["Calls"]),
                    "avg_us": float(r["AverageNs"]) / 1e3,
                    "total_ms": float(r["TotalDurationNs"]) / 1e6,
                    "pct": float(r["Percentage"])})
    return out


def counters(d):
    """{kernel name: {counter: [values per dispatch]}}"""
    acc = collections.defaultdict(lambda: collections.defaultdict(list))
    for r in csv.DictReader(open(_one(d, "_counter_collection.csv"))):
        acc[r["Kernel_Name"]][r["Counter_Name"]].append(float(r["Counter_Value"]))
    return acc


def _match(acc, keys):
    for key in ([keys] if isinstance(keys, str) else keys):
        for name, v in acc.items():
            if key in name:
                return name, v
    return None, None


def mean(x):
    return sum(x) / len(x) if x else float("nan")


def traffic(fetch_dir, write_dir):
    f, w = counters(fetch_dir), counters(write_dir)
    _, cf = _match(f, "__amd_rocclr_copyBuffer")
    _, cw = _match(w, "__amd_rocclr_copyBuffer")
    if cf is None or cw is None:
        raise SystemExit("no calibration copy in the counter runs (bench.py --copy-calibration)")
    # the largest copies are the 1 GiB calibration copies
    kf = max(cf["FETCH_SIZE"]) * 1024.0
    kw = max(cw["WRITE_SIZE"]) * 1024.0
    read_scale, write_scale = GIB / kf, GIB / kw
    out = {}
    for tag, key in KERNELS.items():
        nf, vf = _match(f, key)
        nw, vw = _match(w, key)
        if vf is None or vw is None:
            continue
        rd = mean(vf["FETCH_SIZE"]) * 1024.0 * read_scale
        wr = mean(vw["WRITE_SIZE"]) * 1024.0 * write_scale
        out[tag] = {"kernel": nf, "read": rd, "written": wr, "total": rd + wr,
                    "dispatches": len(vf["FETCH_SIZE"])}
    return out, {"read_bytes_per_fetch_kib": read_scale * 1024.0,
                 "write_bytes_per_write_kib": write_scale * 1024.0}


def source_hash():
    import hashlib
    h = hashlib.sha256()
    for f in ("fv3d_amd/csrc/fv_kernels.hip", "fv3d_amd/csrc/fv_physics.h",
              "fv3d_amd/csrc/fv_device.h"):
        with open(os.path.join(ROOT, f), "rb") as fh:
            h.update(fh.read())
    return h.hexdigest()


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("trace")
    ap.add_argument("fetch")
    ap.add_argument("write")
    ap.add_argument("extra", nargs="*", help="further counter runs to tabulate")
    ap.add_argument("--json")
    ap.add_argument("--n", type=int, default=96)
    a = ap.parse_args(argv)
    st = kernel_stats(a.trace)
    print("| kernel | calls | avg us | total % |\n|---|---|---|---|")
    for r in st:
        if r["pct"] >= 0.05:
            print("| `%s` | %d | %.1f | %.1f |" % (r["name"][:110], r["calls"], r["avg_us"], r["pct"]))
    tr, cal = traffic(a.fetch, a.write)
    print("\nCalibration: %.1f bytes read per FETCH_SIZE KiB, %.1f bytes written per "
          "WRITE_SIZE KiB (1 GiB device copy)\n" % (cal["read_bytes_per_fetch_kib"],
                                                   cal["write_bytes_per_write_kib"]))
    print("| kernel | read (GB) | written (GB) | total (GB) |\n|---|---|---|---|")
    for tag, t in tr.items():
        print("| %s | %.3f | %.3f | %.3f |" % (tag, t["read"] / 1e9, t["written"] / 1e9,
                                               t["total"] / 1e9))
    for d in a.extra:
        acc = counters(d)
        for tag, key in KERNELS.items():
            name, v = _match(acc, key)
            if v is None:
                continue
            print("\n%s (%s):" % (tag, os.path.basename(d.rstrip("/"))))
            for c, vals in sorted(v.items()):
                print("- %s: %.6g (mean of %d dispatches)" % (c, mean(vals), len(vals)))
    if a.json:
        avg = {}
        for tag, keys in KERNELS.items():
            for r in st:
                if any(k in r["name"] for k in keys) and tag not in avg:
                    avg[tag] = r["avg_us"]
        with open(a.json, "w") as f:
            json.dump({"n": a.n, "sources_sha256": source_hash(),
                       "bytes_per_launch": {k: v["total"] for k, v in tr.items()},
                       "detail": tr, "calibration": cal, "trace_avg_us": avg}, f, indent=1)
            f.write("\n")
    return 0


if __name__ == "__main__":
    sys.exit(main())
