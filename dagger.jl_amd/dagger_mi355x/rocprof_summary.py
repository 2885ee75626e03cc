"""Summaries of rocprofv3 CSV output (kernel statistics and PMC counters).

  python -m dagger_mi355x.rocprof_summary stats KERNEL_STATS_CSV [--top 12]
      per-kernel calls, total and average duration, share of GPU time
  python -m dagger_mi355x.rocprof_summary traffic FETCH_CSV WRITE_CSV \
          --kernel REGEX --copy-bytes BYTES
      HBM bytes per dispatch of the kernels matching REGEX from two
      `rocprofv3 --pmc FETCH_SIZE` / `--pmc WRITE_SIZE` runs (KiB per
      dispatch).  Each counter is scaled by what it reads for the copy
      kernel (`__amd_rocclr_copyBuffer`) of the largest device-to-device copy
      in the same run, whose true size is BYTES (read BYTES, write BYTES).

Prints a short text summary; the `traffic` mode ends with machine-readable
lines `traffic_bytes_per_launch: <bytes>` and `fetch_scale`/`write_scale`.
"""
import argparse
import csv
import re
import sys
from collections import defaultdict


def _short(name, width=110):
    name = name.strip('"')
    return name if len(name) <= width else name[:width - 3] + "..."


def stats(path, top):
    rows = list(csv.DictReader(open(path)))
    total = sum(float(r["TotalDurationNs"]) for r in rows)
    rows.sort(key=lambda r: -float(r["TotalDurationNs"]))
    out = [f"{'calls':>7} {'total ms':>10} {'avg us':>10} {'share':>6}  kernel"]
    for r in rows[:top]:
        t = float(r["TotalDurationNs"])
        out.append(f"{int(r['Calls']):>7} {t / 1e6:>10.3f} {float(r['AverageNs']) / 1e3:>10.2f} "
                   f"{100 * t / total:>5.1f}%  {_short(r['Name'])}")
    out.append(f"total kernel time: {total / 1e6:.3f} ms over {sum(int(r['Calls']) for r in rows)} "
               f"dispatches")
    return "\n".join(out)


def _per_dispatch(path, counter):
    """{dispatch id: (kernel name, value)} of one counter."""
    vals = {}
    for r in csv.DictReader(open(path)):
        if r["Counter_Name"] != counter:
            continue
        d = int(r["Dispatch_Id"])
        name, v = r["Kernel_Name"], float(r["Counter_Value"])
        if d in vals:  # per-XCD / per-instance rows of one dispatch: sum them
            vals[d] = (name, vals[d][1] + v)
        else:
            vals[d] = (name, v)
    return vals


def _copy_scale(vals, copy_bytes):
    copies = [v for name, v in vals.values() if "copyBuffer" in name]
    if not copies:
        raise SystemExit("no __amd_rocclr_copyBuffer dispatch in the run")
    big = max(copies)
    same = [v for v in copies if v >= 0.95 * big]
    kib = sum(same) / len(same)
    return copy_bytes / (kib * 1024.0), kib, len(same)


def traffic(fetch_csv, write_csv, kernel, copy_bytes):
    pat = re.compile(kernel)
    f = _per_dispatch(fetch_csv, "FETCH_SIZE")
    w = _per_dispatch(write_csv, "WRITE_SIZE")
    fs, fkib, fn = _copy_scale(f, copy_bytes)
    ws, wkib, wn = _copy_scale(w, copy_bytes)
    fk = [v for name, v in f.values() if pat.search(name)]
    wk = [v for name, v in w.values() if pat.search(name)]
    if not fk or not wk:
        raise SystemExit(f"no dispatch matches {kernel!r}")
    names = sorted({name for name, _ in f.values() if pat.search(name)})
    fetch_b = sum(fk) / len(fk) * 1024.0 * fs
    write_b = sum(wk) / len(wk) * 1024.0 * ws
    per_kernel = defaultdict(list)
    for name, v in f.values():
        per_kernel[name].append(v)
    out = [
        f"copy kernel ({copy_bytes} B read + written): FETCH_SIZE {fkib:.0f} KiB "
        f"(mean of {fn}), WRITE_SIZE {wkib:.0f} KiB (mean of {wn})",
        f"fetch_scale: {fs:.4f}",
        f"write_scale: {ws:.4f}",
        f"kernel: {', '.join(_short(n, 160) for n in names)}",
        f"dispatches: {len(fk)} (FETCH run), {len(wk)} (WRITE run)",
        f"mean FETCH_SIZE per dispatch: {sum(fk) / len(fk):.0f} KiB -> {fetch_b / 1e6:.2f} MB scaled",
        f"mean WRITE_SIZE per dispatch: {sum(wk) / len(wk):.0f} KiB -> {write_b / 1e6:.2f} MB scaled",
        f"traffic_bytes_per_launch: {fetch_b + write_b:.0f}",
    ]
    return "\n".join(out)


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    sub = ap.add_subparsers(dest="mode", required=True)
    s = sub.add_parser("stats")
    s.add_argument("csv")
    s.add_argument("--top", type=int, default=12)
    t = sub.add_parser("traffic")
    t.add_argument("fetch_csv")
    t.add_argument("write_csv")
    t.add_argument("--kernel", required=True)
    t.add_argument("--copy-bytes", type=int, required=True)
    a = ap.parse_args(argv)
    if a.mode == "stats":
        print(stats(a.csv, a.top))
    else:
        print(traffic(a.fetch_csv, a.write_csv, a.kernel, a.copy_bytes))
    return 0


if __name__ == "__main__":
    sys.exit(main())
