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

  python -m dagger_mi355x.rocprof_summary counters COUNTER_CSV [COUNTER_CSV ...] \
          --kernel REGEX [--simds 1024] [--xcds 8]
      per-dispatch means of every counter collected for the kernels matching
      REGEX (rocprofv3 sums a dispatch's XCD and shader-engine instances into
      one value; GRBM_* counters are then divided by the XCD count to give
      one XCD's cycles), and the
      derived figures the counters allow: MFMA utilisation
      sum(SQ_VALU_MFMA_BUSY_CYCLES) / (GRBM_GUI_ACTIVE x SIMDs), the
      average shader clock GRBM_GUI_ACTIVE / dispatch duration, the share of
      LDS-active cycles lost to bank conflicts, and the wave-cycle shares
      spent waiting.

  python -m dagger_mi355x.rocprof_summary window KERNEL_TRACE_CSV \
          --start REGEX --stop REGEX [--index I]
      windows of a kernel trace: each opens at a dispatch whose kernel
      matches --start and closes before the next dispatch matching --stop
      (or --start); prints every window's span (first start to last end) and
      dispatch count, and the per-kernel time of window I (default: the
      last).

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


def _dispatch_rows(paths, pat):
    """{dispatch id: {"name", "dur_ns", counter: [row values]}} of the
    dispatches whose kernel name matches `pat`, over several CSV files."""
    out = {}
    for path in paths:
        for r in csv.DictReader(open(path)):
            name = r["Kernel_Name"]
            if not pat.search(name):
                continue
            key = (path, int(r["Dispatch_Id"]))
            d = out.setdefault(key, {"name": name, "dur_ns": None, "vals": defaultdict(list)})
            if r.get("Start_Timestamp") and r.get("End_Timestamp"):
                d["dur_ns"] = float(r["End_Timestamp"]) - float(r["Start_Timestamp"])
            d["vals"][r["Counter_Name"]].append(float(r["Counter_Value"]))
    return out


def counters(paths, kernel, simds, xcds):
    pat = re.compile(kernel)
    rows = _dispatch_rows(paths, pat)
    if not rows:
        raise SystemExit(f"no dispatch matches {kernel!r}")
    per = defaultdict(list)      # counter -> per-dispatch values
    durs = defaultdict(list)     # counter -> durations of the dispatches that carry it
    for d in rows.values():
        for c, vals in d["vals"].items():
            v = sum(vals) / xcds if c.startswith("GRBM_") else sum(vals)
            per[c].append(v)
            if d["dur_ns"]:
                durs[c].append(d["dur_ns"])
    mean = {c: sum(v) / len(v) for c, v in per.items()}
    names = sorted({d["name"] for d in rows.values()})
    out = [f"kernel: {', '.join(_short(n, 160) for n in names)}",
           f"dispatches: {len(rows)} over {len(paths)} run(s)"]
    for c in sorted(mean):
        out.append(f"  {c:<32} {mean[c]:>18.1f}  (mean per dispatch of {len(per[c])})")
    g = mean.get("GRBM_GUI_ACTIVE")
    if g and "SQ_VALU_MFMA_BUSY_CYCLES" in mean:
        # ratio of sums over the dispatches that carry both counters
        out.append(f"mfma_util: {mean['SQ_VALU_MFMA_BUSY_CYCLES'] / (g * simds):.4f}  "
                   f"(sum SQ_VALU_MFMA_BUSY_CYCLES / (GRBM_GUI_ACTIVE x {simds} SIMDs))")
    if g and durs.get("GRBM_GUI_ACTIVE"):
        ns = sum(durs["GRBM_GUI_ACTIVE"]) / len(durs["GRBM_GUI_ACTIVE"])
        out.append(f"mean_dispatch_us: {ns / 1e3:.1f}")
        out.append(f"avg_clock_mhz: {g / ns * 1e3:.0f}  (GRBM_GUI_ACTIVE / dispatch duration)")
    if mean.get("SQ_LDS_IDX_ACTIVE") and "SQ_LDS_BANK_CONFLICT" in mean:
        out.append(f"lds_bank_conflict_share: "
                   f"{mean['SQ_LDS_BANK_CONFLICT'] / mean['SQ_LDS_IDX_ACTIVE']:.4f}  "
                   f"(SQ_LDS_BANK_CONFLICT / SQ_LDS_IDX_ACTIVE)")
    w = mean.get("SQ_WAVE_CYCLES")
    if w:
        for c in ("SQ_WAIT_ANY", "SQ_WAIT_INST_ANY", "SQ_WAIT_INST_LDS", "SQ_ACTIVE_INST_LDS",
                  "SQ_ACTIVE_INST_VALU", "SQ_INST_CYCLES_VMEM_RD"):
            if c in mean:
                out.append(f"{c.lower()}_per_wave_cycle: {mean[c] / w:.4f}")
    return "\n".join(out)


def window(path, start, stop, index=-1):
    rows = sorted(csv.DictReader(open(path)), key=lambda r: int(r["Start_Timestamp"]))
    ps, pe = re.compile(start), re.compile(stop)
    wins, cur = [], None
    for r in rows:
        name = r["Kernel_Name"]
        if ps.search(name):
            cur = [r]
            wins.append(cur)
        elif pe.search(name):
            cur = None
        elif cur is not None:
            cur.append(r)
    if not wins:
        raise SystemExit(f"no dispatch matches {start!r}")
    out = []
    for i, w in enumerate(wins):
        t0 = int(w[0]["Start_Timestamp"])
        t1 = max(int(r["End_Timestamp"]) for r in w)
        out.append(f"window {i}: span {(t1 - t0) / 1e6:.3f} ms, {len(w)} dispatches")
    agg = defaultdict(lambda: [0, 0.0])
    for r in wins[index]:
        a = agg[r["Kernel_Name"]]
        a[0] += 1
        a[1] += int(r["End_Timestamp"]) - int(r["Start_Timestamp"])
    out.append(f"window {index % len(wins)}, per kernel:")
    out.append(f"{'calls':>7} {'total ms':>10} {'avg us':>10}  kernel")
    for name, (c, t) in sorted(agg.items(), key=lambda kv: -kv[1][1]):
        out.append(f"{c:>7} {t / 1e6:>10.3f} {t / c / 1e3:>10.2f}  {_short(name)}")
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
    c = sub.add_parser("counters")
    c.add_argument("csv", nargs="+")
    c.add_argument("--kernel", required=True)
    c.add_argument("--simds", type=int, default=1024)
    c.add_argument("--xcds", type=int, default=8)
    w = sub.add_parser("window")
    w.add_argument("csv")
    w.add_argument("--start", required=True)
    w.add_argument("--stop", required=True)
    w.add_argument("--index", type=int, default=-1, help="window to break down (default: last)")
    a = ap.parse_args(argv)
    if a.mode == "window":
        print(window(a.csv, a.start, a.stop, a.index))
        return 0
    if a.mode == "stats":
        print(stats(a.csv, a.top))
    elif a.mode == "counters":
        print(counters(a.csv, a.kernel, a.simds, a.xcds))
    else:
        print(traffic(a.fetch_csv, a.write_csv, a.kernel, a.copy_bytes))
    return 0


if __name__ == "__main__":
    sys.exit(main())
