"""Condense rocprofv3 CSV output into the short text summaries kept with the
sources.

  python -m bulk_amd.rocprof_summary stats <dir> [--command "..."] [--last KERNEL:N]
      per-kernel calls / average / total from *_kernel_stats.csv, and for
      the library's kernels the average over the dispatches that ran alone
      (no other dispatch overlapping them in time, from *_kernel_trace.csv):
      graph replays run independent branches concurrently, eager launches
      (the benchmark's per-stage HIP-event timing) one kernel at a time.
      --last adds the average over the last N dispatches of KERNEL (the
      benchmark's eager per-stage launches come after its graph replays)

  python -m bulk_amd.rocprof_summary counters <dir> [--command "..."]
      per library kernel: dispatches, then each counter of a --pmc run
      (*_counter_collection.csv) summed over its instances and averaged
      over the kernel's dispatches

  python -m bulk_amd.rocprof_summary requests <read_dir> <write_dir> [--command "..."]
      request sizes at the L2-to-memory interface per library kernel, from a
      --pmc run of TCC_BUBBLE_sum TCC_EA0_RDREQ_sum TCC_EA0_RDREQ_32B_sum
      (reads: 128-byte requests are TCC_BUBBLE, 32-byte ones RDREQ_32B, the
      rest 64-byte -- the split the FETCH_SIZE expression uses) and one of
      TCC_EA0_WRREQ_sum TCC_EA0_WRREQ_64B_sum (writes: 64 or 32 bytes),
      summed over each kernel's dispatches

  python -m bulk_amd.rocprof_summary traffic <fetch_dir> <write_dir> [--jobs J --seed S]
      HBM bytes per dispatch from a FETCH_SIZE run and a WRITE_SIZE run
      (*_counter_collection.csv).  The counters are scaled by what they read
      for the 1 GiB device copies the profiled process makes
      (bulk_amd.opbench.calibration_copy: 1 GiB read + 1 GiB written):
      factor = known bytes / counter bytes.
"""
from __future__ import annotations

import argparse
import collections
import csv
import glob
import os
import re
import sys

COPY_BYTES = 1 << 30  # the calibration copy: 2^28 int32


def _short(name: str, width: int = 96) -> str:
    name = re.sub(r"\(.*", "", name)  # drop the argument list
    name = name.replace("void ", "").replace("bulkx::", "")
    return name if len(name) <= width else name[: width - 3] + "..."


def _find(d: str, suffix: str) -> str:
    hits = glob.glob(os.path.join(d, "**", f"*{suffix}"), recursive=True)
    if not hits:
        sys.exit(f"no *{suffix} under {d}")
    return hits[0]


def stats(d: str, command: str | None, top: int, last: str | None = None) -> str:
    rows = list(csv.DictReader(open(_find(d, "kernel_stats.csv"))))
    rows.sort(key=lambda r: -float(r["TotalDurationNs"]))
    out = []
    if command:
        out.append(f"command: {command}\n")
    out.append("| kernel | calls | avg us | total ms | % |")
    out.append("|---|---:|---:|---:|---:|")
    for r in rows[:top]:
        out.append(f"| {_short(r['Name'])} | {int(r['Calls'])} | {float(r['AverageNs']) / 1e3:.1f} | "
                   f"{float(r['TotalDurationNs']) / 1e6:.2f} | {float(r['Percentage']):.2f} |")
    hits = glob.glob(os.path.join(d, "**", "*kernel_trace.csv"), recursive=True)
    if hits:
        disp = [(int(r["Start_Timestamp"]), int(r["End_Timestamp"]), r["Kernel_Name"],
                 int(r.get("Grid_Size_X") or r.get("Grid_Size") or 0))
                for r in csv.DictReader(open(hits[0]))]
        disp.sort()
        alone = collections.defaultdict(list)
        by_grid = collections.defaultdict(lambda: collections.defaultdict(list))
        reach = -1  # latest end among the dispatches before i
        for i, (s, e, n, g) in enumerate(disp):
            nxt = disp[i + 1][0] if i + 1 < len(disp) else None
            if s >= reach and (nxt is None or nxt >= e) and "bulkx" in n:
                alone[_short(n)].append(e - s)
                by_grid[_short(n)][g].append(e - s)
            reach = max(reach, e)
        if alone:
            out.append("")
            out.append("Library kernels, dispatches that ran alone:")
            out.append("")
            out.append("| kernel | alone | avg us |")
            out.append("|---|---:|---:|")
            for n, v in sorted(alone.items(), key=lambda kv: -sum(kv[1])):
                out.append(f"| {n} | {len(v)} | {sum(v) / len(v) / 1e3:.1f} |")
        # A kernel launched with several grid sizes (one per merge round)
        # does different work per launch: its plain average depends on which
        # launches ran alone, so the average per grid size and their mean
        # (one launch of each size, as a step makes them) are listed too.
        for n, grids in sorted(by_grid.items(), key=lambda kv: -sum(map(sum, kv[1].values()))):
            if len(grids) < 2:
                continue
            out.append("")
            out.append(f"{n}, dispatches that ran alone, by grid size (workitems):")
            out.append("")
            out.append("| grid | alone | avg us |")
            out.append("|---:|---:|---:|")
            means = []
            for g in sorted(grids, reverse=True):
                v = grids[g]
                means.append(sum(v) / len(v))
                out.append(f"| {g} | {len(v)} | {means[-1] / 1e3:.1f} |")
            out.append("")
            out.append(f"mean over the {len(means)} grid sizes: {sum(means) / len(means) / 1e3:.1f} us")
        if last:
            name, count = last.rsplit(":", 1)
            count = int(count)
            mine = [e - s for s, e, n, _ in disp if _short(n) == name]
            if mine:
                tail = mine[-count:]
                out.append("")
                out.append(f"{name}, the last {len(tail)} of its {len(mine)} dispatches "
                           f"(the eager per-stage launches): avg {sum(tail) / len(tail) / 1e3:.1f} us")
    return "\n".join(out) + "\n"


def _per_dispatch(d: str, counter: str):
    """{dispatch id: (kernel name, grid size, counter value)}"""
    vals = collections.defaultdict(float)
    meta = {}
    for r in csv.DictReader(open(_find(d, "counter_collection.csv"))):
        if r["Counter_Name"] != counter:
            continue
        did = int(r["Dispatch_Id"])
        vals[did] += float(r["Counter_Value"])
        meta[did] = (r["Kernel_Name"], int(r["Grid_Size"]))
    return {k: (meta[k][0], meta[k][1], v) for k, v in vals.items()}


def counters(d: str, command: str | None) -> str:
    out = []
    if command:
        out.append(f"command: {command}\n")
    out.append("Per library kernel: dispatches, then each counter averaged per dispatch "
               "(summed over the counter instances).\n")
    per = collections.defaultdict(float)  # (dispatch, counter) -> value
    name_of = {}
    for r in csv.DictReader(open(_find(d, "counter_collection.csv"))):
        if "bulkx" not in r["Kernel_Name"]:
            continue
        did = int(r["Dispatch_Id"])
        per[(did, r["Counter_Name"])] += float(r["Counter_Value"])
        name_of[did] = _short(r["Kernel_Name"], 200)
    by_kernel = collections.defaultdict(lambda: collections.defaultdict(list))
    for (did, c), v in per.items():
        by_kernel[name_of[did]][c].append(v)
    for k in sorted(by_kernel):
        cs = by_kernel[k]
        n = max(len(v) for v in cs.values())
        fields = " ".join(f"{c}={sum(v) / len(v):.4g}" for c, v in sorted(cs.items()))
        out.append(f"{k} dispatches {n} {fields}")
    return "\n".join(out) + "\n"


def _kernel_counters(d: str) -> dict:
    """{short kernel name: {counter: value summed over instances and dispatches}}"""
    out = collections.defaultdict(lambda: collections.defaultdict(float))
    for r in csv.DictReader(open(_find(d, "counter_collection.csv"))):
        if "bulkx" in r["Kernel_Name"] or "copy" in r["Kernel_Name"].lower():
            out[_short(r["Kernel_Name"], 200)][r["Counter_Name"]] += float(r["Counter_Value"])
    return out


def requests(read_dir: str, write_dir: str, command: str | None) -> str:
    out = []
    if command:
        out.append(f"command: {command}\n")
    out.append("Per library kernel, summed over its dispatches: read requests by size "
               "(128 / 64 / 32 bytes) and the share of read bytes in 128-byte requests; "
               "write requests by size (64 / 32 bytes) and the share of written bytes in "
               "64-byte requests.\n")
    rd, wr = _kernel_counters(read_dir), _kernel_counters(write_dir)
    out.append("| kernel | reads 128 B | 64 B | 32 B | read bytes in 128 B | "
               "writes 64 B | 32 B | written bytes in 64 B |")
    out.append("|---|---:|---:|---:|---:|---:|---:|---:|")
    for k in sorted(set(rd) | set(wr)):
        r, w = rd.get(k, {}), wr.get(k, {})
        n128 = r.get("TCC_BUBBLE_sum", 0.0)
        n32 = r.get("TCC_EA0_RDREQ_32B_sum", 0.0)
        n64 = max(0.0, r.get("TCC_EA0_RDREQ_sum", 0.0) - n128 - n32)
        rb = 128 * n128 + 64 * n64 + 32 * n32
        w64 = w.get("TCC_EA0_WRREQ_64B_sum", 0.0)
        w32 = max(0.0, w.get("TCC_EA0_WRREQ_sum", 0.0) - w64)
        wb = 64 * w64 + 32 * w32
        fr = f"{100 * 128 * n128 / rb:.1f} %" if rb else "-"
        fw = f"{100 * 64 * w64 / wb:.1f} %" if wb else "-"
        out.append(f"| {k} | {n128:.4g} | {n64:.4g} | {n32:.4g} | {fr} | {w64:.4g} | {w32:.4g} | {fw} |")
    return "\n".join(out) + "\n"


def traffic(fetch_dir: str, write_dir: str, jobs: int, seed: int, command: str | None) -> str:
    out = []
    if command:
        out.append(f"command: {command}\n")
    per = {}
    for counter, d in (("FETCH_SIZE", fetch_dir), ("WRITE_SIZE", write_dir)):
        disp = _per_dispatch(d, counter)
        # calibration: the first three large copies of the run (the process
        # makes its 1 GiB calibration copies before any other large copy)
        copies = [(did, v) for did, (n, g, v) in disp.items() if "copy" in n.lower()]
        top = max((v for _, v in copies), default=0.0)
        big = [v for _, v in sorted(c for c in copies if c[1] >= 0.25 * top)][:3]
        kib = sum(big) / len(big) if big else float("nan")
        factor = COPY_BYTES / (kib * 1024.0) if big else float("nan")
        out.append(f"{counter}: copy of {COPY_BYTES} bytes reads {kib:.0f} KiB -> factor {factor:.4f}")
        groups = collections.defaultdict(list)
        for n, g, v in disp.values():
            if "bulkx" in n:
                groups[_short(n, 200)].append(v * 1024.0 * factor)
        per[counter] = groups
    out.append("")
    if jobs is not None:  # the workload of a bench.py run
        out.append(f"jobs: {jobs}")
        out.append(f"seed: {seed}")
        out.append("")
    names = sorted(set(per["FETCH_SIZE"]) | set(per["WRITE_SIZE"]))
    for n in names:
        f = per["FETCH_SIZE"].get(n, [])
        w = per["WRITE_SIZE"].get(n, [])
        if not f and not w:
            continue
        fb = sum(f) / len(f) if f else 0.0
        wb = sum(w) / len(w) if w else 0.0
        out.append(f"kernel: {n}")
        out.append(f"dispatches: {len(f)} / {len(w)}")
        out.append(f"fetch_bytes_per_launch: {fb:.0f}")
        out.append(f"write_bytes_per_launch: {wb:.0f}")
        out.append("")
    return "\n".join(out)


def main(argv=None):
    p = argparse.ArgumentParser()
    sub = p.add_subparsers(dest="mode", required=True)
    s = sub.add_parser("stats")
    s.add_argument("dir")
    s.add_argument("--command")
    s.add_argument("--top", type=int, default=30)
    s.add_argument("--last", default=None, help="KERNEL:N -- the average of its last N dispatches")
    c = sub.add_parser("counters")
    c.add_argument("dir")
    c.add_argument("--command")
    q = sub.add_parser("requests")
    q.add_argument("read_dir")
    q.add_argument("write_dir")
    q.add_argument("--command")
    t = sub.add_parser("traffic")
    t.add_argument("fetch_dir")
    t.add_argument("write_dir")
    t.add_argument("--jobs", type=int, default=None)
    t.add_argument("--seed", type=int, default=None)
    t.add_argument("--command")
    a = p.parse_args(argv)
    if a.mode == "stats":
        print(stats(a.dir, a.command, a.top, a.last), end="")
    elif a.mode == "counters":
        print(counters(a.dir, a.command), end="")
    elif a.mode == "requests":
        print(requests(a.read_dir, a.write_dir, a.command), end="")
    else:
        print(traffic(a.fetch_dir, a.write_dir, a.jobs, a.seed, a.command))


if __name__ == "__main__":
    main()
