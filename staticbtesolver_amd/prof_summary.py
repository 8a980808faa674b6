"""Summarises rocprofv3 databases (kernel trace + FETCH_SIZE / WRITE_SIZE passes) into a short
markdown report: per-kernel call counts and average durations, and HBM bytes per dispatch of the
dominant kernel scaled by the calibration copy kernel (hbm_copy) of known size.

Usage: python -m staticbtesolver_amd.prof_summary <prof_dir> <title> <command> [copy_MiB]
where <prof_dir> holds trace/, pmc_FETCH_SIZE/, pmc_WRITE_SIZE/, pmc_copy_FETCH_SIZE/,
pmc_copy_WRITE_SIZE/ as written by staticbtesolver_amd/profile.sh.
"""
from __future__ import annotations

import glob
import os
import sqlite3
import sys


def _db(path):
    files = sorted(glob.glob(os.path.join(path, "**", "*_results.db"), recursive=True))
    if not files:
        raise FileNotFoundError(path)
    return sqlite3.connect(files[0])


def top_kernels(prof_dir):
    cur = _db(os.path.join(prof_dir, "trace")).cursor()
    cur.execute("select name, total_calls, total_duration, average, percentage from top_kernels "
                "order by total_duration desc")
    return cur.fetchall()


def counter_per_dispatch(path, counter, name_like):
    """Mean counter value (KiB) per dispatch of kernels whose name contains name_like."""
    cur = _db(path).cursor()
    cur.execute("select dispatch_id, sum(value) from counters_collection where counter_name = ? and kernel_name like ? "
                "group by dispatch_id", (counter, f"%{name_like}%"))
    rows = cur.fetchall()
    if not rows:
        return None, 0
    return sum(v for _, v in rows) / len(rows), len(rows)


def short(name):
    return name.split("(")[0].replace("void ", "")


def summary(prof_dir, title, command, copy_mib=1024):
    lines = [f"# {title}", "", "Command (repository root, on the MI355X box):", "", "```", command, "```", "",
             "## Kernel trace (rocprofv3 --kernel-trace --stats)", "",
             "| kernel | calls | total us | avg us | % |", "|---|---|---|---|---|"]
    tk = top_kernels(prof_dir)
    for n, calls, tot, avg, pct in tk[:20]:
        lines.append(f"| `{short(n)}` | {calls} | {tot:.1f} | {avg:.2f} | {pct:.1f} |")
    # the kernel reported: the busiest one, or (PROF_KERNEL=<name part>) the instantiation of that
    # kernel launched most often -- the one the autotuner kept, not the candidates it timed
    row = tk[0]
    want = os.environ.get("PROF_KERNEL")
    if want:
        cands = [r for r in tk if want in r[0]]
        if cands:
            row = max(cands, key=lambda r: r[1])
    full = short(row[0])
    dom = full.split("::")[-1].split("<")[0]
    lines += ["", f"Dominant kernel: `{full}`, average {row[3]:.2f} us per dispatch.", ""]
    res = {"dominant": dom, "instance": full, "avg_us": row[3]}
    try:
        f_k, nf = counter_per_dispatch(os.path.join(prof_dir, "pmc_FETCH_SIZE"), "FETCH_SIZE", full)
        w_k, nw = counter_per_dispatch(os.path.join(prof_dir, "pmc_WRITE_SIZE"), "WRITE_SIZE", full)
        cf, _ = counter_per_dispatch(os.path.join(prof_dir, "pmc_copy_FETCH_SIZE"), "FETCH_SIZE", "k_copy")
        cw, _ = counter_per_dispatch(os.path.join(prof_dir, "pmc_copy_WRITE_SIZE"), "WRITE_SIZE", "k_copy")
        exact_kib = copy_mib * 1024.0
        sf, sw = exact_kib / cf, exact_kib / cw
        traffic = (f_k * sf + w_k * sw) * 1024.0
        lines += ["## HBM bytes (rocprofv3 --pmc, one counter per run)", "",
                  f"- copy kernel of {copy_mib} MiB read + {copy_mib} MiB write: FETCH_SIZE {cf:.0f} KiB, "
                  f"WRITE_SIZE {cw:.0f} KiB per dispatch -> scale factors {sf:.3f} (fetch), {sw:.3f} (write)",
                  f"- `{dom}`: FETCH_SIZE {f_k:.0f} KiB ({nf} dispatches), WRITE_SIZE {w_k:.0f} KiB ({nw} dispatches)"
                  f" per dispatch",
                  f"- scaled HBM traffic per dispatch: {traffic / 1e6:.1f} MB", ""]
        res.update(fetch_kib=f_k, write_kib=w_k, scale_fetch=sf, scale_write=sw, traffic_bytes=traffic)
    except (FileNotFoundError, TypeError, ZeroDivisionError) as e:
        lines += [f"(byte counters unavailable: {e})", ""]
    return "\n".join(lines), res


if __name__ == "__main__":
    text, res = summary(sys.argv[1], sys.argv[2], sys.argv[3], int(sys.argv[4]) if len(sys.argv) > 4 else 1024)
    print(text)
    print(res, file=sys.stderr)
    if len(sys.argv) > 5:  # machine-readable copy for bench.py's roofline.traffic
        import json

        with open(sys.argv[5], "w") as f:
            json.dump(dict(res, title=sys.argv[2], command=sys.argv[3]), f, indent=1)
