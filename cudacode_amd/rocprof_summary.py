"""Summarise rocprofv3 databases (rocpd SQLite) into the short text records
kept under profiles/.

  python -m cudacode_amd.rocprof_summary trace DB [N]        per-kernel time table (top N)
  python -m cudacode_amd.rocprof_summary grid DB [N]         per (kernel, grid size) time table
  python -m cudacode_amd.rocprof_summary pmc DB [DB ...]      mean counter value per kernel
  python -m cudacode_amd.rocprof_summary pmcgrid DB [DB ...]  the same per (kernel, grid size)

Counter passes are scaled by a copy kernel of known size run under the same
counter (see bench.py --calibrate-copy): bytes = KiB * 1024 * scale.
"""
from __future__ import annotations

import sqlite3
import sys


def trace(db: str, limit: int = 12) -> str:
    c = sqlite3.connect(db)
    rows = c.execute("select name, total_calls, total_duration, average, percentage "
                     "from top_kernels limit ?", (limit,)).fetchall()
    # top_kernels reports durations in microseconds.
    out = ["| kernel | calls | total us | avg us | % |", "|---|---|---|---|---|"]
    for n, calls, tot, avg, pct in rows:
        out.append(f"| `{short(n)}` | {calls} | {tot:.1f} | {avg:.3f} | {pct:.1f} |")
    return "\n".join(out)


def grid(db: str, limit: int = 16) -> str:
    """Per (kernel, grid) dispatch statistics: the same kernel launched over
    different replica ranges (e.g. a sweep's two replica halves and a timed
    full-range launch) gets one row per grid size. Durations in us."""
    c = sqlite3.connect(db)
    rows = c.execute("select name, grid_x, grid_y, grid_z, workgroup_x, count(*), sum(duration), "
                     "avg(duration), min(duration) from kernels group by name, grid_x, grid_y, "
                     "grid_z order by sum(duration) desc limit ?", (limit,)).fetchall()
    out = ["| kernel | grid (threads) | block | calls | total us | avg us | min us |",
           "|---|---|---|---|---|---|---|"]
    for n, gx, gy, gz, wx, calls, tot, avg, mn in rows:
        out.append(f"| `{short(n)}` | {gx}x{gy}x{gz} | {wx} | {calls} | {tot / 1e3:.1f} | "
                   f"{avg / 1e3:.3f} | {mn / 1e3:.3f} |")
    return "\n".join(out)


def pmc(db: str) -> dict:
    c = sqlite3.connect(db)
    res = {}
    for name, counter, n, avg in c.execute(
            "select kernel_name, counter_name, count(*), avg(value) from counters_collection "
            "group by kernel_name, counter_name"):
        res[(short(name), counter)] = (n, avg)
    return res


def pmc_grid(db: str) -> list:
    """Mean counter value per (kernel, grid size, counter): dispatches of one
    kernel over different replica ranges kept apart."""
    c = sqlite3.connect(db)
    return c.execute("select kernel_name, grid_size, counter_name, count(*), avg(value) "
                     "from counters_collection group by kernel_name, grid_size, counter_name "
                     "order by avg(value) desc").fetchall()


def short(name: str) -> str:
    return name.split("(")[0].replace("void ", "")


def main(argv):
    if len(argv) >= 2 and argv[0] == "trace":
        print(trace(argv[1], int(argv[2]) if len(argv) > 2 else 12))
    elif len(argv) >= 2 and argv[0] == "grid":
        print(grid(argv[1], int(argv[2]) if len(argv) > 2 else 16))
    elif len(argv) >= 2 and argv[0] == "pmcgrid":
        for db in argv[1:]:
            for k, threads, cnt, n, avg in pmc_grid(db)[:16]:
                print(f"| `{short(k)}` | {threads} | {cnt} | {n} | {avg:.1f} KiB |")
    elif len(argv) >= 2 and argv[0] == "pmc":
        for db in argv[1:]:
            for (k, cnt), (n, avg) in sorted(pmc(db).items(), key=lambda t: -t[1][1])[:8]:
                print(f"| `{k}` | {cnt} | {n} | {avg:.1f} KiB |")
    else:
        print(__doc__)
        return 2
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
