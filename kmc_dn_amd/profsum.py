"""Summarise rocprofv3 SQLite output (python -m kmc_dn_amd.profsum <results.db> ...).

Prints, per kernel: dispatch count, total and average duration; and, when the
run collected counters (--pmc), the average value of every counter per
dispatch of every kernel.  Used to write the text summaries kept in profiles/.
"""
from __future__ import annotations

import argparse
import sqlite3
import sys
from collections import defaultdict


def kernel_stats(con) -> list[tuple]:
    rows = con.execute("select name, total_calls, total_duration, average, percentage "
                       "from top_kernels order by total_duration desc").fetchall()
    return rows


def dispatches(con) -> list[tuple]:
    return con.execute("select name, grid_x, workgroup_x, lds_size, vgpr_count, "
                       "scratch_size, duration, start from kernels order by start").fetchall()


def counter_stats(con) -> dict:
    acc = defaultdict(list)
    try:
        for name, counter, value in con.execute(
                "select kernel_name, counter_name, value from counters_collection"):
            acc[(name, counter)].append(value)
    except sqlite3.OperationalError:
        return {}
    return {k: (len(v), sum(v) / len(v)) for k, v in acc.items()}


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("db", nargs="+")
    ap.add_argument("--dispatches", action="store_true", help="list every dispatch")
    ap.add_argument("--match", default="", help="only kernels whose name contains this")
    args = ap.parse_args(argv)
    for path in args.db:
        con = sqlite3.connect(path)
        print(f"# {path}")
        ks = kernel_stats(con)
        if ks:
            # rocprofv3's top_kernels view reports durations in microseconds
            print("| kernel | calls | total ms | average ms | % |")
            print("|---|---|---|---|---|")
            for name, calls, tot, avg, pct in ks:
                print(f"| `{name}` | {calls} | {tot / 1e3:.3f} | {avg / 1e3:.3f} | {pct:.2f} |")
        if args.dispatches:
            # start: ms after the first listed dispatch (concurrent launches overlap)
            print("| kernel | grid | workgroup | LDS B | VGPRs | scratch | start ms | ms |")
            print("|---|---|---|---|---|---|---|---|")
            t0 = None
            for name, gx, wx, lds, vg, scr, dur, start in dispatches(con):
                if args.match in name:
                    t0 = start if t0 is None else t0
                    print(f"| `{name}` | {gx} | {wx} | {lds} | {vg} | {scr} | "
                          f"{(start - t0) / 1e6:.3f} | {dur / 1e6:.3f} |")
        cs = counter_stats(con)
        if cs:
            print("| kernel | counter | dispatches | average per dispatch |")
            print("|---|---|---|---|")
            for (name, counter), (n, avg) in sorted(cs.items()):
                print(f"| `{name}` | {counter} | {n} | {avg:.6g} |")
        print()
    return 0


if __name__ == "__main__":
    sys.exit(main())
