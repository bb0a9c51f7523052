"""Summarise rocprofv3 output databases (results.db) per kernel.

Measurement infrastructure.  Kernel-trace mode reads the SQLite output of
`rocprofv3 --kernel-trace --stats -d <dir> -o <name> -- <cmd>` and prints a
markdown table (calls, total ms, average/min/max us, share of GPU time).
Counter mode (--pmc) reads the output of `rocprofv3 --pmc <COUNTER> ...` and
prints, per kernel and counter, the number of dispatches and the average and
total counter value per dispatch; --scale multiplies the values (the factor
by which a counter under-reads a plain copy of known size).

    python steiner-edge-linear_amd/profsum.py <results.db> [--title T] [--cmd C]
    python steiner-edge-linear_amd/profsum.py --pmc <results.db> [--scale S]
"""
import argparse
import sqlite3


def short(name):
    return name.split("(")[0].replace("void ", "")


def trace_table(cur, title, cmd):
    rows = cur.execute(
        "select name, count(*), sum(end-start), avg(end-start), min(end-start), max(end-start) "
        "from kernels group by name order by sum(end-start) desc").fetchall()
    total = sum(r[2] for r in rows)
    print(f"# {title}\n")
    if cmd:
        print(f"Command: `{cmd}`\n")
    print("| kernel | calls | total ms | avg us | min us | max us | share |")
    print("|---|---|---|---|---|---|---|")
    for name, cnt, tot, avg, mn, mx in rows:
        print(f"| `{short(name)}` | {cnt} | {tot / 1e6:.2f} | {avg / 1e3:.1f} | {mn / 1e3:.1f} | "
              f"{mx / 1e3:.1f} | {100.0 * tot / total:.1f}% |")
    print(f"\nTotal GPU kernel time: {total / 1e6:.2f} ms over {sum(r[1] for r in rows)} dispatches.")


def pmc_table(cur, scale):
    rows = cur.execute(
        "select name, counter_name, count(*), avg(counter_value), sum(counter_value), "
        "avg(duration) from pmc_events group by name, counter_name "
        "order by sum(counter_value) desc").fetchall()
    print("| kernel | counter | dispatches | avg per dispatch | total | avg us (counting run) |")
    print("|---|---|---|---|---|---|")
    for name, ctr, cnt, avg, tot, dur in rows:
        print(f"| `{short(name)}` | {ctr} | {cnt} | {avg * scale:.1f} | {tot * scale:.1f} | "
              f"{dur / 1e3:.1f} |")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("db")
    ap.add_argument("--title", default="kernel statistics")
    ap.add_argument("--cmd", default="")
    ap.add_argument("--pmc", action="store_true")
    ap.add_argument("--scale", type=float, default=1.0)
    args = ap.parse_args()
    cur = sqlite3.connect(args.db).cursor()
    if args.pmc:
        pmc_table(cur, args.scale)
    else:
        trace_table(cur, args.title, args.cmd)


if __name__ == "__main__":
    main()
