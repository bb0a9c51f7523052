"""Summarise a rocprofv3 kernel-trace database (results.db) per kernel.

Test/measurement infrastructure: reads the SQLite output of
`rocprofv3 --kernel-trace --stats -d <dir> -o <name> -- <cmd>` and prints a
markdown table (calls, total ms, average/min/max us, share of GPU time).

    python steiner-edge-linear_amd/profsum.py <results.db> [--title T] [--cmd C]
"""
import argparse
import sqlite3


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("db")
    ap.add_argument("--title", default="kernel statistics")
    ap.add_argument("--cmd", default="")
    args = ap.parse_args()
    cur = sqlite3.connect(args.db).cursor()
    rows = cur.execute(
        "select name, count(*), sum(end-start), avg(end-start), min(end-start), max(end-start) "
        "from kernels group by name order by sum(end-start) desc").fetchall()
    total = sum(r[2] for r in rows)
    print(f"# {args.title}\n")
    if args.cmd:
        print(f"Command: `{args.cmd}`\n")
    print("| kernel | calls | total ms | avg us | min us | max us | share |")
    print("|---|---|---|---|---|---|---|")
    for name, cnt, tot, avg, mn, mx in rows:
        short = name.split("(")[0].replace("void ", "")
        print(f"| `{short}` | {cnt} | {tot / 1e6:.2f} | {avg / 1e3:.1f} | {mn / 1e3:.1f} | "
              f"{mx / 1e3:.1f} | {100.0 * tot / total:.1f}% |")
    print(f"\nTotal GPU kernel time: {total / 1e6:.2f} ms over {sum(r[1] for r in rows)} dispatches.")


if __name__ == "__main__":
    main()
