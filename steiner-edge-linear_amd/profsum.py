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
    python steiner-edge-linear_amd/profsum.py --busy <results.db> [--cus 256]
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


def busy_table(cur, cus):
    """Per kernel function: TA / TD busy fractions and cache hit rates.  A
    dispatch's counter may come as several rows (one per XCD or shader
    engine): block counters (TA_*, TD_*, TCP_*, TCC_*, SQ_*) are summed over
    them, GRBM_GUI_ACTIVE (the same clock on every XCD) is averaged.  Busy =
    sum over dispatches of the unit-summed busy cycles / (GPU-active cycles x
    CUs)."""
    import collections

    rows = cur.execute("select name, dispatch_id, counter_name, counter_value, duration "
                       "from pmc_events").fetchall()
    per = collections.defaultdict(lambda: collections.defaultdict(list))
    dur = {}
    for name, did, ctr, val, d in rows:
        per[(short(name), did)][ctr].append(val)
        dur[(short(name), did)] = d
    agg = collections.defaultdict(lambda: collections.defaultdict(float))
    tim = collections.defaultdict(float)
    ndisp = collections.defaultdict(int)
    for (k, did), ctrs in per.items():
        for ctr, vals in ctrs.items():
            agg[k][ctr] += (sum(vals) / len(vals)) if ctr.startswith("GRBM") else sum(vals)
        tim[k] += dur[(k, did)]
        ndisp[k] += 1
    print("| kernel | dispatches | time (counting run) | TA busy | TD busy | L1 hit | L2 hit |"
          " VALU insts | VMEM rd insts |")
    print("|---|---|---|---|---|---|---|---|---|")
    for k in sorted(agg, key=lambda x: -tim[x]):
        a = agg[k]
        g = a.get("GRBM_GUI_ACTIVE", 0.0) * cus

        def frac(x, y):
            return f"{x / y:.2f}" if y else "-"

        ta = frac(a["TA_TA_BUSY_sum"], g) if "TA_TA_BUSY_sum" in a else "-"
        td = frac(a["TD_TD_BUSY_sum"], g) if "TD_TD_BUSY_sum" in a else "-"
        l1 = (frac(a["TCP_TOTAL_CACHE_ACCESSES_sum"] - a["TCP_TCC_READ_REQ_sum"],
                   a["TCP_TOTAL_CACHE_ACCESSES_sum"]) if "TCP_TCC_READ_REQ_sum" in a else "-")
        l2 = (frac(a["TCC_HIT_sum"], a["TCC_HIT_sum"] + a["TCC_MISS_sum"])
              if "TCC_HIT_sum" in a else "-")
        valu = f"{a['SQ_INSTS_VALU']:.3g}" if "SQ_INSTS_VALU" in a else "-"
        vmem = f"{a['SQ_INSTS_VMEM_RD']:.3g}" if "SQ_INSTS_VMEM_RD" in a else "-"
        if tim[k] < 1e6:  # under 1 ms in the run
            continue
        print(f"| `{k}` | {ndisp[k]} | {tim[k] / 1e6:.1f} ms | {ta} | {td} | {l1} | {l2} | "
              f"{valu} | {vmem} |")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("db")
    ap.add_argument("--title", default="kernel statistics")
    ap.add_argument("--cmd", default="")
    ap.add_argument("--pmc", action="store_true")
    ap.add_argument("--scale", type=float, default=1.0)
    ap.add_argument("--busy", action="store_true",
                    help="TA/TD busy and L1/L2 hit rates per kernel (counter runs)")
    ap.add_argument("--cus", type=int, default=256)
    args = ap.parse_args()
    cur = sqlite3.connect(args.db).cursor()
    if args.busy:
        busy_table(cur, args.cus)
    elif args.pmc:
        pmc_table(cur, args.scale)
    else:
        trace_table(cur, args.title, args.cmd)


if __name__ == "__main__":
    main()
