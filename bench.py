"""Benchmark of the EMV Steiner DP on one MI355X.

A step is one full EMV solve (singleton shortest paths, every merge and
relaxation layer; no traceback) of the plain single-instance workload:
random connected 8-regular graph, n = 262,144, m = 1,048,576, weights
uniform in [1, 1000], k = 14 uniform terminals (generator seed 2), a table
of 2^14 x 2^18 cells (16-bit: the distance bound is below 0xFFFF).  The
graph is resident in HBM before timing.

Prints one JSON line.  `value` = DP cells (2^k * n) per second of wall time
over the timed steps.  `roofline` covers the dominant kernel (the kernel
function with the largest share of device time: the relaxation round kernel
here; the merge runs as split and whole-row kernels per layer, and is
VALU/cache-bound), with achieved = algorithmic bytes / kernel time from HIP
events on the library's stream, and traffic = HBM bytes per launch measured
by two rocprofv3 counter passes (FETCH_SIZE, WRITE_SIZE) scaled against a
plain copy of known size.  `cpu_baseline` times the reference
reader-el (oracle/_ref, OpenMP) on a bounded sample of the same family on
this host's cores.

Two more legs go into the same line, each compared at identical costs with
the plain one-instance-at-a-time GPU path (Context.solve):
  "batched": 256 uneven random regular instances (n = round(2^U[10,16]),
             d ~ U{3..8}, k ~ U{6..12}, seeds 3000+i) solved by one Batch
             (shared launches, autotuned shape, hipGraph replay) vs one
             plain solve per instance with the graph already loaded;
  "served":  512 x 512 grid (w ~ U[1,100], seed 4), terminal-set queries
             (k ~ U{4..12}, seed 40) on a continuous-batching Server: a
             burst (all queries at t=0) and Poisson streams at rates swept
             as multiples of the plain path's serial capacity, with p50/p99
             latency from arrival to completion; the plain serial figures
             replay the same arrival times through a FIFO single server
             whose service times are the measured plain solve times.
A third leg, "group", solves the group Steiner workload through the
reduction (1024 x 1024 grid, w ~ U[1,1000], 13 groups of 32 vertices,
seed 5; virtual terminals joined by edges of weight M = 1 + sum of w) and
reports its time and OPT_group = OPT(reduced) - 13 M.

    python bench.py [--gpus 1] [--steps 3] [--warmup 1] [--no-cpu] [--no-traffic]
                    [--legs plain,batch,serve,group] [--queries 24]
"""
import argparse
import json
import os
import re
import subprocess
import sys
import tempfile
import time

REPO = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(REPO, "steiner-edge-linear_amd")
sys.path.insert(0, PKG)

HBM_PEAK_GBS = 8000.0
METRIC = "DP cells/s k=14 n=2^18 (% HBM roofline); served solves/s + p99 vs plain GPU"

WORKLOAD = dict(workload="EMV plain single instance: random connected 8-regular graph "
                "n=262144 m=1048576, w~U[1,1000], k=14 uniform terminals, seed 2",
                n=262144, d=8, wmax=1000, k=14, seed=2)


def log(msg):
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def cpu_baseline():
    """Reference reader-el (OpenMP, its own Makefile recipe minus
    -march=native) on an 8-regular n=65536 k=10 instance of the same family
    (about 10-30 s of CPU work); DP cells/s from its [kernel: ms] field."""
    ref = os.path.join(REPO, "oracle", "_ref", "reader-el")
    gen = os.path.join(PKG, "emv-gen")
    n, d, k = 65536, 8, 10
    threads = min(16, os.cpu_count() or 1, 127)
    sample = f"8-regular n={n} m={n * d // 2} w~U[1,1000] k={k} seed 2, full EMV kernel"
    with tempfile.TemporaryDirectory() as td:
        stp = os.path.join(td, "sample.stp")
        with open(stp, "w") as fh:
            subprocess.run([gen, "regular", str(n), str(d), "1000", str(k), "2"], stdout=fh,
                           check=True)
        if os.path.exists(ref):
            env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="spread")
            out = subprocess.run([ref, "-in", stp, "-el"], capture_output=True, text=True,
                                 env=env, cwd=td, timeout=600)
            line = [x for x in out.stdout.splitlines() if x.startswith("erickson")][0]
            ms = float(re.search(r"\[kernel: ([0-9.]+) ms", line).group(1))
            kind = "reference"
        else:
            sys.path.insert(0, os.path.join(REPO, "oracle"))
            import emv
            import oracle

            e = emv.regular_graph(n, d, 1000, 2)
            kk = emv.terminals(n, k, 2)
            t0 = time.perf_counter()
            oracle.emv(n, e, kk, threads)
            ms = (time.perf_counter() - t0) * 1e3
            kind = "port"
    cells = (1 << k) * n
    return dict(value=cells / (ms / 1e3), unit="DP cells/s", cores=threads, kind=kind,
                sample=sample, kernel_ms=ms)


def _pmc_avg(db, prefix):
    """Average counter value per dispatch (KiB) of the kernel instance whose
    name starts with `prefix` and that takes the most time, its name and its
    number of dispatches."""
    import collections
    import sqlite3

    cur = sqlite3.connect(db).cursor()
    rows = cur.execute("select name, counter_value, duration from pmc_events").fetchall()
    by = collections.defaultdict(lambda: [0.0, 0.0, 0])
    for name, v, dur in rows:
        short = name.replace("void ", "", 1).lstrip()
        if short.startswith(prefix):
            t = by[short.split("(")[0]]
            t[0] += v
            t[1] += dur
            t[2] += 1
    if not by:
        return None, None, 0
    name, (tot, _, cnt) = max(by.items(), key=lambda kv: kv[1][1])
    return tot / cnt, name, cnt


def measure_traffic(args, kernel_prefix):
    """HBM bytes per launch of the dominant kernel from two rocprofv3 counter
    passes (FETCH_SIZE, then WRITE_SIZE) over one solve of the same workload,
    each scaled by what the counter reads for a plain copy of known size
    (emv-copyprobe, 1 GiB read + 1 GiB written per launch).  None when
    rocprofv3 is missing or a pass fails."""
    import shutil

    prof = shutil.which("rocprofv3")
    probe = os.path.join(PKG, "emv-copyprobe")
    if not prof or not os.path.exists(probe):
        return None
    copy_bytes = 1 << 30
    out = {}
    with tempfile.TemporaryDirectory(dir="/tmp") as td:
        env = dict(os.environ, TMPDIR="/tmp")
        for ctr in ("FETCH_SIZE", "WRITE_SIZE"):
            runs = {
                "copy": [probe, str(copy_bytes), "3"],
                "solve": [sys.executable, os.path.abspath(__file__), "--legs", "plain", "--steps",
                          "1", "--warmup", "0", "--no-cpu", "--no-traffic", "--n", str(args.n),
                          "--k", str(args.k)],
            }
            for tag, cmd in runs.items():
                name = f"{tag}_{ctr}"
                log(f"traffic: rocprofv3 --pmc {ctr} ({tag})")
                r = subprocess.run([prof, "--pmc", ctr, "-d", td, "-o", name, "--"] + cmd, cwd=td,
                                   env=env, capture_output=True, text=True, timeout=600)
                db = os.path.join(td, f"{name}_results.db")
                if r.returncode != 0 or not os.path.exists(db):
                    log(f"traffic: {name} failed (rc={r.returncode})")
                    return None
                prefix = "k_copy16" if tag == "copy" else kernel_prefix
                out[name] = _pmc_avg(db, prefix)
    fetch_copy = out["copy_FETCH_SIZE"][0]
    write_copy = out["copy_WRITE_SIZE"][0]
    fetch, kname, nd = out["solve_FETCH_SIZE"]
    write = out["solve_WRITE_SIZE"][0]
    if not (fetch_copy and write_copy and fetch is not None and write is not None):
        return None
    fs = copy_bytes / (fetch_copy * 1024.0)
    ws = copy_bytes / (write_copy * 1024.0)
    return dict(kernel=kname, bytes=fetch * 1024.0 * fs + write * 1024.0 * ws,
                fetch_bytes=fetch * 1024.0 * fs, write_bytes=write * 1024.0 * ws,
                fetch_scale=fs, write_scale=ws, dispatches=nd)


def plain_leg(args):
    import emv

    wl = dict(WORKLOAD)
    wl["n"], wl["k"] = args.n, args.k
    t0 = time.perf_counter()
    e = emv.regular_graph(wl["n"], wl["d"], wl["wmax"], wl["seed"])
    kk = emv.terminals(wl["n"], wl["k"], wl["seed"])
    log(f"generated n={wl['n']} m={len(e)} k={wl['k']} in {time.perf_counter() - t0:.1f}s")
    ctx = emv.Context(0)
    ctx.load_edges(wl["n"], e)
    cost = None
    for i in range(args.warmup):
        sol = ctx.solve(kk)
        cost = sol.cost
        log(f"warmup {i} kernel={sol.kernel_ms:.1f} ms merge={sol.merge_ms:.1f} "
            f"relax={sol.relax_ms:.1f} single={sol.singleton_ms:.1f} cost={sol.cost}")
    sols = []
    t0 = time.perf_counter()
    for i in range(args.steps):
        sol = ctx.solve(kk)  # returns after the library's stream has drained
        sols.append(sol)
        log(f"step {i} kernel={sol.kernel_ms:.1f} ms merge={sol.merge_ms:.1f} "
            f"relax={sol.relax_ms:.1f} records={sol.unpack_ms:.1f} single={sol.singleton_ms:.1f} "
            f"rounds={sol.relax_rounds}")
    wall = time.perf_counter() - t0
    ctx.close()
    for s in sols:
        if cost is not None and s.cost != cost:
            raise RuntimeError(f"cost changed between steps: {s.cost} != {cost}")
    cost = sols[-1].cost
    cells = sols[-1].dp_cells
    merge_ms = sum(s.merge_ms for s in sols) / len(sols)
    relax_ms = sum(s.relax_ms for s in sols) / len(sols)
    unpack_ms = sum(s.unpack_ms for s in sols) / len(sols)
    # relax: the layer rounds (the singleton rounds are in singleton_ms);
    # its algorithmic bytes are the per-subset model without the k singleton rows
    k = sols[-1].k
    relax_bytes = sols[-1].relax_alg_bytes * ((1 << k) - 1 - k) / ((1 << k) - 1)
    kern = {
        "merge": dict(bytes=sols[-1].merge_alg_bytes, ms=merge_ms,
                      launches=sols[-1].merge_launches),
        "relax": dict(bytes=relax_bytes, ms=relax_ms, launches=sols[-1].relax_rounds),
    }
    for v in kern.values():
        v["achieved_GBs"] = v["bytes"] / (v["ms"] * 1e6) if v["ms"] > 0 else 0.0
        v["frac"] = v["achieved_GBs"] / HBM_PEAK_GBS
        v["cells_per_s"] = cells / (v["ms"] / 1e3) if v["ms"] > 0 else 0.0
        v["avg_launch_us"] = v["ms"] * 1e3 / max(v["launches"], 1)
    # The merge reads its lower rows from L2/L1 many times per X (cross-X
    # reuse), so the per-X streaming model puts its algorithmic bytes above
    # the HBM peak: it is VALU/cache-bound, not HBM-bound.
    split_ms = sum(s.merge_split_ms for s in sols) / len(sols)
    pairs = float(cells) / (1 << k) * (3 ** k - 2 ** (k + 1) + 1) / 2.0
    kern["merge"].update(
        split_ms=split_ms, whole_ms=merge_ms - split_ms,
        pair_cells_per_s=pairs / (merge_ms / 1e3) if merge_ms > 0 else 0.0,
        note="VALU/cache-bound: lower-layer rows are re-read from L2/L1 across X; the per-X "
             "streaming model exceeds the HBM peak")
    # dominant kernel = the kernel function with the most device time: the
    # relaxation round kernel (k_relax_v), or the split merge (k_merge_split)
    # or the whole-row merge (k_merge_v) when one of them outweighs it
    dom = "relax" if relax_ms >= max(split_ms, merge_ms - split_ms) else "merge"
    roof = dict(bound="hbm", kernel=dom, achieved=round(kern[dom]["achieved_GBs"], 1),
                peak=HBM_PEAK_GBS, unit="GB/s", frac=round(kern[dom]["frac"], 4), traffic=None,
                merge_relax_frac=round((kern["merge"]["bytes"] + kern["relax"]["bytes"]) /
                                       ((merge_ms + relax_ms) * 1e6) / HBM_PEAK_GBS, 4))
    if not args.no_traffic:
        tr = measure_traffic(args, "emv::k_relax" if dom == "relax" else "emv::k_merge")
        if tr and dom == "merge":
            tr = None  # the merge runs as several kernels per layer; no single per-launch figure
        if tr:
            launch_s = kern[dom]["ms"] / 1e3 / max(kern[dom]["launches"], 1)
            roof["traffic"] = tr["bytes"]
            roof["traffic_detail"] = tr
            roof["traffic_GBs"] = round(tr["bytes"] / launch_s / 1e9, 1)
            roof["traffic_frac"] = round(tr["bytes"] / launch_s / 1e9 / HBM_PEAK_GBS, 4)
    return dict(
        value=cells * args.steps / wall,
        ms_per_step=wall * 1e3 / args.steps,
        # table cells: 16-bit when the distance bound allows, else int32 / int64
        dtype={2: "uint16", 4: "int32", 8: "int64"}[sols[-1].table_bytes],
        config={"workload": wl["workload"] if (args.n, args.k) == (262144, 14)
                else f"8-regular n={args.n} k={args.k} seed 2",
                "n": wl["n"], "m": int(len(e)), "k": wl["k"], "dp_cells": cells,
                "table": {2: "uint16", 4: "int32", 8: "int64"}[sols[-1].table_bytes],
                "merge_acc_bits": sols[-1].merge_bits, "label_bits": sols[-1].label_bits},
        cost=cost,
        kernel_ms=sum(s.kernel_ms for s in sols) / len(sols),
        singleton_ms=sum(s.singleton_ms for s in sols) / len(sols),
        merge_ms=merge_ms, relax_ms=relax_ms, unpack_ms=unpack_ms, kernels=kern, roofline=roof)


def batch_leg(args):
    """256 uneven instances: plain per-instance solves vs one Batch."""
    import numpy as np

    import autotune
    import emv
    import workloads

    insts = workloads.batch_instances(args.instances)
    cells = sum((1 << len(kk)) * n for n, e, kk in insts)
    ctx = emv.Context(0)
    ctx.load_edges(insts[0][0], insts[0][1])
    ctx.solve(insts[0][2])  # warm: code objects, workspace
    plain_costs, t_solve, t_load = [], 0.0, 0.0
    for n, e, kk in insts:
        t0 = time.perf_counter()
        ctx.load_edges(n, e)
        t1 = time.perf_counter()
        plain_costs.append(ctx.solve(kk).cost)
        t_solve += time.perf_counter() - t1
        t_load += t1 - t0
    ctx.close()
    plain_costs = np.array(plain_costs)
    log(f"batch leg: plain {len(insts)} solves {t_solve:.3f}s (+{t_load:.3f}s graph loads)")
    b = emv.Batch(insts)
    t0 = time.perf_counter()
    params, table = autotune.tune_batch(b, insts, reps=1, log=log)
    tune_s = time.perf_counter() - t0
    times, used = [], []
    for i in range(args.batch_steps):
        t0 = time.perf_counter()
        costs, st = b.solve()
        times.append(time.perf_counter() - t0)
        used.append(int(st.graph_used))
        if not np.array_equal(costs, plain_costs):
            raise RuntimeError("batched costs differ from the plain per-instance costs")
        log(f"batch leg: solve {i} {times[-1]:.3f}s graph={st.graph_used} "
            f"fallback={st.graph_fallback} rounds={st.relax_rounds}")
    b.close()
    bt = sum(times) / len(times)
    return dict(
        workload=f"{len(insts)} random connected d-regular instances, n=round(2^U[10,16]), "
                 "d~U{3..8}, k~U{6..12}, w~U[1,1000], seeds 3000+i",
        instances=len(insts), dp_cells=cells, costs_equal=True,
        plain_s=t_solve, plain_graph_load_s=t_load,
        plain_instances_per_s=len(insts) / t_solve, plain_cells_per_s=cells / t_solve,
        batched_s=bt, batched_instances_per_s=len(insts) / bt, batched_cells_per_s=cells / bt,
        speedup=t_solve / bt, graph_replay=used,
        autotune=dict(relax_blocks_per_cu=params[0], merge_rows_per_wave=params[1],
                      seconds=round(tune_s, 2), table=table))


def group_leg(args):
    """Group Steiner through the reduction: 1024 x 1024 grid, 13 groups of 32."""
    import emv
    import workloads

    R = C = 1024
    g, size = 13, 32
    n, e, grp, nn, red, kk, M = workloads.group_instance(R, C, g, size, 1000, 5)
    ctx = emv.Context(0)
    t0 = time.perf_counter()
    ctx.load_edges(nn, red)
    t_load = time.perf_counter() - t0
    t0 = time.perf_counter()
    sol = ctx.solve(kk)  # warm: table and workspace allocations, code objects
    log(f"group leg: warm-up solve {time.perf_counter() - t0:.3f}s")
    times = []
    for i in range(args.group_steps):
        t0 = time.perf_counter()
        sol = ctx.solve(kk)
        times.append(time.perf_counter() - t0)
        log(f"group leg: solve {i} {times[-1]:.3f}s kernel={sol.kernel_ms:.1f} ms "
            f"merge={sol.merge_ms:.1f} relax={sol.relax_ms:.1f} records={sol.unpack_ms:.1f} "
            f"relative_label_layers={sol.relative_label_layers} cost={sol.cost}")
    ctx.close()
    return dict(
        workload=f"group Steiner via reduction: {R}x{C} grid, 4-neighbour, w~U[1,1000], "
                 f"{g} groups of {size} distinct vertices, seed 5; virtual terminals joined "
                 "by weight-M edges, M = 1 + sum of w",
        n=n, m=len(e), reduced_n=nn, reduced_m=len(red), k=len(kk), M=M,
        dp_cells=(1 << len(kk)) * nn, solve_s=sum(times) / len(times), graph_load_s=t_load,
        kernel_ms=sol.kernel_ms, merge_ms=sol.merge_ms, relax_ms=sol.relax_ms,
        records_ms=sol.unpack_ms, relax_rounds=sol.relax_rounds, table_bytes=sol.table_bytes,
        label_bits=sol.label_bits, relative_label_layers=sol.relative_label_layers,
        compact_merge_layers=sol.compact_merge_layers,
        relax_blocks=sol.relax_blocks, cost=sol.cost, group_cost=sol.cost - g * M)


def _percentile(xs, p):
    import numpy as np

    return float(np.percentile(np.asarray(xs), p)) if len(xs) else None


def _serve_stream(srv, qs, arr):
    """Submits query i at arrival time arr[i] (s, from now), steps the server
    while anything is in flight; returns latencies (arrival -> completion
    seen after a step), costs and the completion time of the last query."""
    Q = len(qs)
    lat, costs = [None] * Q, [None] * Q
    pending = {}
    i = 0
    t0 = time.perf_counter()
    last = 0.0
    while i < Q or pending:
        now = time.perf_counter() - t0
        while i < Q and arr[i] <= now:
            pending[srv.submit(qs[i])] = i
            i += 1
        if not pending:
            time.sleep(max(0.0, arr[i] - (time.perf_counter() - t0)))
            continue
        srv.step()
        tnow = time.perf_counter() - t0
        for tk in list(pending):
            info = srv.poll(tk)
            if info.state == 2:
                j = pending.pop(tk)
                if info.status != 0:
                    raise RuntimeError(f"served query {j} failed with status {info.status}")
                lat[j] = tnow - arr[j]
                costs[j] = info.cost
                last = tnow
    return lat, costs, last


def _fifo_latencies(arr, service):
    """Latencies of a FIFO single server with the given service times."""
    end, lat = 0.0, []
    for a, s in zip(arr, service):
        end = max(a, end) + s
        lat.append(end - a)
    return lat, end


def serve_leg(args):
    import numpy as np

    import emv
    import workloads

    R = C = 512
    n = R * C
    e = emv.grid_graph(R, C, 100, 4)
    qs = workloads.grid_queries(n, args.queries, 40)
    ctx = emv.Context(0)
    ctx.load_edges(n, e)
    ctx.solve(qs[0])  # warm
    service, plain_costs = [], []
    for q in qs:
        t0 = time.perf_counter()
        plain_costs.append(ctx.solve(q).cost)
        service.append(time.perf_counter() - t0)
    ctx.close()
    total = sum(service)
    log(f"serve leg: plain serial {len(qs)} queries {total:.2f}s "
        f"(p50 {_percentile(service, 50) * 1e3:.0f} ms)")
    srv = emv.Server(n, e, max_inflight=64)
    _serve_stream(srv, qs[:1], [0.0])  # warm
    lat, costs, last = _serve_stream(srv, qs, [0.0] * len(qs))
    if costs != plain_costs:
        raise RuntimeError("served costs differ from the plain costs (burst)")
    burst = dict(solves_per_s=len(qs) / last, plain_solves_per_s=len(qs) / total,
                 speedup=total / last, p50_ms=_percentile(lat, 50) * 1e3,
                 p99_ms=_percentile(lat, 99) * 1e3)
    log(f"serve leg: burst {last:.2f}s ({burst['speedup']:.2f}x plain)")
    sweep = []
    for idx, f in enumerate(args.rates):
        lam = f * len(qs) / total
        arr = list(workloads.poisson_arrivals(len(qs), lam, 40 + idx))
        lat, costs, last = _serve_stream(srv, qs, arr)
        if costs != plain_costs:
            raise RuntimeError(f"served costs differ from the plain costs (rate {lam:.3f}/s)")
        plat, pend = _fifo_latencies(arr, service)
        sweep.append(dict(rate_per_s=lam, rate_vs_plain_capacity=f,
                          solves_per_s=len(qs) / last, p50_ms=_percentile(lat, 50) * 1e3,
                          p99_ms=_percentile(lat, 99) * 1e3,
                          plain_solves_per_s=len(qs) / pend,
                          plain_p50_ms=_percentile(plat, 50) * 1e3,
                          plain_p99_ms=_percentile(plat, 99) * 1e3))
        log(f"serve leg: rate {lam:.2f}/s served p50 {sweep[-1]['p50_ms']:.0f} ms "
            f"p99 {sweep[-1]['p99_ms']:.0f} ms; plain p50 {sweep[-1]['plain_p50_ms']:.0f} ms "
            f"p99 {sweep[-1]['plain_p99_ms']:.0f} ms")
    srv.close()
    return dict(
        workload=f"512x512 grid, 4-neighbour, w~U[1,100], seed 4; {len(qs)} terminal-set "
                 "queries k~U{4..12}, seed 40",
        queries=len(qs), costs_equal=True, plain_serial_s=total,
        plain_service_p50_ms=_percentile(service, 50) * 1e3,
        plain_service_p99_ms=_percentile(service, 99) * 1e3,
        burst=burst, poisson=sweep,
        latency="arrival to completion, p99 over the stream's queries")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--n", type=int, default=WORKLOAD["n"])
    ap.add_argument("--k", type=int, default=WORKLOAD["k"])
    ap.add_argument("--legs", default="plain,batch,serve,group")
    ap.add_argument("--group-steps", type=int, default=2)
    ap.add_argument("--instances", type=int, default=256)
    ap.add_argument("--batch-steps", type=int, default=3)
    ap.add_argument("--queries", type=int, default=24)
    ap.add_argument("--rates", default="0.5,1,2")
    ap.add_argument("--no-traffic", action="store_true",
                    help="skip the rocprofv3 counter passes behind roofline.traffic")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); the "
              "multi-GPU sharding is described, not built", file=sys.stderr)
        sys.exit(2)
    args.rates = [float(x) for x in args.rates.split(",") if x]
    legs = set(args.legs.split(","))

    out = {"metric": METRIC, "value": None, "unit": "DP cells/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": None,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "int32",
           "data": "synthetic"}
    if "plain" in legs:
        p = plain_leg(args)
        out["value"] = p.pop("value")
        out["ms_per_step"] = p.pop("ms_per_step")
        out["dtype"] = p.pop("dtype")
        out["config"] = p.pop("config")
        out.update(p)
    if "batch" in legs:
        out["batched"] = batch_leg(args)
    if "serve" in legs:
        out["served"] = serve_leg(args)
    if "group" in legs:
        out["group"] = group_leg(args)
    if not args.no_cpu and "plain" in legs:
        log("cpu baseline (reference reader-el)...")
        out["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
