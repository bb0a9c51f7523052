"""Benchmark of the EMV Steiner DP on one MI355X.

A step is one full EMV solve (singleton shortest paths, every merge and
relaxation layer; no traceback) of the plain single-instance workload:
random connected 8-regular graph, n = 262,144, m = 1,048,576, weights
uniform in [1, 1000], k = 14 uniform terminals (generator seed 2), a table
of 2^14 x 2^18 cells (16-bit: the distance bound is below 0xFFFF).  The
graph is resident in HBM before timing.

Prints one JSON line.  `value` = DP cells (2^k * n) per second of wall time
over the timed steps.  `roofline` covers the dominant kernel (the
relaxation round kernel), with achieved = algorithmic bytes / kernel time
from HIP events on the library's stream, and traffic = HBM bytes per launch
measured by two rocprofv3 counter passes (FETCH_SIZE, WRITE_SIZE) scaled
against a plain copy of known size.  The merge is reported beside it
against its measured HBM traffic and its VALU floor (it re-reads lower rows
from cache, so its per-X streaming model is no HBM figure).  `table32`
repeats the solve with an int32 table.  `cpu_baseline` times the reference
reader-el (oracle/_ref, OpenMP on every physical core up to 127) on a
bounded sample of the same family, with the CPU model and a row on the
n = 1,000 4-regular k = 8 instance.

More legs go into the same line, each compared at identical costs with the
plain one-instance-at-a-time GPU path (Context.solve):
  "batched": 256 uneven random regular instances (n = round(2^U[10,16]),
             d ~ U{3..8}, k ~ U{6..12}, seeds 3000+i) solved by one Batch
             (shared launches, autotuned shape, hipGraph replay; the eager
             first solve is reported too) vs one plain solve per instance
             with the graph already loaded;
  "served":  512 x 512 grid (w ~ U[1,100], seed 4), terminal-set queries
             (k ~ U{4..12}, seed 40) and group queries on a
             continuous-batching Server: a burst (all queries at t=0) and
             Poisson streams at rates swept as multiples of the plain path's
             serial capacity (up to saturation), with p50/p99 latency from
             arrival to completion; the plain serial figures replay the same
             arrival times through a FIFO single server whose service times
             are the measured plain solve times;
  "group":   the group Steiner workload (1024 x 1024 grid, w ~ U[1,1000],
             13 groups of 32 vertices, seed 5) solved through the reduction
             (virtual terminals joined by edges of weight M = 1 + sum of w)
             with its listed tree validated as a group Steiner tree, and the
             same query on the Server's resident grid (the DP on the grid
             itself, no int64 table).

    python bench.py [--gpus 1] [--steps 3] [--warmup 1] [--no-cpu] [--no-traffic]
                    [--legs plain,batch,serve,group] [--queries 24]
"""
import argparse
import json
import os
import re
import signal
import subprocess
import sys
import tempfile
import time

REPO = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(REPO, "steiner-edge-linear_amd")
sys.path.insert(0, PKG)

HBM_PEAK_GBS = 8000.0
# dense integer VALU: 256 CUs x 4 SIMDs x 16 lanes x 2.4 GHz lane-ops per second
VALU_LANE_OPS = 256 * 64 * 2.4e9
METRIC = "DP cells/s k=14 n=2^18 (% HBM roofline); served solves/s + p99 vs plain GPU"

WORKLOAD = dict(workload="EMV plain single instance: random connected 8-regular graph "
                "n=262144 m=1048576, w~U[1,1000], k=14 uniform terminals, seed 2",
                n=262144, d=8, wmax=1000, k=14, seed=2)


def _comb(n, r):
    import math

    return math.comb(n, r)


def log(msg):
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def _lscpu():
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=30).stdout
    except (OSError, subprocess.SubprocessError):
        return {}
    info = {}
    for line in out.splitlines():
        if ":" in line:
            k, v = line.split(":", 1)
            info[k.strip()] = v.strip()
    return info


def physical_cores(info):
    try:
        return int(info["Core(s) per socket"]) * int(info["Socket(s)"])
    except (KeyError, ValueError):
        return os.cpu_count() or 1


def _run_reference(stp, threads, timeout=900, binary="reader-el"):
    ref = os.path.join(REPO, "oracle", "_ref", binary)
    env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="spread")
    out = subprocess.run([ref, "-in", stp, "-el"], capture_output=True, text=True, env=env,
                         cwd=os.path.dirname(stp), timeout=timeout)
    line = [x for x in out.stdout.splitlines() if x.startswith("erickson")][0]
    kms = float(re.search(r"\[kernel: ([0-9.]+) ms", line).group(1))
    done = float(re.search(r"done\. \[([0-9.]+) ms\]", line).group(1))
    cost = int(re.search(r"\[cost: (-?[0-9]+)\]", line).group(1))
    return kms, done, cost


def cpu_baseline(config2=False):
    """Reference reader-el (OpenMP on min(physical cores, 127) threads,
    reader-el.c:62,2213; its Makefile recipe minus -march=native) on an
    8-regular n=65536 k=10 instance of the same family (the bounded sample,
    about 10-30 s of CPU work), on the n=1000 4-regular k=8 instance next to
    the GPU solve of the same file, and with config2 on the full n=2^18 k=14
    instance; DP cells/s from its [kernel: ms] field."""
    import emv

    info = _lscpu()
    threads = min(physical_cores(info), 127)
    model = info.get("Model name", "unknown")
    ref = os.path.join(REPO, "oracle", "_ref", "reader-el")
    gen = os.path.join(PKG, "emv-gen")
    n, d, k = 65536, 8, 10
    sample = f"8-regular n={n} m={n * d // 2} w~U[1,1000] k={k} seed 2, full EMV kernel"
    out = dict(unit="DP cells/s", cores=threads, cpu_model=model,
               threads_per_core=info.get("Thread(s) per core"), sample=sample)
    with tempfile.TemporaryDirectory() as td:
        stp = os.path.join(td, "sample.stp")
        with open(stp, "w") as fh:
            subprocess.run([gen, "regular", str(n), str(d), "1000", str(k), "2"], stdout=fh,
                           check=True)
        if os.path.exists(ref):
            kms, done, cost0 = _run_reference(stp, threads)
            out.update(kind="reference", kernel_ms=kms, done_ms=done)
            # the same reference build at -O3 (the published recipe has no -O
            # flag), a separately labelled row on the same sample
            if os.path.exists(ref + "-O3"):
                kms3, done3, cost3 = _run_reference(stp, threads, binary="reader-el-O3")
                if cost3 != cost0:
                    raise RuntimeError(f"-O3 reference cost {cost3} != {cost0}")
                out["O3"] = dict(kernel_ms=kms3, done_ms=done3,
                                 value=(1 << k) * n / (kms3 / 1e3))
        else:
            sys.path.insert(0, os.path.join(REPO, "oracle"))
            import oracle

            e = emv.regular_graph(n, d, 1000, 2)
            kk = emv.terminals(n, k, 2)
            t0 = time.perf_counter()
            oracle.emv(n, e, kk, threads)
            kms = (time.perf_counter() - t0) * 1e3
            out.update(kind="port", kernel_ms=kms)
        out["value"] = (1 << k) * n / (kms / 1e3)
        # the n=1000 4-regular k=8 instance (the reference's own STP file and
        # the GPU solve of the same graph)
        cfg1 = os.path.join(REPO, "tests", "golden", "regular_n1000_d4_k8_s1.stp")
        if os.path.exists(ref) and os.path.exists(cfg1):
            kms1, done1, cost1 = _run_reference(cfg1, threads)
            sys.path.insert(0, os.path.join(REPO, "tests"))
            from stputil import parse_stp

            inst = parse_stp(cfg1)
            with emv.Context(0) as ctx:
                ctx.load_edges(inst.n, inst.edges)
                ctx.solve(inst.terminals)  # warm
                sols = [ctx.solve(inst.terminals) for _ in range(5)]
            gms = min(s.kernel_ms for s in sols)
            if sols[-1].cost != cost1:
                raise RuntimeError(f"n=1000 instance: GPU cost {sols[-1].cost} != reference {cost1}")
            cells = (1 << inst.k) * inst.n
            out["n1000_k8"] = dict(workload="4-regular n=1000 m=2000 w~U[1,100] k=8 seed 1 "
                                   "(tests/golden)", cost=cost1, cpu_kernel_ms=kms1,
                                   cpu_done_ms=done1, cpu_cells_per_s=cells / (kms1 / 1e3),
                                   gpu_kernel_ms=gms, gpu_cells_per_s=cells / (gms / 1e3))
        if config2 and os.path.exists(ref):
            stp2 = os.path.join(td, "cfg2.stp")
            with open(stp2, "w") as fh:
                subprocess.run([gen, "regular", "262144", "8", "1000", "14", "2"], stdout=fh,
                               check=True)
            log("cpu baseline: reference on n=2^18 k=14 (minutes)")
            kms2, done2, cost2 = _run_reference(stp2, threads, timeout=3000)
            out["n262144_k14"] = dict(cost=cost2, kernel_ms=kms2, done_ms=done2,
                                      cells_per_s=(1 << 14) * 262144 / (kms2 / 1e3))
    return out


def _pmc_rows(db):
    import sqlite3

    cur = sqlite3.connect(db).cursor()
    return cur.execute("select name, counter_value, duration from pmc_events").fetchall()


def _short(name):
    return name.replace("void ", "", 1).lstrip().split("(")[0]


def _pmc_avg(rows, prefix):
    """Average counter value per dispatch (KiB) of the kernel instance whose
    name starts with `prefix` and that takes the most time, its name and its
    number of dispatches."""
    import collections

    by = collections.defaultdict(lambda: [0.0, 0.0, 0])
    for name, v, dur in rows:
        short = _short(name)
        if short.startswith(prefix):
            t = by[short]
            t[0] += v
            t[1] += dur
            t[2] += 1
    if not by:
        return None, None, 0
    name, (tot, _, cnt) = max(by.items(), key=lambda kv: kv[1][1])
    return tot / cnt, name, cnt


def _pmc_sum(rows, prefixes):
    """Counter total (KiB) over every dispatch whose name starts with one of
    the prefixes."""
    return sum(v for name, v, _ in rows if _short(name).startswith(prefixes))


def _run_group(cmd, cwd, env, timeout):
    """Runs cmd in its own process group and reaps the whole group after it
    returns, so that no helper it started outlives it."""
    p = subprocess.Popen(cmd, cwd=cwd, env=env, stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                         text=True, start_new_session=True)
    try:
        out, err = p.communicate(timeout=timeout)
    except subprocess.TimeoutExpired:
        os.killpg(p.pid, signal.SIGKILL)
        out, err = p.communicate()
    try:
        os.killpg(p.pid, signal.SIGKILL)
    except ProcessLookupError:
        pass
    return p.returncode, out, err


MERGE_KERNELS = ("emv::k_merge", "emv::k_pack")


def measure_traffic(args, kernel_prefix):
    """HBM bytes of one solve from two rocprofv3 counter passes (FETCH_SIZE,
    then WRITE_SIZE), each scaled by what the counter reads for a plain copy
    of known size (emv-copyprobe, 1 GiB read + 1 GiB written per launch): per
    launch of the dominant kernel, and per solve summed over the merge
    kernels.  None when rocprofv3 is missing or a pass fails."""
    import shutil

    prof = shutil.which("rocprofv3")
    probe = os.path.join(PKG, "emv-copyprobe")
    if not prof or not os.path.exists(probe):
        return None
    copy_bytes = 1 << 30
    rows = {}
    with tempfile.TemporaryDirectory(dir="/tmp") as td:
        env = dict(os.environ, TMPDIR="/tmp")
        for ctr in ("FETCH_SIZE", "WRITE_SIZE"):
            runs = {
                "copy": [probe, str(copy_bytes), "3"],
                "solve": [sys.executable, os.path.abspath(__file__), "--legs", "plain", "--steps",
                          "1", "--warmup", "0", "--no-cpu", "--no-traffic", "--no-table32",
                          "--no-visits",
                          "--n", str(args.n), "--k", str(args.k)],
            }
            for tag, cmd in runs.items():
                name = f"{tag}_{ctr}"
                log(f"traffic: rocprofv3 --pmc {ctr} ({tag})")
                rc, _, _ = _run_group([prof, "--pmc", ctr, "-d", td, "-o", name, "--"] + cmd,
                                      td, env, 600)
                db = os.path.join(td, f"{name}_results.db")
                if rc != 0 or not os.path.exists(db):
                    log(f"traffic: {name} failed (rc={rc})")
                    return None
                rows[name] = _pmc_rows(db)
    fetch_copy = _pmc_avg(rows["copy_FETCH_SIZE"], "k_copy16")[0]
    write_copy = _pmc_avg(rows["copy_WRITE_SIZE"], "k_copy16")[0]
    fetch, kname, nd = _pmc_avg(rows["solve_FETCH_SIZE"], kernel_prefix)
    write = _pmc_avg(rows["solve_WRITE_SIZE"], kernel_prefix)[0]
    if not (fetch_copy and write_copy and fetch is not None and write is not None):
        return None
    fs = copy_bytes / (fetch_copy * 1024.0)
    ws = copy_bytes / (write_copy * 1024.0)
    mf = _pmc_sum(rows["solve_FETCH_SIZE"], MERGE_KERNELS) * 1024.0 * fs
    mw = _pmc_sum(rows["solve_WRITE_SIZE"], MERGE_KERNELS) * 1024.0 * ws
    return dict(kernel=kname, bytes=fetch * 1024.0 * fs + write * 1024.0 * ws,
                fetch_bytes=fetch * 1024.0 * fs, write_bytes=write * 1024.0 * ws,
                fetch_scale=fs, write_scale=ws, dispatches=nd,
                merge_solve_bytes=mf + mw, merge_fetch_bytes=mf, merge_write_bytes=mw)


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
    # one more solve counting the relaxation's item visits (not timed)
    visits = -1
    if not args.no_visits:
        vsol = ctx.solve(kk, emv.COUNT_VISITS)
        visits = vsol.relax_item_visits
        if vsol.cost != sols[-1].cost:
            raise RuntimeError("visit-counting solve changed the cost")
    t32 = None
    if not args.no_table32:
        # the same solve on an int32 table (the 16-bit cells are exact because
        # the distance bound is below 0xFFFF; this is the int32 figure)
        ctx.solve(kk, emv.TABLE32)
        s32 = []
        t1 = time.perf_counter()
        for _ in range(2):
            s32.append(ctx.solve(kk, emv.TABLE32))
        w32 = (time.perf_counter() - t1) / len(s32)
        if any(s.cost != sols[-1].cost for s in s32):
            raise RuntimeError("int32 table changed the cost")
        t32 = dict(table="int32", ms_per_step=w32 * 1e3, cells_per_s=s32[-1].dp_cells / w32,
                   merge_ms=s32[-1].merge_ms, relax_ms=s32[-1].relax_ms,
                   records_ms=s32[-1].unpack_ms, label_bits=s32[-1].label_bits, cost=s32[-1].cost)
        log(f"table32: {w32 * 1e3:.1f} ms per solve")
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
    relax = dict(bytes=relax_bytes, ms=relax_ms, launches=sols[-1].relax_rounds)
    relax["achieved_GBs"] = relax_bytes / (relax_ms * 1e6) if relax_ms > 0 else 0.0
    relax["frac"] = relax["achieved_GBs"] / HBM_PEAK_GBS
    relax["cells_per_s"] = cells / (relax_ms / 1e3) if relax_ms > 0 else 0.0
    relax["avg_launch_us"] = relax_ms * 1e3 / max(relax["launches"], 1)
    # item visits: a (chunk, vertex) item is the chunk's rows (128 or 64) of
    # one vertex; one pass over every item of layers 2..k is the one-pass
    # model's work
    cw = sols[-1].relax_chunk_rows or 64
    one_pass = sum(-(-_comb(k, mm) // cw) for mm in range(2, k + 1)) * wl["n"]
    relax["chunk_rows"] = cw
    relax["item_visits"] = visits
    relax["one_pass_items"] = one_pass
    relax["visits_per_item"] = visits / one_pass if visits > 0 else None
    deg = wl["d"]
    lb = sols[-1].label_bits // 8
    relax["visit_bytes_model"] = cw * lb * (deg + 2) + 8 * deg  # own rows read+write, neighbours, adjacency
    relax["visit_bytes_per_solve"] = visits * relax["visit_bytes_model"] if visits > 0 else None
    # merge: unordered pair-cells; its per-X streaming model counts lower rows
    # re-read from cache, so it is reported as a model ratio, not a roofline
    # fraction; the merge is bounded by its VALU work (about two lane-ops per
    # pair-cell with two 16-bit cells per lane) and by its measured traffic
    split_ms = sum(s.merge_split_ms for s in sols) / len(sols)
    pairs = float(cells) / (1 << k) * (3 ** k - 2 ** (k + 1) + 1) / 2.0
    valu_floor_ms = pairs * 2.0 / VALU_LANE_OPS * 1e3
    merge = dict(ms=merge_ms, launches=sols[-1].merge_launches, split_ms=split_ms,
                 whole_ms=merge_ms - split_ms, pair_cells=pairs,
                 pair_cells_per_s=pairs / (merge_ms / 1e3) if merge_ms > 0 else 0.0,
                 cells_per_s=cells / (merge_ms / 1e3) if merge_ms > 0 else 0.0,
                 model_bytes=sols[-1].merge_alg_bytes,
                 model_ratio=sols[-1].merge_alg_bytes / (merge_ms * 1e6) / HBM_PEAK_GBS,
                 valu_floor_ms=valu_floor_ms, valu_floor_frac=valu_floor_ms / merge_ms,
                 note="per-X streaming model over the HBM peak: lower rows are re-read from "
                      "L2/L1 across X, so model_ratio is no roofline fraction; measured_* is HBM")
    roof = dict(bound="hbm", kernel="relax", achieved=round(relax["achieved_GBs"], 1),
                peak=HBM_PEAK_GBS, unit="GB/s", frac=round(relax["frac"], 4), traffic=None)
    if not args.no_traffic:
        tr = measure_traffic(args, "emv::k_relax")
        if tr:
            launch_s = relax_ms / 1e3 / max(relax["launches"], 1)
            roof["traffic"] = tr["bytes"]
            roof["traffic_detail"] = tr
            roof["traffic_GBs"] = round(tr["bytes"] / launch_s / 1e9, 1)
            roof["traffic_frac"] = round(tr["bytes"] / launch_s / 1e9 / HBM_PEAK_GBS, 4)
            merge["measured_bytes"] = tr["merge_solve_bytes"]
            merge["measured_GBs"] = tr["merge_solve_bytes"] / (merge_ms * 1e6)
            merge["measured_frac"] = merge["measured_GBs"] / HBM_PEAK_GBS
    roof["merge"] = {key: merge[key] for key in ("ms", "valu_floor_ms", "valu_floor_frac")
                     if key in merge}
    for key in ("measured_GBs", "measured_frac"):
        if key in merge:
            roof["merge"][key] = round(merge[key], 4)
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
        merge_ms=merge_ms, relax_ms=relax_ms, unpack_ms=unpack_ms,
        kernels=dict(merge=merge, relax=relax), roofline=roof, table32=t32)


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
    t0 = time.perf_counter()
    b = emv.Batch(insts)
    create_s = time.perf_counter() - t0
    # the first solve of a fresh batch: eager steps (default shape), rounds recorded
    t0 = time.perf_counter()
    costs, st = b.solve()
    first_s = time.perf_counter() - t0
    if not np.array_equal(costs, plain_costs):
        raise RuntimeError("batched costs differ from the plain per-instance costs (first solve)")
    log(f"batch leg: first (eager) solve {first_s:.3f}s")
    t0 = time.perf_counter()
    if args.batch_no_graph:
        # eager steps only (no hipGraph capture or replay), default shape
        shape, table = dict(autotune.DEFAULTS), []
        b.set_params(use_graph=0, **shape)
    else:
        shape, table = autotune.tune_batch(b, insts, reps=1, log=log)
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
    # the tuned shape without the replay (eager steps)
    b.set_params(use_graph=0, **shape)
    t0 = time.perf_counter()
    costs, _ = b.solve()
    eager_s = time.perf_counter() - t0
    if not np.array_equal(costs, plain_costs):
        raise RuntimeError("batched costs differ from the plain per-instance costs (eager)")
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
        first_solve_s=first_s, first_solve_speedup=t_solve / first_s,
        eager_tuned_s=eager_s, eager_tuned_speedup=t_solve / eager_s, create_s=create_s,
        autotune=dict(shape=shape, seconds=round(tune_s, 2), evaluated=len(table), table=table))


def _group_queries(n, count, seed):
    import numpy as np

    rng = np.random.default_rng(seed)
    return [[rng.choice(n, size=int(rng.integers(4, 17)), replace=False)
             for _ in range(int(rng.integers(3, 9)))] for _ in range(count)]


def group_leg(args):
    """Group Steiner through the reduction: 1024 x 1024 grid, 13 groups of 32;
    then the same query on a Server holding the grid."""
    import emv
    import workloads

    sys.path.insert(0, os.path.join(REPO, "tests"))
    from bigcheck import validate_group_tree

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
    lsol = ctx.solve(kk, emv.LIST)
    ctx.close()
    ok, why, gcost, _ = validate_group_tree(n, e, grp, nn, red, M, lsol.pairs, lsol.cost)
    if not ok or lsol.cost != sol.cost:
        raise RuntimeError(f"group leg: listed tree invalid: {why}")
    log(f"group leg: tree validated, group cost {gcost}")
    # the same query on a Server holding the grid (the DP on the grid itself)
    srv = emv.Server(n, e, max_inflight=4)
    served = []
    for i in range(1 + args.group_steps):
        t0 = time.perf_counter()
        t = srv.submit_groups([list(x) for x in grp])
        while srv.poll(t).state != 2:
            srv.step()
        served.append(time.perf_counter() - t0)
        info = srv.poll(t)
        if info.status != 0 or info.reduced_cost != sol.cost:
            raise RuntimeError(f"group leg: served cost {info.reduced_cost} != {sol.cost}")
    srv.close()
    log(f"group leg: served on the grid {min(served[1:]):.3f}s")
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
        relax_blocks=sol.relax_blocks, cost=sol.cost, group_cost=sol.cost - g * M,
        validated=True, tree_pairs=int(len(lsol.pairs)),
        served_on_grid_s=sum(served[1:]) / len(served[1:]),
        served_note="Server.submit_groups on the resident grid: the DP runs on the grid with "
                    "singleton rows = distance to the nearest member (int32 cells), same cost")


def _percentile(xs, p):
    import numpy as np

    return float(np.percentile(np.asarray(xs), p)) if len(xs) else None


def _serve_stream(srv, qs, arr):
    """Submits query i at arrival time arr[i] (s, from now), steps the server
    while anything is in flight; returns latencies (arrival -> completion
    seen after a step), reduced costs and the completion time of the last
    query.  A query is a terminal list or ("groups", member lists)."""
    Q = len(qs)
    lat, costs = [None] * Q, [None] * Q
    pending = {}
    i = 0
    t0 = time.perf_counter()
    last = 0.0
    while i < Q or pending:
        now = time.perf_counter() - t0
        while i < Q and arr[i] <= now:
            q = qs[i]
            tk = srv.submit_groups(q[1]) if isinstance(q, tuple) else srv.submit(q)
            pending[tk] = i
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
                costs[j] = info.reduced_cost
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
    import emv
    import workloads

    R = C = 512
    n = R * C
    e = emv.grid_graph(R, C, 100, 4)
    qs = list(workloads.grid_queries(n, args.queries, 40))
    # group queries join the stream, spread over it
    gq = _group_queries(n, args.group_queries, 42)
    for j, g in enumerate(gq):
        qs.insert((j + 1) * len(qs) // (len(gq) + 1), ("groups", g))
    ctx = emv.Context(0)
    ctx.load_edges(n, e)
    ctx.solve(qs[0])  # warm
    service, plain_costs, group_loads = [], [], 0.0
    for q in qs:
        if isinstance(q, tuple):
            # plain path: the reference's reduction, then a solve on the reduced graph
            nn, red, kk, _ = emv.group_reduce(n, e, q[1])
            t0 = time.perf_counter()
            ctx.load_edges(nn, red)
            t1 = time.perf_counter()
            plain_costs.append(ctx.solve(kk).cost)
            service.append(time.perf_counter() - t1)
            group_loads += t1 - t0
            ctx.load_edges(n, e)
        else:
            t0 = time.perf_counter()
            plain_costs.append(ctx.solve(q).cost)
            service.append(time.perf_counter() - t0)
    ctx.close()
    total = sum(service)
    log(f"serve leg: plain serial {len(qs)} queries {total:.2f}s "
        f"(p50 {_percentile(service, 50) * 1e3:.0f} ms; group graph loads {group_loads:.2f}s "
        "not counted)")
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
                          plain_p99_ms=_percentile(plat, 99) * 1e3,
                          throughput_vs_plain=(len(qs) / last) / (len(qs) / pend)))
        log(f"serve leg: rate {lam:.2f}/s ({f}x) served {len(qs) / last:.1f}/s p50 "
            f"{sweep[-1]['p50_ms']:.0f} ms p99 {sweep[-1]['p99_ms']:.0f} ms; plain "
            f"{len(qs) / pend:.1f}/s p50 {sweep[-1]['plain_p50_ms']:.0f} ms p99 "
            f"{sweep[-1]['plain_p99_ms']:.0f} ms")
    srv.close()
    small = _serve_small(args, n, e) if args.small_queries > 0 else None
    return dict(
        workload=f"512x512 grid, 4-neighbour, w~U[1,100], seed 4; {args.queries} terminal-set "
                 f"queries k~U{{4..12}}, seed 40, and {len(gq)} group queries (3-8 groups of "
                 "4-16 vertices, seed 42); plain path = one solve per query (group queries "
                 "through the reduction, graph loads not counted)",
        queries=len(qs), group_queries=len(gq), costs_equal=True, plain_serial_s=total,
        plain_service_p50_ms=_percentile(service, 50) * 1e3,
        plain_service_p99_ms=_percentile(service, 99) * 1e3,
        burst=burst, poisson=sweep, target_speedup=5.0, small_queries=small,
        latency="arrival to completion, p99 over the stream's queries")


def _serve_small(args, n, e):
    """The small-query regime on the same grid: terminal sets with k ~
    U{4..8} (seed 43), where one query's launches leave most of the GPU idle.
    Burst and a Poisson stream at 4x the plain serial capacity."""
    import emv
    import workloads

    qs = list(workloads.grid_queries(n, args.small_queries, 43, kmin=4, kmax=8))
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
    srv = emv.Server(n, e, max_inflight=128)
    _serve_stream(srv, qs[:1], [0.0])  # warm
    lat, costs, last = _serve_stream(srv, qs, [0.0] * len(qs))
    if costs != plain_costs:
        raise RuntimeError("served costs differ from the plain costs (small queries, burst)")
    out = dict(workload=f"{len(qs)} terminal-set queries k~U{{4..8}}, seed 43, same grid",
               queries=len(qs), plain_serial_s=total, plain_solves_per_s=len(qs) / total,
               burst=dict(solves_per_s=len(qs) / last, speedup=total / last,
                          p50_ms=_percentile(lat, 50) * 1e3, p99_ms=_percentile(lat, 99) * 1e3))
    lam = 4.0 * len(qs) / total
    arr = list(workloads.poisson_arrivals(len(qs), lam, 49))
    lat, costs, last = _serve_stream(srv, qs, arr)
    if costs != plain_costs:
        raise RuntimeError("served costs differ from the plain costs (small queries, stream)")
    plat, pend = _fifo_latencies(arr, service)
    out["poisson_4x"] = dict(rate_per_s=lam, solves_per_s=len(qs) / last,
                             p50_ms=_percentile(lat, 50) * 1e3, p99_ms=_percentile(lat, 99) * 1e3,
                             plain_solves_per_s=len(qs) / pend,
                             plain_p50_ms=_percentile(plat, 50) * 1e3,
                             plain_p99_ms=_percentile(plat, 99) * 1e3,
                             throughput_vs_plain=(len(qs) / last) / (len(qs) / pend))
    srv.close()
    log(f"serve leg: small queries burst {out['burst']['speedup']:.2f}x plain, stream at 4x "
        f"{out['poisson_4x']['throughput_vs_plain']:.2f}x, p99 {out['poisson_4x']['p99_ms']:.0f} ms")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-config2", action="store_true",
                    help="also time the reference on the full n=2^18 k=14 instance (minutes)")
    ap.add_argument("--n", type=int, default=WORKLOAD["n"])
    ap.add_argument("--k", type=int, default=WORKLOAD["k"])
    ap.add_argument("--legs", default="plain,batch,serve,group")
    ap.add_argument("--group-steps", type=int, default=2)
    ap.add_argument("--instances", type=int, default=256)
    ap.add_argument("--batch-steps", type=int, default=3)
    ap.add_argument("--queries", type=int, default=24)
    ap.add_argument("--group-queries", type=int, default=6)
    ap.add_argument("--small-queries", type=int, default=64)
    ap.add_argument("--rates", default="0.5,1,2,4")
    ap.add_argument("--no-traffic", action="store_true",
                    help="skip the rocprofv3 counter passes behind roofline.traffic")
    ap.add_argument("--no-table32", action="store_true")
    ap.add_argument("--no-visits", action="store_true",
                    help="skip the extra solve that counts the relaxation's item visits")
    ap.add_argument("--batch-no-graph", action="store_true",
                    help="batch leg with eager steps only (no hipGraph capture), for profiling")
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
        out["cpu_baseline"] = cpu_baseline(args.cpu_config2)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
