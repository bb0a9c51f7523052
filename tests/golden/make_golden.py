"""Writes the golden fixtures of tests/golden/ from the reference itself.

Run in a container that holds the reference tree (default /root/reference)
after `make -C oracle ref`:

    python tests/golden/make_golden.py [--ref /root/reference]

For every SteinLib instance of the reference testset
(testset/small-instances/*/*.stp and testset/path/*.stp) it

  * runs oracle/_ref/reader-el (the reference's reader-el recipe) with
    `-el -list` and records the printed cost and solution pairs;
  * runs oracle/_ref/ref_dump (reference emv_kernel, called directly) and
    records SHA-256 digests of the full DP table (int64 LE, INF = INT64_MAX,
    rows X = 0 .. 2^k-1) and of the record kinds (Y for a merge record,
    2^k for an edge record, -1 unset);
  * stores the instance itself (edges, terminals, STP cost line) in
    testset.npz, since the GPU tests run where the reference tree is absent.

It also generates the 4-regular n=1000 k=8 instance (emv-gen, seed 1) and
stores its STP text, cost, solution and full table (tables.npz), together
with the full tables of every testset instance of at most 2^16 cells.
"""
import argparse
import glob
import hashlib
import json
import os
import re
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(HERE))
from stputil import parse_stp  # noqa: E402

REF_READER = os.path.join(REPO, "oracle", "_ref", "reader-el")
REF_DUMP = os.path.join(REPO, "oracle", "_ref", "ref_dump")
GEN = os.path.join(REPO, "steiner-edge-linear_amd", "emv-gen")


def table_digest(f):
    return hashlib.sha256(np.ascontiguousarray(f, dtype="<i8").tobytes()).hexdigest()


def kinds(code, k):
    out = code.copy()
    out[code >= (1 << k)] = 1 << k
    return out


def read_dump(path):
    raw = np.fromfile(path, dtype="<i8")
    magic, k, n, cost = (int(x) for x in raw[:4])
    assert magic == 0x454D5652454631
    cells = (1 << k) * n
    f = raw[4 : 4 + cells].reshape(1 << k, n)
    code = raw[4 + cells : 4 + 2 * cells].reshape(1 << k, n)
    return k, n, cost, f, code


def run_reader(stp):
    out = subprocess.run([REF_READER, "-in", stp, "-el", "-list"], capture_output=True, text=True,
                         cwd="/tmp", env=dict(os.environ, OMP_NUM_THREADS="4"))
    cost, sol = None, []
    for line in out.stdout.splitlines():
        if line.startswith("erickson"):
            cost = int(re.search(r"\[cost: (-?\d+)\]", line).group(1))
        elif line.startswith("solution"):
            sol = [tuple(int(x) for x in p.split()) for p in re.findall(r'"([^"]*)"', line)]
    return out.returncode, cost, sol


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--ref", default="/root/reference")
    args = ap.parse_args()
    files = sorted(glob.glob(os.path.join(args.ref, "testset/small-instances/*/*.stp")))
    files += sorted(glob.glob(os.path.join(args.ref, "testset/path/*.stp")))
    meta, arrays, tables = {}, {}, {}
    tmp = "/tmp/emv_golden_dump.bin"
    for path in files:
        rel = os.path.relpath(path, os.path.join(args.ref, "testset"))
        name = rel[:-4]
        inst = parse_stp(path)
        rc, rcost, sol = run_reader(path)
        subprocess.run([REF_DUMP, path, tmp], check=True, cwd="/tmp")
        k, n, dcost, f, code = read_dump(tmp)
        assert n == inst.n and k == inst.k
        meta[name] = dict(n=inst.n, m=inst.m, k=inst.k, stp_cost=inst.cost, ref_cost=rcost,
                          ref_rc=rc, kernel_cost=dcost, f_sha256=table_digest(f),
                          kind_sha256=table_digest(kinds(code, k)), ref_solution=sol)
        key = name.replace("/", "__")
        arrays[key + "__edges"] = inst.edges.astype(np.int32)
        arrays[key + "__terms"] = inst.terminals.astype(np.int32)
        if (1 << k) * n <= 1 << 16:
            tables[key + "__f"] = f
            tables[key + "__code"] = code
        print(f"{name}: n={n} k={k} stp_cost={inst.cost} ref={rcost} kernel={dcost} rc={rc}",
              flush=True)

    # the 4-regular n=1000 k=8 instance, seed 1 (resampled until simple and connected)
    cfg1 = subprocess.run([GEN, "regular", "1000", "4", "100", "8", "1", "reject"],
                          capture_output=True, text=True, check=True).stdout
    cfg1_path = os.path.join(HERE, "regular_n1000_d4_k8_s1.stp")
    with open(cfg1_path, "w") as fh:
        fh.write(cfg1)
    rc, rcost, sol = run_reader(cfg1_path)
    subprocess.run([REF_DUMP, cfg1_path, tmp], check=True, cwd="/tmp")
    k, n, dcost, f, code = read_dump(tmp)
    meta["regular_n1000_d4_k8_s1"] = dict(n=n, m=2000, k=k, stp_cost=-1, ref_cost=rcost,
                                          ref_rc=rc, kernel_cost=dcost,
                                          f_sha256=table_digest(f),
                                          kind_sha256=table_digest(kinds(code, k)),
                                          ref_solution=sol,
                                          stp_sha256=hashlib.sha256(cfg1.encode()).hexdigest())
    tables["regular_n1000_d4_k8_s1__f"] = f.astype(np.int64)
    tables["regular_n1000_d4_k8_s1__code"] = code.astype(np.int64)
    os.remove(tmp)

    with open(os.path.join(HERE, "ref_golden.json"), "w") as fh:
        json.dump(meta, fh, indent=0, sort_keys=True)
    np.savez_compressed(os.path.join(HERE, "testset.npz"), **arrays)
    # INF is stored as -1 so the int32 cast is lossless for these tables
    small = {}
    for key, val in tables.items():
        v = val.copy()
        v[v == np.iinfo(np.int64).max] = -1
        small[key] = v.astype(np.int32)
    np.savez_compressed(os.path.join(HERE, "tables.npz"), **small)
    print(f"{len(meta)} instances written")


if __name__ == "__main__":
    main()
