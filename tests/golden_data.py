"""Access to the committed golden fixtures (tests/golden/)."""
import hashlib
import json
import os

import numpy as np

from stputil import INF, Instance, parse_stp

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")

with open(os.path.join(GOLD, "ref_golden.json")) as _fh:
    META = json.load(_fh)
TESTSET = np.load(os.path.join(GOLD, "testset.npz"))
TABLES = np.load(os.path.join(GOLD, "tables.npz"))
NAMES = sorted(n for n in META if not n.startswith("regular"))
CFG1 = "regular_n1000_d4_k8_s1"


def load_instance(name):
    if name.startswith("regular"):
        return parse_stp(os.path.join(GOLD, name + ".stp"))
    key = name.replace("/", "__")
    return Instance(META[name]["n"], TESTSET[key + "__edges"].astype(np.int64),
                    TESTSET[key + "__terms"].astype(np.int64), META[name]["stp_cost"], name)


def golden_table(name):
    key = name.replace("/", "__")
    f = TABLES[key + "__f"].astype(np.int64)
    f[f == -1] = INF
    return f, TABLES[key + "__code"].astype(np.int64)


def digest(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype="<i8").tobytes()).hexdigest()


def kinds(code, k):
    out = np.array(code, dtype=np.int64, copy=True)
    out[out >= (1 << k)] = 1 << k
    return out
