"""The synthetic workloads, built with the seeded generator (libemvgen.so).

regular_single(n, d, wmax, k, seed)   random connected d-regular graph + k terminals
batch_instances(count, seed0)          uneven instances: n = round(2^U[10,16]),
                                       d ~ U{3..8}, k ~ U{6..12}, w ~ U[1,1000]
grid_queries(count, seed)              k ~ U{4..12} uniform terminal sets on an
                                       n-vertex graph
poisson_arrivals(count, rate, seed)    arrival times (s) with exponential gaps
group_instance(rows, cols, g, size, wmax, seed)
                                       grid + g groups of `size` distinct vertices,
                                       reduced to a Steiner instance
"""
import math

import numpy as np

import emv

STREAM_PARAMS = 7
STREAM_QUERY = 8
STREAM_ARRIVAL = 9


def _unit(seed, stream, index):
    return (emv.uniform(seed, stream, index) >> 11) * (1.0 / (1 << 53))


def regular_single(n, d, wmax, k, seed, method="repair"):
    e = emv.regular_graph(n, d, wmax, seed, method)
    return n, e, emv.terminals(n, k, seed)


def batch_params(i, seed0=3000):
    s = seed0 + i
    n = int(round(2.0 ** (10.0 + 6.0 * _unit(s, STREAM_PARAMS, 0))))
    d = 3 + int(6 * _unit(s, STREAM_PARAMS, 1))
    if (n * d) % 2:
        n += 1
    k = 6 + int(7 * _unit(s, STREAM_PARAMS, 2))
    return n, d, k, s


def batch_instances(count, seed0=3000, wmax=1000):
    out = []
    for i in range(count):
        n, d, k, s = batch_params(i, seed0)
        out.append(regular_single(n, d, wmax, k, s))
    return out


def grid_queries(n, count, seed, kmin=4, kmax=12):
    qs = []
    for i in range(count):
        k = kmin + int((kmax - kmin + 1) * _unit(seed, STREAM_QUERY, i))
        qs.append(emv.distinct(n, k, seed * 1000 + i, 2))
    return qs


def poisson_arrivals(count, rate, seed):
    t, out = 0.0, []
    for i in range(count):
        t += -math.log(1.0 - _unit(seed, STREAM_ARRIVAL, i)) / rate
        out.append(t)
    return np.array(out)


def group_instance(rows, cols, g, size, wmax, seed):
    e = emv.grid_graph(rows, cols, wmax, seed)
    n = rows * cols
    grp = emv.groups(n, g, size, seed)
    nn, red, kk, M = emv.group_reduce(n, e, grp)
    return n, e, grp, nn, red, kk, M
