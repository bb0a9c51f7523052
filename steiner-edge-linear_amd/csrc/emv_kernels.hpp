// emv_kernels.hpp -- gfx950 device code of the EMV Steiner DP.
//
// Table layout in HBM (subset-major, as reader-el.c:690-692, rows padded):
//   f[X * np + v]   cost, int32 or int64         X in [0, 2^k), v in [0, np)
//   b[X * np + v]   int32 record code: Y (merge split, Y == X terminates)
//                   or 2^k + u (edge from predecessor u)
// np is n rounded up to a multiple of 1024; padding cells stay 0.
//
// Relaxation staging: one uint64 per cell of the layer batch,
//   S = (d << CB) | code,  code < 2^CB
// so that a single 64-bit atomicMin keeps the smallest distance and, among
// equal distances, the smallest code: the merge record (code = Y < 2^k) beats
// every edge record (code = 2^k + u), and edge records prefer the smallest u.
// That reproduces the reference's rule that the merge record survives unless
// the shortest path strictly improves on the merged value (reader-el.c:2122)
// and makes the records independent of the schedule.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace emv {

struct Edge {
    int32_t v;
    int32_t w;
};

constexpr int kBlock = 256;

typedef int32_t v4i32 __attribute__((ext_vector_type(4)));
typedef int64_t v2i64 __attribute__((ext_vector_type(2)));

template <typename T> struct VecOf;
template <> struct VecOf<int32_t> { static constexpr int N = 4; typedef v4i32 type; };
template <> struct VecOf<int64_t> { static constexpr int N = 2; typedef v2i64 type; };

// Work-item mapping that keeps neighbouring items on one XCD.  Workgroups
// are dealt to the XCDs round-robin (workgroup b runs on XCD b % nxcd), so
// XCD x receives the contiguous item range [x*per, (x+1)*per).
__device__ __forceinline__ int64_t xcd_item(int nxcd, int64_t per) {
    int64_t b = blockIdx.x;
    return (b % nxcd) * per + (b / nxcd);
}

// ---------------------------------------------------------------- fills
template <typename T>
__global__ void k_fill(T *__restrict__ p, int64_t count, T val) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; i < count; i += stride) p[i] = val;
}

// ------------------------------------------------------ singleton seeding
// Row t of the staging holds the SSSP from kk[t]: all INF except the source,
// whose code 1<<t is the reference's (kk[t], {t}) record (reader-el.c:2025).
__global__ void k_seed(uint64_t *__restrict__ S, int64_t np, const int32_t *__restrict__ src,
                       int nsrc, uint32_t *__restrict__ q, uint32_t *__restrict__ cnt,
                       uint32_t *__restrict__ stamp, uint32_t stamp_val, int VS,
                       uint32_t code_base_is_bit) {
    int t = threadIdx.x;
    if (t < nsrc) {
        int64_t cell = (int64_t)t * np + src[t];
        uint64_t code = code_base_is_bit ? (1ull << t) : 1ull;
        S[cell] = code;  // d = 0
        stamp[cell] = stamp_val;
        q[t] = ((uint32_t)t << VS) | (uint32_t)src[t];
    }
    if (t == 0) {
        cnt[0] = (uint32_t)nsrc;
        cnt[1] = 0;
        cnt[2] = 0;
    }
}

// Wave-aggregated append of `cell` to a queue (one atomic per wave).
__device__ __forceinline__ void wave_append(bool want, uint32_t entry, uint32_t *__restrict__ q,
                                            uint32_t *__restrict__ cnt) {
    uint64_t mask = __ballot(want);
    if (mask == 0) return;
    int lane = __lane_id();
    int leader = __ffsll((unsigned long long)mask) - 1;
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(cnt, (uint32_t)__popcll(mask));
    base = __shfl(base, leader);
    if (want) {
        uint64_t below = mask & ((1ull << lane) - 1ull);
        q[base + (uint32_t)__popcll(below)] = entry;
    }
}

// ------------------------------------------------------------ relaxation
// Pull sweep over every cell of a batch (first round of a layer).  Each
// thread owns one cell, so its write needs no atomic; concurrent readers see
// the old or the new label, both valid upper bounds.
__global__ __launch_bounds__(kBlock) void k_pull(uint64_t *__restrict__ S, int64_t np, int nrows,
                                                 int n, const int32_t *__restrict__ rp,
                                                 const Edge *__restrict__ adj, int CB,
                                                 uint64_t pred_base, uint32_t *__restrict__ qout,
                                                 uint32_t *__restrict__ cnt, int jout,
                                                 uint32_t *__restrict__ stamp, uint32_t stamp_val,
                                                 int VS, int nxcd, int64_t per_xcd,
                                                 int64_t tiles_per_row) {
    if (blockIdx.x == 0 && threadIdx.x == 0) cnt[(jout + 1) % 3] = 0;
    int64_t item = xcd_item(nxcd, per_xcd);
    if ((int64_t)(blockIdx.x / nxcd) >= per_xcd) return;
    int64_t row = item / tiles_per_row;
    int64_t tile = item - row * tiles_per_row;
    bool want = false;
    uint32_t entry = 0;
    int v = (int)(tile * kBlock + threadIdx.x);
    if (row < nrows && v < n) {
        uint64_t *Sr = S + row * np;
        uint64_t cur = Sr[v];
        uint64_t best = cur;
        const uint64_t dinf = ~0ull >> CB;
        for (int e = rp[v]; e < rp[v + 1]; e++) {
            Edge ed = adj[e];
            uint64_t du = Sr[ed.v] >> CB;
            if (du == dinf) continue;
            uint64_t cand = ((du + (uint64_t)ed.w) << CB) | (pred_base + (uint64_t)ed.v);
            best = cand < best ? cand : best;
        }
        if (best < cur) {
            Sr[v] = best;
            if ((best >> CB) < (cur >> CB)) {
                want = true;
                entry = ((uint32_t)row << VS) | (uint32_t)v;
                stamp[row * np + v] = stamp_val;
            }
        }
    }
    wave_append(want, entry, qout, &cnt[jout]);
}

// Push round: every queued cell offers d + w to its neighbours with a 64-bit
// atomicMin; a neighbour whose distance drops is queued once for the next
// round (stamp dedup).  Input/output counters rotate over cnt[0..2]; this
// round clears the counter the next round will write.
__global__ __launch_bounds__(kBlock) void k_push(uint64_t *__restrict__ S, int64_t np,
                                                 const int32_t *__restrict__ rp,
                                                 const Edge *__restrict__ adj, int CB,
                                                 uint64_t pred_base,
                                                 const uint32_t *__restrict__ qin,
                                                 uint32_t *__restrict__ qout,
                                                 uint32_t *__restrict__ cnt, int j,
                                                 uint32_t *__restrict__ stamp, uint32_t stamp_val,
                                                 int VS) {
    const uint32_t nin = cnt[j % 3];
    if (blockIdx.x == 0 && threadIdx.x == 0) cnt[(j + 2) % 3] = 0;
    uint32_t *cout = &cnt[(j + 1) % 3];
    const uint32_t vmask = (1u << VS) - 1u;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    // Round the trip count up to a whole number of waves so that every lane
    // of a wave reaches the wave-wide append together.
    for (uint64_t base = (uint64_t)blockIdx.x * blockDim.x; base < nin; base += stride) {
        uint64_t i = base + threadIdx.x;
        bool active = i < nin;
        uint32_t ent = active ? qin[i] : 0u;
        int64_t row = ent >> VS;
        int u = (int)(ent & vmask);
        uint64_t *Sr = S + row * np;
        uint64_t du = active ? (Sr[u] >> CB) : 0;
        int e0 = active ? rp[u] : 0, e1 = active ? rp[u + 1] : 0;
        uint64_t code = pred_base + (uint64_t)u;
        // Lanes walk their own adjacency; the append is done once per edge
        // slot with all lanes of the wave present.
        int deg = e1 - e0;
        int maxdeg = deg;
        for (int off = 32; off > 0; off >>= 1) {
            int o = __shfl_xor(maxdeg, off);
            maxdeg = o > maxdeg ? o : maxdeg;
        }
        for (int t = 0; t < maxdeg; t++) {
            bool want = false;
            uint32_t entry = 0;
            if (t < deg) {
                Edge ed = adj[e0 + t];
                uint64_t cand = ((du + (uint64_t)ed.w) << CB) | code;
                uint64_t *dst = Sr + ed.v;
                if (cand < *dst) {
                    uint64_t old = atomicMin((unsigned long long *)dst, (unsigned long long)cand);
                    if ((cand >> CB) < (old >> CB)) {
                        uint32_t prev = atomicMax(&stamp[row * np + ed.v], stamp_val);
                        if (prev < stamp_val) {
                            want = true;
                            entry = ((uint32_t)row << VS) | (uint32_t)ed.v;
                        }
                    }
                }
            }
            wave_append(want, entry, qout, cout);
        }
    }
}

// ------------------------------------------------------------------ merge
// f[X][v] = min over nonempty Y strictly inside X of f[Y][v] + f[X\Y][v]
// (reader-el.c:2068-2088).  Only Y without the top bit of X is enumerated:
// the pair {Y, X\Y} is visited once and Y is the smaller of the two.  The
// accumulator packs (sum << k) | Y, so an unsigned min keeps the smallest
// sum and, among equal sums, the smallest Y -- the reference's first strict
// minimum over increasing Y.  Each workgroup handles one X on a tile of
// 256*VEC vertices; the tiles of one XCD are consecutive, so the rows a tile
// needs are shared in that XCD's L2.
template <typename T, typename P>
__global__ __launch_bounds__(kBlock) void k_merge(const T *__restrict__ f, int64_t np,
                                                  const uint32_t *__restrict__ xs, int nx,
                                                  int64_t ntiles, int k, int CB,
                                                  uint64_t *__restrict__ S, int nxcd,
                                                  int64_t per_xcd) {
    constexpr int VEC = VecOf<T>::N;
    typedef typename VecOf<T>::type VT;
    if ((int64_t)(blockIdx.x / nxcd) >= per_xcd) return;
    int64_t item = xcd_item(nxcd, per_xcd);
    int64_t tile = item / nx;
    int xi = (int)(item - tile * nx);
    if (tile >= ntiles) return;
    const uint32_t X = xs[xi];
    const uint32_t top = 1u << (31 - __clz(X));
    const uint32_t Xp = X ^ top;
    const int64_t v0 = tile * (kBlock * VEC) + (int64_t)threadIdx.x * VEC;
    P acc[VEC];
#pragma unroll
    for (int j = 0; j < VEC; j++) acc[j] = ~(P)0;
    for (uint32_t Y = Xp; Y != 0; Y = (Y - 1) & Xp) {
        VT a = *(const VT *)(f + (int64_t)Y * np + v0);
        VT c = *(const VT *)(f + (int64_t)(X ^ Y) * np + v0);
#pragma unroll
        for (int j = 0; j < VEC; j++) {
            P s = (((P)a[j] + (P)c[j]) << k) | (P)Y;
            acc[j] = s < acc[j] ? s : acc[j];
        }
    }
    uint64_t *Sr = S + (int64_t)xi * np + v0;
    const P ymask = ((P)1 << k) - 1;
#pragma unroll
    for (int j = 0; j < VEC; j++) {
        uint64_t d = (uint64_t)(acc[j] >> k);
        uint64_t y = (uint64_t)(acc[j] & ymask);
        Sr[j] = (d << CB) | y;
    }
}

// ----------------------------------------------------------------- unpack
// Staging rows -> table rows f[X], b[X] for a batch of layer rows.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_unpack(const uint64_t *__restrict__ S, int64_t np,
                                                   const uint32_t *__restrict__ xs, int nrows,
                                                   int n, int CB, T *__restrict__ f,
                                                   int32_t *__restrict__ b, int nxcd,
                                                   int64_t per_xcd, int64_t tiles_per_row) {
    if ((int64_t)(blockIdx.x / nxcd) >= per_xcd) return;
    int64_t item = xcd_item(nxcd, per_xcd);
    int64_t row = item / tiles_per_row;
    int64_t tile = item - row * tiles_per_row;
    int v = (int)(tile * kBlock + threadIdx.x);
    if (row >= nrows || v >= n) return;
    uint64_t s = S[row * np + v];
    int64_t X = xs[row];
    f[X * np + v] = (T)(s >> CB);
    b[X * np + v] = (int32_t)(s & ((1ull << CB) - 1ull));
}

// Singleton rows: f[1<<t][v] = dist(kk[t], v); record (kk[t], {t}) for every
// v (reader-el.c:2015-2027).  Also the per-row maximum distance (bounds the
// table: f[X][v] <= sum over t in X of dist(kk[t], v)).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_unpack_singletons(
    const uint64_t *__restrict__ S, int64_t np, const int32_t *__restrict__ kk, int k, int n,
    int CB, T *__restrict__ f, int32_t *__restrict__ b, unsigned long long *__restrict__ rowmax) {
    int t = blockIdx.y;
    int v = blockIdx.x * kBlock + threadIdx.x;
    unsigned long long d = 0;
    if (v < n) {
        d = S[(int64_t)t * np + v] >> CB;
        int64_t X = 1ll << t;
        if (f) {
            f[X * np + v] = (T)d;
            b[X * np + v] = (v == kk[t]) ? (int32_t)X : (int32_t)(((int64_t)1 << k) + kk[t]);
        }
    }
    // wave max then one atomic per wave
    for (int off = 32; off > 0; off >>= 1) {
        unsigned long long o = __shfl_xor(d, off);
        d = o > d ? o : d;
    }
    if (__lane_id() == 0) atomicMax(&rowmax[t], d);
}

// Distances and predecessors of one SSSP row (for -dijkstra / k == 2).
__global__ void k_unpack_sssp(const uint64_t *__restrict__ S, int n, int CB, uint64_t pred_base,
                              int64_t *__restrict__ d, int64_t *__restrict__ pred) {
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n) return;
    uint64_t s = S[v];
    uint64_t dd = s >> CB;
    uint64_t code = s & ((1ull << CB) - 1ull);
    d[v] = (dd == (~0ull >> CB)) ? INT64_MAX : (int64_t)dd;
    pred[v] = (code >= pred_base) ? (int64_t)(code - pred_base) : -1;
}

// -------------------------------------------------------------- traceback
// Depth-first walk of the records from (q, C) in the reference's order
// (reader-el.c:1907-1932): an edge record emits (v,u) and continues at
// (u, X); a merge record visits (v, Y) then (v, X\Y); Y == X stops.
// One lane walks; the records are a few hundred steps for real trees.
__global__ void k_traceback(const int32_t *__restrict__ b, int64_t np, int k, int64_t q,
                            int64_t C, int64_t *__restrict__ pairs, int64_t cap,
                            int64_t *__restrict__ stack, int64_t stcap,
                            int64_t *__restrict__ result) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const int64_t S = 1ll << k;
    int64_t sp = 0, np_ = 0;
    stack[0] = q;
    stack[1] = C;
    sp = 1;
    while (sp > 0) {
        sp--;
        int64_t v = stack[2 * sp], X = stack[2 * sp + 1];
        if (X == 0 || v < 0) continue;
        int64_t r = b[X * np + v];
        if (r < 0) continue;
        if (r >= S) {
            int64_t u = r - S;
            if (np_ >= cap) { result[0] = -1; return; }
            pairs[2 * np_] = v;
            pairs[2 * np_ + 1] = u;
            np_++;
            stack[2 * sp] = u;
            stack[2 * sp + 1] = X;
            sp++;
        } else {
            if (r == X) continue;
            if (sp + 2 > stcap) { result[0] = -2; return; }
            stack[2 * sp] = v;
            stack[2 * sp + 1] = X & ~r;
            sp++;
            stack[2 * sp] = v;
            stack[2 * sp + 1] = r;
            sp++;
        }
    }
    result[0] = np_;
}

}  // namespace emv
