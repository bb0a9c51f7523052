// emv_kernels.hpp -- gfx950 device code of the EMV Steiner DP.
//
// Table in HBM (subset-major as reader-el.c:690-692, rows padded to np, an
// odd number of 64-vertex tiles; the padding columns v >= n hold whatever
// the allocation held and only feed lanes whose results are dropped):
//   f[X * np + v]   cost: 16-bit (0xFFFF = INF), int32 or int64, the
//                   narrowest that holds the distance bound
//   b[X * np + v]   int32 record code: Y for a merge split (Y == X stops the
//                   traceback), 2^k + u for an edge from predecessor u
// An int64 table may carry a compact int32 copy for the merges,
// fcopy[X * np + v] = f[X][v] - fbase[X].
//
// Layer staging (vertex-major): the rows of one layer batch are grouped in
// chunks of 64; chunk c holds, for every vertex v, the 64 rows' labels
// side by side:
//   D[(c * np + v) * 64 + lane]     distance label of row c*64+lane at v
//                                   (or its offset from the row's base)
//   CODE[(c * np + v) * 64 + lane]  the merge's packed word (sum << k) | Y
// One wave owns one (chunk, vertex) item: lane = row.  A neighbour's labels
// for the whole chunk are one coalesced 128 B (u16), 256 B (u32) or 512 B
// (u64) load, and the adjacency of v is read once for 64 rows.
//
// Relaxation rounds move labels only.  Records are decided once, from the
// converged labels, when the staging is unpacked: the merge record (split
// Y) stays when the label still equals the merged value -- the reference
// keeps it exactly when no path strictly improves on the merge
// (reader-el.c:2117-2126) -- and otherwise the record is the edge from the
// smallest neighbour u with label(u) + w(u, v) == label(v).  That is the
// lexicographic minimum of (distance, code) over the merge candidate
// (merged value, Y) and the edge candidates (label(u) + w, 2^k + u), with
// merge codes below every edge code, so it does not depend on the order in
// which the rounds ran.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cstring>
#include <type_traits>
#include <vector>

namespace emv {

struct Edge {
    int32_t v;
    int32_t w;
};

constexpr int kBlock = 256;
constexpr int kWave = 64;

// Work-item mapping that keeps neighbouring items on one XCD.  Workgroups
// are dealt to the XCDs round-robin (workgroup b runs on XCD b % nxcd), so
// XCD x receives the contiguous item range [x*per, (x+1)*per).
__device__ __forceinline__ int64_t xcd_item(int nxcd, int64_t per) {
    int64_t b = blockIdx.x;
    return (b % nxcd) * per + (b / nxcd);
}

// Cell layout of a DP table: cell (X, v) at (v >> ts) * tcells + X * rcells +
// (v & (2^ts - 1)).  Rows (ts = 62, tcells = 0, rcells = np): row X is
// contiguous, the reference's f[X*n + v] with a padded stride.  Tiles
// (16-bit tables, ts = 7): the 2^k rows of a 128-vertex tile are one
// contiguous block of 2^k * 256 bytes (tcells = 2^k * 128, rcells = 128), so
// the merges, which read one tile's blocks of many rows, stay within a few
// pages instead of touching one page per row.
struct FLayout {
    int64_t tcells;
    int64_t rcells;
    int32_t ts;
    int32_t pad_;
    __host__ __device__ int64_t cell(int64_t X, int64_t v) const {
        return (v >> ts) * tcells + X * rcells + (v & ((1ll << ts) - 1));
    }
};

inline FLayout rows_layout(int64_t np) { return FLayout{0, np, 62, 0}; }
// tiles of 2^ts vertices (7 for 16-bit cells, 6 for 32-bit: 256 bytes a row)
inline FLayout tiles_layout(int k, int ts = 7) {
    return FLayout{(int64_t)1 << (k + ts), (int64_t)1 << ts, ts, 0};
}

// ---------------------------------------------------------------- fills
template <typename T>
__global__ void k_fill(T *__restrict__ p, int64_t count, T val) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; i < count; i += stride) p[i] = val;
}

// Row X of a table in layout fl, cells v < nv, set to val.
template <typename T>
__global__ void k_fill_row(T *__restrict__ f, FLayout fl, int64_t X, int64_t nv, T val) {
    int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; v < nv; v += stride) f[fl.cell(X, v)] = val;
}

// ------------------------------------------------------------- seeding
// Chunk 0, lane t: the shortest-path labels from src[t] (label 0 at the
// source; every other label starts at INF).
template <typename DT>
__global__ void k_seed_v(DT *__restrict__ D, const int32_t *__restrict__ src, int nsrc) {
    int t = threadIdx.x;
    if (t < nsrc) D[(int64_t)src[t] * kWave + t] = 0;
}

// ------------------------------------------------------------ relaxation
// One pull round over (chunk, vertex) items, item i = (chunk i / n, vertex
// i % n).  Every item whose neighbour's label dropped in the previous round
// is flagged (a word in F_cur); round 0 (`dense`) takes every item.  A round
// takes its items in one of two ways, decided on the device from the
// previous round's counters:
//   worklist : the flagged items were also appended to Q_cur (cnt[r&3] of
//              them, at most qcap) -- a thin frontier costs only its size;
//   flag scan: too many were flagged for the list, so the wave reads 64
//              consecutive flags at once and walks the set ones.
// Per item the wave recomputes, for its 64 rows, the minimum of its own label
// and (neighbour label + w) over the whole adjacency and writes the rows that
// dropped.  If any label dropped it flags every neighbour in F_next (a writer
// that finds the flag clear also appends the neighbour to Q_next while the
// list has room) and raises any[(r+1)&3].  Each flagged item is processed by
// exactly one wave per round: the flag scan reads each flag once, and a
// listed item is taken only by the wave whose atomic exchange clears its
// flag, so duplicate list entries are skipped.  A round clears the flags it
// consumed, so F_next is clean when the round starts, and clears
// any/cnt[(r+2)&3] for the round after next.  Labels only decrease and every
// drop re-flags the neighbours, so the rounds end at the shortest-path
// fixed point whatever order the waves run in.
//
// Adjacency: DMAX > 0 uses the padded fixed-degree copy ell[v * DMAX + j]
// (missing neighbours have v = -1).  The 64 items a wave holds load their
// rows into registers with one coalesced load, so an item's neighbour labels
// are gathered with no dependent load in front, and two items are processed
// together to keep 2 * DMAX gathers in flight.  DMAX == 0 walks the CSR.
template <typename DT>
__device__ __forceinline__ DT relax_min(DT du, int u, int w, DT INF, DT best) {
    if (u >= 0 && du != INF) {
        const DT nd = du + (DT)w;
        best = nd < best ? nd : best;
    }
    return best;
}

// Packed pairs (PK): with 16-bit labels in chunks of 128 rows, one 32-bit
// word holds the labels of rows 2j and 2j + 1 of a (chunk, vertex) item, and
// a lane takes both: a neighbour's labels for the whole chunk are one 256-byte
// load (random 256-byte reads run at 1.4x the bytes per second of 128-byte
// ones), and one packed saturating add and one packed min relax two rows.
// Every finite label + w stays below 0xFFFF (the 16-bit label bound), and
// INF (0xFFFF) + w saturates to INF, so the halves never interact.
typedef unsigned short emv_u16x2 __attribute__((ext_vector_type(2)));

template <typename DT, bool PK>
__device__ __forceinline__ DT relax_lbl(DT du, int u, int w, DT INF, DT best) {
    if constexpr (PK) {
        const emv_u16x2 ww = {(unsigned short)w, (unsigned short)w};
        const emv_u16x2 nd = __builtin_elementwise_add_sat(__builtin_bit_cast(emv_u16x2, du), ww);
        return __builtin_bit_cast(DT, __builtin_elementwise_min(__builtin_bit_cast(emv_u16x2, best), nd));
    } else {
        return relax_min(du, u, w, INF, best);
    }
}

// Whether some label of the item dropped (labels only decrease).
template <typename DT, bool PK>
__device__ __forceinline__ bool lbl_dropped(DT best, DT cur) {
    if constexpr (PK) return best != cur;
    else return best < cur;
}

// Whether a row that dropped now improves the neighbour whose labels are d
// (label + w below the neighbour's label).  A dropped label is finite.
template <typename DT, bool PK>
__device__ __forceinline__ bool lbl_improves(DT best, DT cur, int w, DT d) {
    if constexpr (PK) {
        const emv_u16x2 b = __builtin_bit_cast(emv_u16x2, best);
        const emv_u16x2 c = __builtin_bit_cast(emv_u16x2, cur);
        const emv_u16x2 dd = __builtin_bit_cast(emv_u16x2, d);
        const emv_u16x2 ww = {(unsigned short)w, (unsigned short)w};
        const auto m = (b < c) & (__builtin_elementwise_add_sat(b, ww) < dd);
        return (m.x | m.y) != 0;
    } else {
        return best < cur && (DT)(best + (DT)w) < d;
    }
}

// Wave-aggregated append of `entry` to a list of capacity cap (one atomic per
// wave).  Every lane of the wave must call it.  Entries past cap are dropped;
// the counter still counts them, which tells the next round to scan flags.
__device__ __forceinline__ void wave_append(bool want, uint32_t entry, uint32_t *__restrict__ q,
                                            uint32_t *__restrict__ cnt, uint32_t cap) {
    const uint64_t mask = __ballot(want);
    if (mask == 0) return;
    const int lane = __lane_id();
    const int leader = __ffsll((unsigned long long)mask) - 1;
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(cnt, (uint32_t)__popcll(mask));
    base = __shfl(base, leader);
    if (want) {
        const uint32_t at = base + (uint32_t)__popcll(mask & ((1ull << lane) - 1ull));
        if (at < cap) q[at] = entry;
    }
}

// Flags one neighbour per lane (nb < 0: none) of a changed item of chunk c;
// a neighbour whose flag was clear is also appended to the list while it has
// room.  Writers racing on one flag may both append it; the consumer's
// atomic exchange lets only one of them through.
__device__ __forceinline__ void flag_neighbours(int64_t c, int64_t np, int n, int nb,
                                                uint32_t *__restrict__ Fnext,
                                                uint32_t *__restrict__ Qnext,
                                                uint32_t *__restrict__ cnext, uint32_t qcap,
                                                bool list_open) {
    if (!list_open) {
        // flag scan only: a plain store, no read of the flag (list_open is
        // wave-uniform)
        if (nb >= 0) Fnext[c * np + nb] = 1;
        return;
    }
    bool want = false;
    uint32_t ent = 0;
    if (nb >= 0) {
        uint32_t *fp = Fnext + c * np + nb;
        if (*fp == 0) {
            *fp = 1;
            want = true;
            ent = (uint32_t)(c * n + nb);
        }
    }
    wave_append(want, ent, Qnext, cnext, qcap);
}

// Dense stretch (dense_rounds = R > 0): rounds 0 .. R-1 write no flags and
// rounds 1 .. R take every item.  Early in a layer almost every item is
// flagged anyway, so the few extra visits cost less than the random flag
// stores they replace.  Round R writes flags again for round R + 1; the
// rounds without flags leave F_cur / F_next clean, as a dense round reads
// no flags.
template <typename DT, int DMAX, bool PK = false>
__global__ __launch_bounds__(kBlock) void k_relax_v(
    DT *__restrict__ D, int64_t np, int n, const int32_t *__restrict__ rp,
    const Edge *__restrict__ adj, const Edge *__restrict__ ell, int64_t nitems, int dense0,
    uint32_t *__restrict__ Fcur, uint32_t *__restrict__ Fnext, const uint32_t *__restrict__ Qcur,
    uint32_t *__restrict__ Qnext, uint32_t qcap, uint32_t *__restrict__ any,
    uint32_t *__restrict__ cnt, int r, int nxcd, int dense_rounds,
    unsigned long long *__restrict__ visits) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        any[(r + 2) & 3] = 0;
        cnt[(r + 2) & 3] = 0;
    }
    const bool dense = dense0 || (r >= 1 && r <= dense_rounds);
    const bool noflags = r < dense_rounds;  // the next round takes every item
    if (!dense && any[r & 3] == 0) return;  // nothing flagged for this round
    const uint32_t qn = dense ? 0u : cnt[r & 3];
    const bool use_list = !dense && qn < qcap;
    uint32_t *cnext = &cnt[(r + 1) & 3];
    const int lane = threadIdx.x & (kWave - 1);
    const int wid = threadIdx.x / kWave;
    // nxcd > 1: each XCD sweeps its own contiguous range of item blocks;
    // nxcd == 1: the whole grid sweeps them in order.
    const int64_t waves_per_block = kBlock / kWave;
    const int64_t per = gridDim.x / nxcd;
    const int64_t b = blockIdx.x;
    const int64_t wave0 = ((b % nxcd) * per + b / nxcd) * waves_per_block + wid;
    const int64_t wstride = (int64_t)gridDim.x * waves_per_block;
    const int64_t span = use_list ? (int64_t)qn : nitems;
    const int64_t nblk = (span + kWave - 1) / kWave;
    const DT INF = (DT)~(DT)0;
    bool raised = false;
    for (int64_t blk = wave0; blk < nblk; blk += wstride) {
        const int64_t j = blk * kWave + lane;
        int64_t mi = -1;
        if (j < span) mi = use_list ? (int64_t)Qcur[j] : j;
        int mc = 0, mv = 0;
        bool fl = false;
        if (mi >= 0) {
            mc = (int)(mi / n);
            mv = (int)(mi - (int64_t)mc * n);
            const int64_t mf = (int64_t)mc * np + mv;
            if (dense) {
                fl = true;
            } else if (use_list) {
                fl = atomicExch(&Fcur[mf], 0u) != 0;  // one of the listed copies
            } else if (Fcur[mf] != 0) {
                fl = true;
                Fcur[mf] = 0;
            }
        }
        uint64_t act = __ballot(fl);
        if (act == 0) continue;
        // optional count of the items visited (one atomic per 64 items)
        if (visits && lane == 0) atomicAdd(visits, (unsigned long long)__popcll(act));
        const bool list_open = *(volatile uint32_t *)cnext < qcap;
        if constexpr (DMAX > 0) {
            Edge e[DMAX];
#pragma unroll
            for (int q = 0; q < DMAX; q++) e[q] = Edge{-1, 0};
            if (fl) {
#pragma unroll
                for (int q = 0; q < DMAX; q++) e[q] = ell[(int64_t)mv * DMAX + q];
            }
            while (act) {
                const int t1 = __ffsll((unsigned long long)act) - 1;
                act &= act - 1;
                const bool two = act != 0;
                const int t2 = two ? __ffsll((unsigned long long)act) - 1 : t1;
                if (two) act &= act - 1;
                const int64_t ca = __builtin_amdgcn_readlane(mc, t1);
                const int64_t cb = __builtin_amdgcn_readlane(mc, t2);
                const int va = __builtin_amdgcn_readlane(mv, t1);
                const int vb = __builtin_amdgcn_readlane(mv, t2);
                DT *Da = D + ca * np * kWave;
                DT *Db = D + cb * np * kWave;
                const int64_t ia = (int64_t)va * kWave + lane, ib = (int64_t)vb * kWave + lane;
                const DT cura = Da[ia];
                const DT curb = two ? Db[ib] : INF;
                int ua[DMAX], wa[DMAX], ub[DMAX], wb[DMAX];
                DT da[DMAX], db[DMAX];
#pragma unroll
                for (int q = 0; q < DMAX; q++) {
                    ua[q] = __builtin_amdgcn_readlane(e[q].v, t1);
                    wa[q] = __builtin_amdgcn_readlane(e[q].w, t1);
                    ub[q] = __builtin_amdgcn_readlane(e[q].v, t2);
                    wb[q] = __builtin_amdgcn_readlane(e[q].w, t2);
                }
#pragma unroll
                for (int q = 0; q < DMAX; q++) {
                    da[q] = ua[q] >= 0 ? Da[(int64_t)ua[q] * kWave + lane] : INF;
                    db[q] = (two && ub[q] >= 0) ? Db[(int64_t)ub[q] * kWave + lane] : INF;
                }
                DT besta = cura, bestb = curb;
#pragma unroll
                for (int q = 0; q < DMAX; q++) {
                    besta = relax_lbl<DT, PK>(da[q], ua[q], wa[q], INF, besta);
                    bestb = relax_lbl<DT, PK>(db[q], ub[q], wb[q], INF, bestb);
                }
                const bool chga = lbl_dropped<DT, PK>(besta, cura);
                const bool chgb = two && lbl_dropped<DT, PK>(bestb, curb);
                if (chga) Da[ia] = besta;
                if (chgb) Db[ib] = bestb;
                const bool anya = __ballot(chga) != 0, anyb = __ballot(chgb) != 0;
                if (noflags) {
                    if (anya || anyb) raised = true;
                } else if (anya || anyb) {
                    // Only neighbours that a dropped row now improves are
                    // flagged: u is left alone when label + w >= u's label
                    // (as read above, and labels only drop) in every dropped
                    // row.  A row that did not drop offers u nothing new: its
                    // last drop was checked the same way.  The dropped labels
                    // are finite and label + w stays below INF.
                    uint32_t fa = 0, fb = 0;
#pragma unroll
                    for (int q = 0; q < DMAX; q++) {
                        const bool ia = chga && lbl_improves<DT, PK>(besta, cura, wa[q], da[q]);
                        const bool ib = chgb && lbl_improves<DT, PK>(bestb, curb, wb[q], db[q]);
                        if (__ballot(ia) != 0) fa |= 1u << q;
                        if (__ballot(ib) != 0) fb |= 1u << q;
                    }
                    // lane q < DMAX takes neighbour q
                    int na = -1, nb = -1;
#pragma unroll
                    for (int q = 0; q < DMAX; q++) {
                        if (lane == q) {
                            na = (fa >> q) & 1u ? ua[q] : -1;
                            nb = (fb >> q) & 1u ? ub[q] : -1;
                        }
                    }
                    if (fa) flag_neighbours(ca, np, n, na, Fnext, Qnext, cnext, qcap, list_open);
                    if (fb) flag_neighbours(cb, np, n, nb, Fnext, Qnext, cnext, qcap, list_open);
                    if (fa | fb) raised = true;
                }
            }
        } else {
            int me0 = 0, me1 = 0;
            if (fl) {
                me0 = rp[mv];
                me1 = rp[mv + 1];
            }
            while (act) {
                const int t = __ffsll((unsigned long long)act) - 1;
                act &= act - 1;
                const int64_t c = __builtin_amdgcn_readlane(mc, t);
                const int v = __builtin_amdgcn_readlane(mv, t);
                DT *Dc = D + c * np * kWave;
                const int64_t idx = (int64_t)v * kWave + lane;
                const DT cur = Dc[idx];
                DT best = cur;
                const int e0 = __builtin_amdgcn_readlane(me0, t);
                const int e1 = __builtin_amdgcn_readlane(me1, t);
                for (int eb = e0; eb < e1; eb += kWave) {
                    const int cnt_e = min(kWave, e1 - eb);
                    Edge my = {0, 0};
                    if (lane < cnt_e) my = adj[eb + lane];
                    for (int t0 = 0; t0 < cnt_e; t0 += 4) {
                        // four neighbours per step: their loads issue together
                        DT du[4];
                        int uu[4], ww[4];
#pragma unroll
                        for (int q = 0; q < 4; q++) {
                            const int tt = min(t0 + q, cnt_e - 1);
                            uu[q] = __builtin_amdgcn_readlane(my.v, tt);
                            ww[q] = __builtin_amdgcn_readlane(my.w, tt);
                            du[q] = Dc[(int64_t)uu[q] * kWave + lane];
                        }
#pragma unroll
                        for (int q = 0; q < 4; q++)
                            if (t0 + q < cnt_e) best = relax_lbl<DT, PK>(du[q], uu[q], ww[q], INF, best);
                    }
                }
                const bool dchg = lbl_dropped<DT, PK>(best, cur);
                if (dchg) Dc[idx] = best;
                if (__ballot(dchg) != 0) {
                    if (!noflags) {
                        for (int eb = e0; eb < e1; eb += kWave) {
                            const int e = eb + lane;
                            flag_neighbours(c, np, n, e < e1 ? adj[e].v : -1, Fnext, Qnext, cnext,
                                            qcap, list_open);
                        }
                    }
                    raised = true;
                }
            }
        }
    }
    if (raised && lane == 0) any[(r + 1) & 3] = 1;
}

// ------------------------------------------------------ block relaxation
// Graphs with locality (grids, meshes) are cut into blocks of at most kBV
// vertices grown breadth-first.  Slots s0 .. s0+cnt-1 of block b (s0 =
// boff[b]) hold its vertices bvert[s]; the adjacency of slot s lists the
// neighbours inside the block first, as local slots (badj[brp[s] ..
// bext[s])), then the others by vertex id (badj[bext[s] .. brp[s+1])).
// One workgroup takes one (chunk, block) item per visit:
//   1. loads the block's labels of the chunk's 64 rows into LDS and folds in
//      every neighbour outside the block (label + w, read once);
//   2. sweeps the block in LDS, forward and backward in turn; a sweep takes
//      the slots marked dirty (all of them in the first sweep), lowers each
//      label to the minimum over its block neighbours of (label + w), and
//      marks the block neighbours of every slot that dropped for the next
//      sweep -- until a sweep lowers nothing: the block is then at its fixed
//      point for the outside labels it read;
//   3. writes back the labels that dropped and, for a dropped vertex with
//      neighbours outside the block, flags those neighbours' blocks for the
//      next round.
// A round visits every flagged item once (round 0 every item); workgroup g
// takes items g, g + G, g + 2G, ... so a cluster of flagged blocks spreads
// over the grid.  Labels only decrease; a drop inside a block is settled
// within the visit and a drop on its border re-flags the blocks it touches,
// so the rounds end at the same fixed point as the vertex-wise rounds, after
// about (hop length of the longest path) / (block diameter) rounds instead of
// the hop length.  Each item is visited by one workgroup per round, which is
// the only writer of its labels.
constexpr int kBV = 248;  // vertices per block: 32-bit labels of a block fit 64 KB of LDS
constexpr int kBD = 4;    // adjacency entries per vertex and kind held in registers
constexpr int kBW = 16;   // waves per block workgroup (one workgroup fills a CU's LDS)
constexpr int kBT = kBW * kWave;
constexpr int kBPiece = 256;  // flags read per collection step
// Full sweeps of a block's first visit before its worklists take over
// (EMV_BLOCK_FULL_SWEEPS for the single-instance path).
constexpr int kBlkFullSweeps = 12;
constexpr int kBQ = (kBV + kBW - 1) / kBW;  // slots per wave of a block, at most
constexpr int kBRow = kBW * kBQ;             // entries per row of an in-block distance matrix

// LDS of a block workgroup: the labels, then 1 KB of words: 3 sweep flags,
// the piece counter, the piece list (256 B), 2 x kBV sweep marks, and from
// word kBMiscChg the kBW per-wave masks of lowered border slots.
constexpr int kBMiscChg = 200;
static_assert(8 + 64 + (2 * kBV + 3) / 4 <= kBMiscChg && kBMiscChg * 4 + 8 * kBW <= 1024,
              "block LDS words overlap");

template <typename DT>
__host__ __device__ constexpr size_t blk_lds_bytes() {
    return sizeof(DT) * kBV * kWave + 1024;
}

// Waves per SIMD a block workgroup's registers must leave room for: two
// workgroups of kBW waves per CU (8 per SIMD) while two fit the CU's LDS
// (labels up to 32 bits), else one.
template <typename DT>
constexpr int kBlkWaves() {
    return sizeof(DT) <= 4 ? 2 * kBW / 4 : kBW / 4;
}

// A 64-bit value the wave holds in every lane, made explicitly wave-uniform
// (scalar registers), so that masks built with cross-lane reductions keep the
// loops that walk them scalar.
__device__ __forceinline__ uint64_t uniform64(uint64_t x) {
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)x);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(x >> 32));
    return ((uint64_t)hi << 32) | lo;
}

// Lowest (or highest) set bit of a wave-uniform mask, as a scalar.
__device__ __forceinline__ int pick_bit(uint64_t m, bool lowest) {
    const int j = lowest ? __ffsll((unsigned long long)m) - 1 : 63 - __clzll((unsigned long long)m);
    return __builtin_amdgcn_readfirstlane(j);
}

// Puts the block neighbours of slot j (wave-local index, registers ic of its
// wave; the part of its list past the register copy, badj[ea..eb), one lane
// per entry) on the worklists: the slots of the calling wave's band (slots
// wid, wid + kBW, ...) into its mask, the others as LDS marks for their wave.
__device__ __forceinline__ void mark_block_neighbours(int j, const int (&ic)[kBD], int wid,
                                                      int ea, int eb,
                                                      const Edge *__restrict__ badj,
                                                      uint8_t *marks, uint64_t &wmask,
                                                      bool &cross) {
    const int lane = __lane_id();
#pragma unroll
    for (int q = 0; q < kBD; q++) {
        const int sl = __builtin_amdgcn_readlane(ic[q], j);
        if (sl >= 0) {
            if (sl % kBW == wid) {
                wmask |= 1ull << (sl / kBW);
            } else {
                if (lane == 0) marks[sl] = 1;
                cross = true;
            }
        }
    }
    if (ea < eb) {
        uint64_t add = 0;
        bool xo = false;
        for (int e = ea + lane; e < eb; e += kWave) {
            const int sl = badj[e].v;
            if (sl % kBW == wid) {
                add |= 1ull << (sl / kBW);
            } else {
                marks[sl] = 1;
                xo = true;
            }
        }
        for (int off = 32; off > 0; off >>= 1) add |= __shfl_xor(add, off);
        wmask = uniform64(wmask | add);
        cross = cross || (__ballot(xo) != 0);
    }
}

// min(x, lb + d) for a label lb and an in-block distance d.  lb is INF in
// the lanes (rows) where the border slot holds no path yet, and the sum must
// then lower nothing: 16-bit labels add in 32 bits (INF + d >= INF), 32-bit
// ones saturate to INF, and a 64-bit sum that wraps (only INF + d can:
// finite labels stay far below 2^63) is dropped.
template <typename DT, typename DW>
__device__ __forceinline__ DT closure_min(DT x, DT lb, DW d) {
    if constexpr (sizeof(DT) == 2) {
        const uint32_t c = (uint32_t)lb + (uint32_t)d;
        return c < (uint32_t)x ? (DT)c : x;
    } else if constexpr (sizeof(DT) == 4) {
        const uint32_t c = __builtin_elementwise_add_sat((uint32_t)lb, (uint32_t)d);
        return c < x ? c : x;
    } else {
        const uint64_t c = lb + (uint64_t)d;
        return (c >= lb && c < x) ? c : x;
    }
}

// A wave's kBQ entries of a distance row, read as whole dwords (one scalar
// load of 32 or 64 bytes: the address is wave-uniform and the row padded).
template <typename DW>
__device__ __forceinline__ void band_row(const DW *__restrict__ p, DW (&dv)[kBQ]) {
    // the matrices are read-only while any kernel runs: constant address space
    typedef __attribute__((address_space(4))) const uint32_t *const_u32p;
    const const_u32p q = (const_u32p)(p);
    if constexpr (sizeof(DW) == 2) {
#pragma unroll
        for (int t = 0; t < kBQ / 2; t++) {
            const uint32_t w = q[t];
            dv[2 * t] = (DW)(w & 0xFFFFu);
            dv[2 * t + 1] = (DW)(w >> 16);
        }
    } else {
#pragma unroll
        for (int t = 0; t < kBQ; t++) dv[t] = q[t];
    }
}

// One visit of a (chunk, block) item: Dc the chunk's labels, s0 .. s0+cnt-1
// the block's slots, vblk the block of each vertex of the graph and Fc the
// chunk's block flags of the next round.  All threads of the workgroup take
// part; raised is set when a block beyond this one was flagged.
//
// Db (optional) is the block's matrix of in-block distances from each border
// slot (the first slots of a block) to every slot: row b holds the entry of
// slot j * kBW + w (wave w's band) at kBQ * w + j (kBRow entries a row).  With it, a visit of a
// later round -- the block sits at its in-block fixed point since its last
// visit, only border labels can drop when the outside is folded in -- takes
// every slot to min over the lowered border slots b of (label of b + Db[b][
// slot]) in one pass, instead of sweeping: that is the in-block fixed point
// again, since in-block distances obey the triangle inequality over every
// in-block edge.
template <typename DT, typename DW>
__device__ __forceinline__ void blk_visit(DT *__restrict__ Dc, int s0, int cnt,
                                          const int32_t *__restrict__ bvert,
                                          const int32_t *__restrict__ brp,
                                          const int32_t *__restrict__ bext,
                                          const Edge *__restrict__ badj,
                                          const int32_t *__restrict__ vblk,
                                          const DW *__restrict__ Db,
                                          uint32_t *__restrict__ Fc, int dense, DT *L,
                                          uint32_t *s_flag, uint8_t *s_dirty, uint64_t *s_chg,
                                          unsigned long long *__restrict__ stats, bool &raised,
                                          int fsw = kBlkFullSweeps) {
    const int tid = threadIdx.x;
    const int lane = tid & (kWave - 1);
    const int wid = __builtin_amdgcn_readfirstlane(tid / kWave);
    const DT INF = (DT)~(DT)0;
    const long long t_start = stats ? clock64() : 0;
    // wave w owns the slots w, w + kBW, w + 2 kBW, ... (its band, strided so
    // that the border slots, which come first, spread over all waves); lane
    // j holds slot j * kBW + w
    const int nv = wid < cnt ? (cnt - 1 - wid) / kBW + 1 : 0;
    int vj = 0, e0 = 0, em = 0, e1 = 0, ovf = 0;
    int ic[kBD], iw[kBD], xu[kBD], xw[kBD], xb[kBD];
#pragma unroll
    for (int q = 0; q < kBD; q++) {
        ic[q] = -1;
        iw[q] = 0;
        xu[q] = -1;
        xw[q] = 0;
        xb[q] = -1;
    }
    if (lane < nv) {
        const int s = s0 + lane * kBW + wid;
        vj = bvert[s];
        e0 = brp[s];
        em = bext[s];
        e1 = brp[s + 1];
        ovf = (em - e0 > kBD || e1 - em > kBD) ? 1 : 0;
#pragma unroll
        for (int q = 0; q < kBD; q++) {
            if (e0 + q < em) {
                const Edge ed = badj[e0 + q];
                ic[q] = ed.v;
                iw[q] = ed.w;
            }
            if (em + q < e1) {
                const Edge ed = badj[em + q];
                xu[q] = ed.v;
                xw[q] = ed.w;
            }
        }
#pragma unroll
        for (int q = 0; q < kBD; q++)
            if (xu[q] >= 0) xb[q] = vblk[xu[q]];
    }
    const bool any_ovf = __ballot(ovf != 0) != 0;
    if (tid < 3) s_flag[tid] = 0;
    const uint64_t own_mask = nv >= 64 ? ~0ull : ((1ull << nv) - 1ull);
    // 1a. the block's labels, all of the wave's band (nv <= kBQ slots) in
    //     one step, so that its loads are in flight together
    if (nv > 0) {
        DT own[kBQ];
#pragma unroll
        for (int t = 0; t < kBQ; t++) {
            const int j = min(t, nv - 1);
            own[t] = Dc[(int64_t)__builtin_amdgcn_readlane(vj, j) * kWave + lane];
        }
#pragma unroll
        for (int t = 0; t < kBQ; t++)
            if (t < nv) L[(t * kBW + wid) * kWave + lane] = own[t];
    }
    // 1b. fold in the outside neighbours of the border slots, four per
    //     step; a slot that drops puts its block neighbours on the
    //     worklists of the first sweep
    uint64_t chg_mask = 0;  // slots whose label dropped during this visit
    uint64_t wmask = dense ? own_mask : 0ull;
    bool cross0 = false;
    {
        uint64_t bmask = uniform64(__ballot(lane < nv && e1 > em));
        while (bmask) {
            int jj[4];
            int nt = 0;
#pragma unroll
            for (int t = 0; t < 4; t++) {
                if (bmask) {
                    const int j = pick_bit(bmask, true);
                    bmask &= bmask - 1;
                    jj[t] = j;
                    nt++;
                } else {
                    jj[t] = jj[0];
                }
            }
            DT ext[4][kBD];
#pragma unroll
            for (int t = 0; t < 4; t++)
#pragma unroll
                for (int q = 0; q < kBD; q++) {
                    const int u = __builtin_amdgcn_readlane(xu[q], jj[t]);
                    ext[t][q] = u >= 0 ? Dc[(int64_t)u * kWave + lane] : INF;
                }
#pragma unroll
            for (int t = 0; t < 4; t++) {
                if (t >= nt) break;
                const int j = jj[t];
                const DT x0 = L[(j * kBW + wid) * kWave + lane];
                DT x = x0;
#pragma unroll
                for (int q = 0; q < kBD; q++)
                    x = relax_min(ext[t][q], __builtin_amdgcn_readlane(xu[q], j),
                                  __builtin_amdgcn_readlane(xw[q], j), INF, x);
                if (any_ovf && __builtin_amdgcn_readlane(ovf, j)) {
                    const int ea = __builtin_amdgcn_readlane(em, j) + kBD;
                    const int eb = __builtin_amdgcn_readlane(e1, j);
                    for (int e = ea; e < eb; e++) {
                        const Edge ed = badj[e];
                        x = relax_min(Dc[(int64_t)ed.v * kWave + lane], ed.v, ed.w, INF, x);
                    }
                }
                const bool lower = x < x0;
                if (lower) L[(j * kBW + wid) * kWave + lane] = x;
                if (__ballot(lower) != 0) {
                    chg_mask |= 1ull << j;
                    if (!dense && Db == nullptr) {
                        const bool ov = any_ovf && __builtin_amdgcn_readlane(ovf, j);
                        mark_block_neighbours(
                            j, ic, wid, ov ? __builtin_amdgcn_readlane(e0, j) + kBD : 0,
                            ov ? __builtin_amdgcn_readlane(em, j) : 0, badj, s_dirty, wmask,
                            cross0);
                    }
                }
            }
        }
    }
    const bool closure = !dense && Db != nullptr;
    if (closure && lane == 0) s_chg[wid] = chg_mask;
    __syncthreads();
    const long long t_loaded = stats ? clock64() : 0;
    if (closure) {
        // 2'. closure from the border slots the fold lowered (each wave's
        //     in s_chg), one distance row per lowered slot, scalar loads of
        //     the wave's band of it
        DT x[kBQ], x0[kBQ];
#pragma unroll
        for (int t = 0; t < kBQ; t++) x0[t] = x[t] = t < nv ? L[(t * kBW + wid) * kWave + lane] : INF;
        // lowered slots NB at a time, taken across the waves' masks by a
        // scalar cursor (w2, m) over s_chg (a short last step repeats its
        // first slot: min is idempotent), so that the row loads of NB slots
        // are in flight together; entries past the band (t >= nv) are
        // padding and their results are dropped below
        constexpr int NB = sizeof(DW) == 2 ? 4 : 2;
        int w2 = 0;
        uint64_t m = uniform64(s_chg[0]);
        auto next_low = [&]() -> int {
            while (m == 0 && w2 < kBW - 1) m = uniform64(s_chg[++w2]);
            if (m == 0) return -1;
            const int b = pick_bit(m, true) * kBW + w2;
            m &= m - 1;
            return b;
        };
        for (int b0 = next_low(); b0 >= 0; b0 = next_low()) {
            int bb[NB];
            bb[0] = b0;
#pragma unroll
            for (int t = 1; t < NB; t++) {
                const int bt = next_low();
                bb[t] = bt >= 0 ? bt : b0;
            }
            DT lv[NB];
            DW dv[NB][kBQ];
#pragma unroll
            for (int t = 0; t < NB; t++) {
                lv[t] = L[bb[t] * kWave + lane];
                band_row<DW>(Db + (int64_t)bb[t] * kBRow + wid * kBQ, dv[t]);
            }
#pragma unroll
            for (int q = 0; q < kBQ; q++) {
                DT v = x[q];
#pragma unroll
                for (int t = 0; t < NB; t++) v = closure_min<DT, DW>(v, lv[t], dv[t][q]);
                x[q] = v;
            }
        }
#pragma unroll
        for (int t = 0; t < kBQ; t++) {
            const bool lower = t < nv && x[t] < x0[t];
            if (lower) L[(t * kBW + wid) * kWave + lane] = x[t];
            if (__ballot(lower) != 0) chg_mask |= 1ull << t;
        }
        if (stats && tid == 0) {
            // visits, visits whose fold lowered nothing, lowered border slots
            unsigned long long low = 0;
            for (int w2 = 0; w2 < kBW; w2++) low += __popcll(s_chg[w2]);
            atomicAdd(&stats[0], 1ull);
            if (low == 0) atomicAdd(&stats[6], 1ull);
            atomicAdd(&stats[7], low);
        }
    } else {
        // 2. sweeps.  A block's first visit of a relaxation (dense) starts
        //    with full sweeps: every wave relaxes all its slots over their
        //    block neighbours, in place, a barrier per sweep; odd sweeps take
        //    a wave's slots in reverse order, so that a drop runs against the
        //    slot order (breadth-first from the block's seed) as fast as along
        //    it (16-bit labels with 32-bit distances keep one order and full
        //    sweeps to the end: there the second loop spilled registers to
        //    scratch).  After fsw full sweeps, the last of which marks the
        //    block neighbours of the slots it lowered, and on every later
        //    visit of a round without in-block distances, each wave drains a
        //    worklist of its own slots: a slot whose label drops puts its
        //    block neighbours on the list -- its own wave's at once, another
        //    wave's through the LDS marks, which that wave takes up after the
        //    barrier.  The visit ends when a full sweep lowers nothing or a
        //    worklist sweep leaves no mark for another wave.
        constexpr bool one_order = sizeof(DT) == 2 && sizeof(DW) == 4;
        const int K = dense ? (one_order ? (1 << 30) : fsw) : 0;
        unsigned long long processed = 0;
        for (int it = 0;; it++) {
            uint8_t *dcur = s_dirty + (it & 1) * kBV;
            uint8_t *dnext = s_dirty + ((it + 1) & 1) * kBV;
            if (tid == 0) s_flag[(it + 1) % 3] = 0;
            bool cross = false, flag;
            if (it < K) {
                const bool last = it == K - 1;
                if (last) wmask = 0;
                bool any_low = false;
                if constexpr (one_order) {
#pragma unroll
                    for (int t = 0; t < kBQ; t++) {
                        if (t >= nv) break;
                        const int sl0 = t * kBW + wid;
                        const DT x0 = L[sl0 * kWave + lane];
                        DT nb[kBD];
#pragma unroll
                        for (int q = 0; q < kBD; q++) {
                            const int sl = __builtin_amdgcn_readlane(ic[q], t);
                            nb[q] = sl >= 0 ? L[sl * kWave + lane] : INF;
                        }
                        DT best = x0;
#pragma unroll
                        for (int q = 0; q < kBD; q++)
                            best = relax_min(nb[q], __builtin_amdgcn_readlane(ic[q], t),
                                             __builtin_amdgcn_readlane(iw[q], t), INF, best);
                        if (any_ovf && __builtin_amdgcn_readlane(ovf, t)) {
                            const int ea = __builtin_amdgcn_readlane(e0, t) + kBD;
                            const int eb = __builtin_amdgcn_readlane(em, t);
                            for (int e = ea; e < eb; e++) {
                                const Edge ed = badj[e];
                                best = relax_min(L[ed.v * kWave + lane], ed.v, ed.w, INF, best);
                            }
                        }
                        const bool lower = best < x0;
                        if (lower) L[sl0 * kWave + lane] = best;
                        if (__ballot(lower) != 0) {
                            chg_mask |= 1ull << t;
                            any_low = true;
                        }
                    }
                } else {
                    auto relax_slot = [&](int t) {
                        const int sl0 = t * kBW + wid;
                        const DT x0 = L[sl0 * kWave + lane];
                        DT nb[kBD];
#pragma unroll
                        for (int q = 0; q < kBD; q++) {
                            const int sl = __builtin_amdgcn_readlane(ic[q], t);
                            nb[q] = sl >= 0 ? L[sl * kWave + lane] : INF;
                        }
                        DT best = x0;
#pragma unroll
                        for (int q = 0; q < kBD; q++)
                            best = relax_min(nb[q], __builtin_amdgcn_readlane(ic[q], t),
                                             __builtin_amdgcn_readlane(iw[q], t), INF, best);
                        if (any_ovf && __builtin_amdgcn_readlane(ovf, t)) {
                            const int ea = __builtin_amdgcn_readlane(e0, t) + kBD;
                            const int eb = __builtin_amdgcn_readlane(em, t);
                            for (int e = ea; e < eb; e++) {
                                const Edge ed = badj[e];
                                best = relax_min(L[ed.v * kWave + lane], ed.v, ed.w, INF, best);
                            }
                        }
                        const bool lower = best < x0;
                        if (lower) L[sl0 * kWave + lane] = best;
                        if (__ballot(lower) != 0) {
                            chg_mask |= 1ull << t;
                            any_low = true;
                            if (last) {
                                const bool ov = any_ovf && __builtin_amdgcn_readlane(ovf, t);
                                mark_block_neighbours(t, ic, wid, ov ? __builtin_amdgcn_readlane(e0, t) + kBD : 0,
                                                      ov ? __builtin_amdgcn_readlane(em, t) : 0, badj, dnext,
                                                      wmask, cross);
                            }
                        }
                    };
                    if (it & 1) {
#pragma unroll
                        for (int t = kBQ - 1; t >= 0; t--)
                            if (t < nv) relax_slot(t);
                    } else {
#pragma unroll
                        for (int t = 0; t < kBQ; t++) {
                            if (t >= nv) break;
                            relax_slot(t);
                        }
                    }
                }
                flag = any_low;
            } else {
                {
                    const bool dl = lane < nv && dcur[lane * kBW + wid] != 0;
                    if (dl) dcur[lane * kBW + wid] = 0;
                    wmask = uniform64(wmask | __ballot(dl));
                }
                bool cross = false;
                while (wmask) {
                    int jj[4];
                    int nt = 0;
#pragma unroll
                    for (int t = 0; t < 4; t++) {
                        if (wmask) {
                            const int j = pick_bit(wmask, true);
                            wmask &= wmask - 1;
                            jj[t] = j;
                            nt++;
                        } else {
                            jj[t] = jj[0];
                        }
                    }
                    processed += nt;
                    DT x[4], nb[4][kBD];
#pragma unroll
                    for (int t = 0; t < 4; t++) {
                        x[t] = L[(jj[t] * kBW + wid) * kWave + lane];
#pragma unroll
                        for (int q = 0; q < kBD; q++) {
                            const int sl = __builtin_amdgcn_readlane(ic[q], jj[t]);
                            nb[t][q] = sl >= 0 ? L[sl * kWave + lane] : INF;
                        }
                    }
#pragma unroll
                    for (int t = 0; t < 4; t++) {
                        if (t >= nt) break;
                        const int j = jj[t];
                        DT best = x[t];
#pragma unroll
                        for (int q = 0; q < kBD; q++)
                            best = relax_min(nb[t][q], __builtin_amdgcn_readlane(ic[q], j),
                                             __builtin_amdgcn_readlane(iw[q], j), INF, best);
                        const bool ov = any_ovf && __builtin_amdgcn_readlane(ovf, j);
                        const int ea = ov ? __builtin_amdgcn_readlane(e0, j) + kBD : 0;
                        const int eb = ov ? __builtin_amdgcn_readlane(em, j) : 0;
                        for (int e = ea; e < eb; e++) {
                            const Edge ed = badj[e];
                            best = relax_min(L[ed.v * kWave + lane], ed.v, ed.w, INF, best);
                        }
                        const bool lower = best < x[t];
                        if (lower) L[(j * kBW + wid) * kWave + lane] = best;
                        if (__ballot(lower) != 0) {
                            chg_mask |= 1ull << j;
                            mark_block_neighbours(j, ic, wid, ea, eb, badj, dnext, wmask, cross);
                        }
                    }
                }
                flag = cross;
            }
            if (flag && lane == 0) s_flag[it % 3] = 1;
            __syncthreads();
            if (s_flag[it % 3] == 0) {
                if (stats && lane == 0) {
                    if (tid == 0) {
                        atomicAdd(&stats[0], 1ull);
                        atomicAdd(&stats[1], (unsigned long long)(it + 1));
                    }
                    atomicAdd(&stats[2], processed);
                }
                break;
            }
        }
    }
    const long long t_swept = stats ? clock64() : 0;
    // 3. write back the slots that dropped, eight per step; flag the
    //    blocks beyond a dropped border slot (lane q takes its outside
    //    neighbour q, the rest of the list one lane each)
    chg_mask = uniform64(chg_mask);
    while (chg_mask) {
        int jj[8];
        int nt = 0;
#pragma unroll
        for (int t = 0; t < 8; t++) {
            if (chg_mask) {
                const int j = pick_bit(chg_mask, true);
                chg_mask &= chg_mask - 1;
                jj[t] = j;
                nt++;
            } else {
                jj[t] = jj[0];
            }
        }
        DT x[8];
#pragma unroll
        for (int t = 0; t < 8; t++) x[t] = L[(jj[t] * kBW + wid) * kWave + lane];
#pragma unroll
        for (int t = 0; t < 8; t++) {
            if (t >= nt) break;
            const int j = jj[t];
            Dc[(int64_t)__builtin_amdgcn_readlane(vj, j) * kWave + lane] = x[t];
            const int ea = __builtin_amdgcn_readlane(em, j);
            const int eb = __builtin_amdgcn_readlane(e1, j);
            if (ea < eb) {
                int fb = -1;
#pragma unroll
                for (int q = 0; q < kBD; q++) {
                    const int bq = __builtin_amdgcn_readlane(xb[q], j);
                    if (lane == q) fb = bq;
                }
                if (fb >= 0 && Fc[fb] == 0) Fc[fb] = 1;
                for (int e = ea + kBD + lane; e < eb; e += kWave) {
                    const int ob = vblk[badj[e].v];
                    if (Fc[ob] == 0) Fc[ob] = 1;
                }
                raised = true;
            }
        }
    }
    if (stats && tid == 0) {
        const long long t_end = clock64();
        atomicAdd(&stats[3], (unsigned long long)(t_loaded - t_start));
        atomicAdd(&stats[4], (unsigned long long)(t_swept - t_loaded));
        atomicAdd(&stats[5], (unsigned long long)(t_end - t_swept));
    }
}

// Dm / bdoff (optional, Dm null: sweeps only): the in-block distance
// matrices, block b's at Dm + bdoff[b].
template <typename DT, typename DW>
__global__ __launch_bounds__(kBT) __attribute__((amdgpu_waves_per_eu(kBlkWaves<DT>()))) void k_relax_blk(
    DT *__restrict__ D, int64_t np, int nblk, const int32_t *__restrict__ boff,
    const int32_t *__restrict__ bvert, const int32_t *__restrict__ brp,
    const int32_t *__restrict__ bext, const Edge *__restrict__ badj,
    const int32_t *__restrict__ vblk, const DW *__restrict__ Dm,
    const int64_t *__restrict__ bdoff, int64_t nitems, int dense, uint32_t *__restrict__ Fcur,
    uint32_t *__restrict__ Fnext, uint32_t *__restrict__ any, int r,
    unsigned long long *__restrict__ stats, int fsw) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    DT *L = reinterpret_cast<DT *>(lds_raw);  // L[slot * 64 + row]
    uint32_t *s_misc = reinterpret_cast<uint32_t *>(lds_raw + sizeof(DT) * kBV * kWave);
    uint32_t *s_flag = s_misc;     // 3 rotating "sweep lowered something" words
    uint32_t *s_cnt = s_misc + 4;  // flagged items of the current piece
    uint8_t *s_list = reinterpret_cast<uint8_t *>(s_misc + 8);      // 256 entries
    uint8_t *s_dirty = reinterpret_cast<uint8_t *>(s_misc + 8 + 64);  // 2 x kBV marks
    uint64_t *s_chg = reinterpret_cast<uint64_t *>(s_misc + kBMiscChg);  // kBW masks
    const int tid = threadIdx.x;
    const int lane = tid & (kWave - 1);
    if (blockIdx.x == 0 && tid == 0) any[(r + 2) & 3] = 0;
    if (!dense && any[r & 3] == 0) return;
    const int64_t G = gridDim.x;
    const int64_t mine = (nitems > (int64_t)blockIdx.x) ? (nitems - 1 - blockIdx.x) / G + 1 : 0;
    // the marks start clear and are clear again whenever a visit's sweeps end
    for (int i = tid; i < 2 * kBV; i += kBT) s_dirty[i] = 0;
    bool raised = false;
    for (int64_t p0 = 0; p0 < mine; p0 += kBPiece) {
        if (tid == 0) *s_cnt = 0;
        __syncthreads();
        if (tid < kBPiece && p0 + tid < mine) {
            const int64_t it = blockIdx.x + G * (p0 + tid);
            bool f = dense != 0;
            if (!dense && Fcur[it] != 0) {
                f = true;
                Fcur[it] = 0;
            }
            if (f) s_list[atomicAdd(s_cnt, 1u)] = (uint8_t)tid;
        }
        __syncthreads();
        const int nl = (int)*s_cnt;
        for (int qi = 0; qi < nl; qi++) {
            const int64_t item = blockIdx.x + G * (p0 + s_list[qi]);
            const int c = (int)(item / nblk);
            const int b = (int)(item - (int64_t)c * nblk);
            blk_visit<DT, DW>(D + (int64_t)c * np * kWave, boff[b], boff[b + 1] - boff[b], bvert,
                              brp, bext, badj, vblk, Dm ? Dm + bdoff[b] : nullptr,
                              Fnext + (int64_t)c * nblk, dense, L, s_flag, s_dirty, s_chg, stats,
                              raised, fsw);
            __syncthreads();  // LDS is reused by the next item
        }
        __syncthreads();  // s_cnt is reset by the next piece
    }
    if (raised && lane == 0) any[(r + 1) & 3] = 1;
}

// ------------------------------------------------------------------ merge
// f[X][v] = min over nonempty Y strictly inside X of f[Y][v] + f[X\Y][v]
// (reader-el.c:2068-2088).  Only Y without the top bit of X is enumerated:
// each pair {Y, X\Y} is visited once and Y is the smaller of the two.  The
// accumulator packs (sum << k) | Y, so an unsigned min keeps the smallest
// sum and, among equal sums, the smallest Y -- the reference's first strict
// minimum over increasing Y (strict '<' at 2078).
// The subset of `mask` whose bits, read from the lowest, spell x.
__device__ __forceinline__ uint32_t deposit_bits(uint32_t x, uint32_t mask) {
    uint32_t r = 0;
    for (uint32_t bb = 1; mask != 0; bb <<= 1) {
        const uint32_t low = mask & (~mask + 1u);
        if (x & bb) r |= low;
        mask ^= low;
    }
    return r;
}

// The eight subsets of a three-bit mask L, yl[j] = the bits of L that j
// selects (bit i of j picks the i-th lowest bit of L): deposit_bits(j, L)
// without its loop.
__device__ __forceinline__ void low_subsets(uint32_t L, uint32_t yl[8]) {
    const uint32_t b0 = L & (~L + 1u);
    const uint32_t r = L ^ b0;
    const uint32_t b1 = r & (~r + 1u);
    const uint32_t b2 = r ^ b1;
#pragma unroll
    for (int j = 0; j < 8; j++) yl[j] = ((j & 1) ? b0 : 0u) | ((j & 2) ? b1 : 0u) | ((j & 4) ? b2 : 0u);
}

// Pairs of X, enumerated for cheap addressing.  Y runs over the nonempty
// subsets of Xp = X without its top bit (each unordered pair once, Y the side
// without the top bit, which is the smaller one).  Row offsets of disjoint
// sets add up, off(Y | Z) = off(Y) + off(Z), so with Xp split into L (its
// lowest set bits, at most three) and H = Xp \ L, Y = Yh | Yl is walked with
// Yh over the subsets of H ("outer" steps, highest first: Yh = deposit(nH - 1
// - o, H)) and Yl over the subsets of L unrolled: one 64-bit multiply serves
// 2^|L| pairs, and the partner row X \ Y sits at off(X) - off(Y).  The packed
// minimum does not depend on the order, so any slice [o0, o1) of the outer
// steps can be folded separately.
template <typename T, typename P>
__device__ __forceinline__ P merge_outer(const T *__restrict__ fv, int64_t np, uint32_t X, int k,
                                         uint32_t o0, uint32_t o1) {
    const uint32_t Xp = X ^ (1u << (31 - __clz(X)));
    uint32_t L = 0, H = Xp;
    for (int i = 0; i < 3 && H; i++) {
        const uint32_t lo = H & (~H + 1u);
        L |= lo;
        H ^= lo;
    }
    const int nl = __popc(L);
    const uint32_t nH = 1u << __popc(H);
    const int64_t oX = (int64_t)X * np;
    P acc = ~(P)0;
    uint32_t Yh = deposit_bits(nH - 1u - o0, H);
    if (nl == 3) {
        uint32_t yl[8];
        int64_t ol[8];
        low_subsets(L, yl);
#pragma unroll
        for (int j = 0; j < 8; j++) ol[j] = (int64_t)yl[j] * np;
        const uint32_t oend = o1 < nH - 1u ? o1 : nH - 1u;  // Yh == 0 last, apart
        for (uint32_t o = o0; o < oend; o++) {
            const int64_t oh = (int64_t)Yh * np;
            const T *pa = fv + oh;
            const T *pc = fv + (oX - oh);
            T a[8], c[8];
#pragma unroll
            for (int j = 0; j < 8; j++) {
                a[j] = pa[ol[j]];
                c[j] = pc[-ol[j]];
            }
#pragma unroll
            for (int j = 0; j < 8; j++) {
                const P s = (((P)a[j] + (P)c[j]) << k) | (P)(Yh | yl[j]);
                acc = s < acc ? s : acc;
            }
            Yh = (Yh - 1u) & H;
        }
        if (o1 == nH) {  // Yh == 0: the seven nonempty subsets of L
            const T *pc = fv + oX;
            T a[8], c[8];
#pragma unroll
            for (int j = 1; j < 8; j++) {
                a[j] = fv[ol[j]];
                c[j] = pc[-ol[j]];
            }
#pragma unroll
            for (int j = 1; j < 8; j++) {
                const P s = (((P)a[j] + (P)c[j]) << k) | (P)yl[j];
                acc = s < acc ? s : acc;
            }
        }
        return acc;
    }
    // |Xp| <= 2 (layers 2 and 3): H is empty, one outer step over L
    for (uint32_t Y = L; Y != 0; Y = (Y - 1u) & L) {
        const P s = (((P)fv[(int64_t)Y * np] + (P)fv[(int64_t)(X ^ Y) * np]) << k) | (P)Y;
        acc = s < acc ? s : acc;
    }
    return acc;
}

// Rows with at least this many pairs (the upper layers) take the split merge.
constexpr uint32_t kSplitPairs = 128;

// Number of outer steps of X in merge_outer.
__device__ __forceinline__ uint32_t merge_steps(uint32_t X) {
    const uint32_t Xp = X ^ (1u << (31 - __clz(X)));
    const int r = __popc(Xp);
    return r <= 3 ? 1u : (1u << (r - 3));
}

// All pairs of X.
template <typename T, typename P>
__device__ __forceinline__ P merge_one(const T *__restrict__ fv, int64_t np, uint32_t X, int k) {
    return merge_outer<T, P>(fv, np, X, k, 0u, merge_steps(X));
}

// merge_outer over a compact copy of an int64 table: row Y holds int32
// cells f[Y][v] - fbase[Y] (the row's base, int64).  A pair's sum adds the
// two rows' bases back, so the packed minimum is the one over the int64
// cells, at half their bytes.
template <typename P>
__device__ __forceinline__ P merge_outer_rel(const int32_t *__restrict__ fv,
                                             const int64_t *__restrict__ fbase, int64_t np,
                                             uint32_t X, int k, uint32_t o0, uint32_t o1) {
    const uint32_t Xp = X ^ (1u << (31 - __clz(X)));
    uint32_t L = 0, H = Xp;
    for (int i = 0; i < 3 && H; i++) {
        const uint32_t lo = H & (~H + 1u);
        L |= lo;
        H ^= lo;
    }
    const int nl = __popc(L);
    const uint32_t nH = 1u << __popc(H);
    const int64_t oX = (int64_t)X * np;
    P acc = ~(P)0;
    if (nl == 3) {
        uint32_t yl[8];
        int64_t ol[8];
        low_subsets(L, yl);
#pragma unroll
        for (int j = 0; j < 8; j++) ol[j] = (int64_t)yl[j] * np;
        uint32_t Yh = deposit_bits(nH - 1u - o0, H);
        const uint32_t oend = o1 < nH - 1u ? o1 : nH - 1u;  // Yh == 0 last, apart
        for (uint32_t o = o0; o < oend; o++) {
            const int64_t oh = (int64_t)Yh * np;
            const int32_t *pa = fv + oh;
            const int32_t *pc = fv + (oX - oh);
            int32_t a[8], c[8];
            int64_t bs[8];
#pragma unroll
            for (int j = 0; j < 8; j++) {
                a[j] = pa[ol[j]];
                c[j] = pc[-ol[j]];
                const uint32_t Y = Yh | yl[j];
                bs[j] = fbase[Y] + fbase[X ^ Y];
            }
#pragma unroll
            for (int j = 0; j < 8; j++) {
                const P sum = (P)(uint32_t)a[j] + (P)(uint32_t)c[j] + (P)bs[j];
                const P s = (sum << k) | (P)(Yh | yl[j]);
                acc = s < acc ? s : acc;
            }
            Yh = (Yh - 1u) & H;
        }
        if (o1 == nH) {  // Yh == 0: the seven nonempty subsets of L
            const int32_t *pc = fv + oX;
#pragma unroll
            for (int j = 1; j < 8; j++) {
                const P sum = (P)(uint32_t)fv[ol[j]] + (P)(uint32_t)pc[-ol[j]] +
                              (P)(fbase[yl[j]] + fbase[X ^ yl[j]]);
                const P s = (sum << k) | (P)yl[j];
                acc = s < acc ? s : acc;
            }
        }
        return acc;
    }
    for (uint32_t Y = L; Y != 0; Y = (Y - 1u) & L) {
        const P sum = (P)(uint32_t)fv[(int64_t)Y * np] + (P)(uint32_t)fv[(int64_t)(X ^ Y) * np] +
                      (P)(fbase[Y] + fbase[X ^ Y]);
        const P s = (sum << k) | (P)Y;
        acc = s < acc ? s : acc;
    }
    return acc;
}

// All pairs of X from the table (fbase null) or from its compact copy.
template <typename T, typename P>
__device__ __forceinline__ P merge_any(const T *__restrict__ fv, const int64_t *__restrict__ fbase,
                                       int64_t np, uint32_t X, int k, uint32_t o0, uint32_t o1) {
    if constexpr (std::is_same<T, int32_t>::value) {
        if (fbase) return merge_outer_rel<P>(fv, fbase, np, X, k, o0, o1);
    }
    return merge_outer<T, P>(fv, np, X, k, o0, o1);
}

// Work item = one wave: 64 consecutive vertices (one per lane, 256 B
// coalesced row loads) x XG consecutive rows of the layer batch.  Items are
// tile-major and dealt to XCDs in contiguous ranges, so the waves of an XCD
// sweep the rows of one vertex tile together and the tile's lower-layer rows
// stay in that XCD's L2.  Results go straight to the vertex-major staging:
// each lane writes its XG labels (and packed words) as one contiguous run.
// Rows past nx (the tail of the last chunk) are written as INF.
template <typename T, typename P, typename DT, int XG, int CW = kWave>
__global__ __launch_bounds__(kBlock) void k_merge_v(const T *__restrict__ f, FLayout fl, int64_t np,
                                                    const uint32_t *__restrict__ xs, int nx,
                                                    int nsub, int64_t ntiles, int k,
                                                    DT *__restrict__ D, P *__restrict__ CODE,
                                                    int nxcd, int64_t per_xcd,
                                                    const int64_t *__restrict__ fbase) {
    const int lane = threadIdx.x & (kWave - 1);
    const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    if ((int64_t)(blockIdx.x / nxcd) >= per_xcd) return;
    const int64_t item = xcd_item(nxcd, per_xcd) * (kBlock / kWave) + wid;
    const int64_t tile = item / nsub;
    const int sub = (int)(item - tile * nsub);
    if (tile >= ntiles) return;
    const int64_t v = tile * kWave + lane;
    const T *fv = f + fl.cell(0, v);
    const DT INF = (DT)~(DT)0;
    DT dv[XG];
    P pv[XG];
#pragma unroll
    for (int jx = 0; jx < XG; jx++) {
        const int xi = sub * XG + jx;
        dv[jx] = INF;
        pv[jx] = ~(P)0;
        if (xi < nx) {
            const uint32_t X = xs[xi];
            pv[jx] = merge_any<T, P>(fv, fbase, fl.rcells, X, k, 0u, merge_steps(X));
            dv[jx] = (DT)(pv[jx] >> k);
        }
    }
    // the XG rows of the item lie in one chunk of CW rows
    const int chunk = sub / (CW / XG);
    const int loff = (sub % (CW / XG)) * XG;
    DT *dd = D + ((int64_t)chunk * np + v) * CW + loff;
    P *cd = CODE + ((int64_t)chunk * np + v) * CW + loff;
#pragma unroll
    for (int jx = 0; jx < XG; jx++) {
        dd[jx] = dv[jx];
        cd[jx] = pv[jx];
    }
}

// Split merge for the upper layers (few rows, many pairs per row): the outer
// steps of row xi are cut into S parts, and item = (tile, xi, part) is
// numbered tile-major and dealt to XCDs in contiguous ranges, so the waves
// of an XCD work through one 64-vertex tile at a time and read its
// lower-layer rows from that XCD's L2.  Each item folds its part into the
// packed accumulator M[xi * np + v] with an unsigned atomicMin; the minimum
// does not depend on the order, so the split leaves every result unchanged.
// M starts at all ones; k_pack_v moves it to the staging.
template <typename T, typename P>
__global__ __launch_bounds__(kBlock) void k_merge_split(const T *__restrict__ f, FLayout fl, int64_t np,
                                                        const uint32_t *__restrict__ xs, int nx,
                                                        int S, int64_t ntiles, int k,
                                                        P *__restrict__ M, int nxcd,
                                                        int64_t per_xcd,
                                                        const int64_t *__restrict__ fbase) {
    const int lane = threadIdx.x & (kWave - 1);
    const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    if ((int64_t)(blockIdx.x / nxcd) >= per_xcd) return;
    const int64_t item = xcd_item(nxcd, per_xcd) * (kBlock / kWave) + wid;
    const int64_t per_tile = (int64_t)nx * S;
    const int64_t tile = item / per_tile;
    if (tile >= ntiles) return;
    const int rem = (int)(item - tile * per_tile);
    const int xi = rem / S;
    const int part = rem - xi * S;
    const uint32_t X = xs[xi];
    const uint32_t steps = merge_steps(X);
    const uint32_t o0 = (uint32_t)((uint64_t)steps * part / S);
    const uint32_t o1 = (uint32_t)((uint64_t)steps * (part + 1) / S);
    if (o0 >= o1) return;
    const int64_t v = tile * kWave + lane;
    const P acc = merge_any<T, P>(f + fl.cell(0, v), fbase, fl.rcells, X, k, o0, o1);
    atomicMin(&M[(int64_t)xi * np + v], acc);
}

// ----------------------------------------- merge, 16-bit cells, two vertices
// With 16-bit table cells every pair sum is at most the distance bound, below
// 0xFFFF, so one 32-bit load brings the cells of two adjacent vertices and one
// packed 16-bit add gives both sums without a carry between the halves.  Each
// half is keyed as (sum << 16) | Y (one shift-or, or one and-or for the high
// half), which orders like (sum << k) | Y for k <= 16: the smallest sum, then
// the smallest Y.  Half the load instructions of merge_outer per cell pair.
__device__ __forceinline__ uint32_t add_u16x2(uint32_t a, uint32_t c) {
    const emv_u16x2 s = __builtin_bit_cast(emv_u16x2, a) + __builtin_bit_cast(emv_u16x2, c);
    return __builtin_bit_cast(uint32_t, s);
}

// Rows are read through raw buffer loads: a wave-uniform buffer resource on
// the row base (SGPRs, num_records = 0xFFFFFFFF, the gfx9 raw-buffer word 3)
// plus the lane's 32-bit byte offset, so a load needs no 64-bit vector
// address arithmetic.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t row_rsrc(const char *p) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<char *>(p), (short)0, -1, 0x00020000);
}

__device__ __forceinline__ uint32_t ld_pair(__amdgpu_buffer_rsrc_t r, uint32_t voff) {
    return __builtin_amdgcn_raw_buffer_load_b32(r, (int)voff, 0, 0);
}

// As merge_outer over the outer steps [o0, o1), for the vertex pair at byte
// offset lb (4-aligned, < 256) from the wave-uniform byte base fb; klo/khi
// fold the keys of the low / high vertex; rows are rb bytes apart.  Row
// offsets of the three low bits L stay in per-lane 32-bit offsets, which
// needs L * rb + 256 < 2^32 for the largest L of the solve (the host checks
// it).
__device__ __forceinline__ void merge_outer_h2(const char *__restrict__ fb, uint32_t lb,
                                               int64_t rb, uint32_t X, uint32_t o0, uint32_t o1,
                                               uint32_t &klo, uint32_t &khi) {
    // kept in a VGPR so that (sum & hmask) | Y is one v_and_or_b32 with Y in an SGPR
    uint32_t hmask = 0xFFFF0000u;
    asm volatile("" : "+v"(hmask));
    // rb: bytes from a row's cell to the next row's (2 np rows, 256 tiles)
    const uint32_t Xp = X ^ (1u << (31 - __clz(X)));
    uint32_t L = 0, H = Xp;
    for (int i = 0; i < 3 && H; i++) {
        const uint32_t lo = H & (~H + 1u);
        L |= lo;
        H ^= lo;
    }
    const int nl = __popc(L);
    const uint32_t nH = 1u << __popc(H);
    const int64_t oX = (int64_t)X * rb;
    if (nl == 3) {
        uint32_t yl[8], va[8], vc[8];
        low_subsets(L, yl);
        // a row Yh|yl[j] at (fb + Yh*rb) + va[j]; its partner X\Y at
        // (fb + oX - Yh*rb - L*rb) + vc[j]
#pragma unroll
        for (int j = 0; j < 8; j++) {
            va[j] = lb + (uint32_t)((int64_t)yl[j] * rb);
            vc[j] = lb + (uint32_t)((int64_t)(L - yl[j]) * rb);
        }
        const int64_t oL = (int64_t)L * rb;
        // per-j keys (sum << 16) | Yh: for a fixed j, Y = Yh | yl[j] orders
        // like Yh, so yl[j] is OR-ed in once at the end
        uint32_t alo[8], ahi[8];
#pragma unroll
        for (int j = 0; j < 8; j++) alo[j] = ahi[j] = 0xFFFFFFFFu;
        uint32_t Yh = deposit_bits(nH - 1u - o0, H);
        const uint32_t oend = o1 < nH - 1u ? o1 : nH - 1u;  // Yh == 0 last, apart
        uint32_t o = o0;
        for (; o + 2u <= oend; o += 2u) {  // two outer steps per iteration
            const uint32_t Y2 = (Yh - 1u) & H;
            const int64_t oh = (int64_t)Yh * rb, oh2 = (int64_t)Y2 * rb;
            const __amdgpu_buffer_rsrc_t pa = row_rsrc(fb + oh);
            const __amdgpu_buffer_rsrc_t pc = row_rsrc(fb + (oX - oh - oL));
            const __amdgpu_buffer_rsrc_t pa2 = row_rsrc(fb + oh2);
            const __amdgpu_buffer_rsrc_t pc2 = row_rsrc(fb + (oX - oh2 - oL));
            uint32_t a[8], c[8], a2[8], c2[8];
#pragma unroll
            for (int j = 0; j < 8; j++) {
                a[j] = ld_pair(pa, va[j]);
                c[j] = ld_pair(pc, vc[j]);
            }
#pragma unroll
            for (int j = 0; j < 8; j++) {
                a2[j] = ld_pair(pa2, va[j]);
                c2[j] = ld_pair(pc2, vc[j]);
            }
#pragma unroll
            for (int j = 0; j < 8; j++) {
                const uint32_t sm = add_u16x2(a[j], c[j]);
                const uint32_t sm2 = add_u16x2(a2[j], c2[j]);
                alo[j] = min(alo[j], min((sm << 16) | Yh, (sm2 << 16) | Y2));
                ahi[j] = min(ahi[j], min((sm & hmask) | Yh, (sm2 & hmask) | Y2));
            }
            Yh = (Y2 - 1u) & H;
        }
        if (o < oend) {
            const int64_t oh = (int64_t)Yh * rb;
            const __amdgpu_buffer_rsrc_t pa = row_rsrc(fb + oh);
            const __amdgpu_buffer_rsrc_t pc = row_rsrc(fb + (oX - oh - oL));
            uint32_t a[8], c[8];
#pragma unroll
            for (int j = 0; j < 8; j++) {
                a[j] = ld_pair(pa, va[j]);
                c[j] = ld_pair(pc, vc[j]);
            }
#pragma unroll
            for (int j = 0; j < 8; j++) {
                const uint32_t sm = add_u16x2(a[j], c[j]);
                alo[j] = min(alo[j], (sm << 16) | Yh);
                ahi[j] = min(ahi[j], (sm & hmask) | Yh);
            }
        }
        if (o1 == nH) {  // Yh == 0: the seven nonempty subsets of L
            const __amdgpu_buffer_rsrc_t pa = row_rsrc(fb);
            const __amdgpu_buffer_rsrc_t pc = row_rsrc(fb + (oX - oL));
            uint32_t a[8], c[8];
#pragma unroll
            for (int j = 1; j < 8; j++) {
                a[j] = ld_pair(pa, va[j]);
                c[j] = ld_pair(pc, vc[j]);
            }
#pragma unroll
            for (int j = 1; j < 8; j++) {
                const uint32_t sm = add_u16x2(a[j], c[j]);
                alo[j] = min(alo[j], sm << 16);
                ahi[j] = min(ahi[j], sm & hmask);
            }
        }
#pragma unroll
        for (int j = 0; j < 8; j++) {
            klo = min(klo, alo[j] | yl[j]);
            khi = min(khi, ahi[j] | yl[j]);
        }
        return;
    }
    for (uint32_t Y = L; Y != 0; Y = (Y - 1u) & L) {
        const uint32_t sm = add_u16x2(ld_pair(row_rsrc(fb + (int64_t)Y * rb), lb),
                                      ld_pair(row_rsrc(fb + (int64_t)(X ^ Y) * rb), lb));
        klo = min(klo, (sm << 16) | Y);
        khi = min(khi, (sm & hmask) | Y);
    }
}

// Host: whether a solve with tbits-bit cells, k terminals and row stride np
// takes the two-vertices-per-lane merge: 16-bit cells, keys that hold Y in 16
// bits, and lane offsets of the low rows within 32 bits (the low rows are
// the three lowest bits of X without its top bit, at most bits k-4..k-2).
// EMV_MERGE_H2=0 keeps one vertex per lane.
inline bool merge_h2_ok(int tbits, int k, int64_t rcells) {
    static const bool on = [] {
        const char *e = getenv("EMV_MERGE_H2");
        return !(e && atoi(e) == 0);
    }();
    if (!on || tbits != 16 || k > 16 || k < 2) return false;
    const uint64_t lmax = k >= 4 ? 7ull << (k - 4) : (1ull << (k - 1)) - 1ull;
    return 2ull * lmax * (uint64_t)rcells + 256ull < (1ull << 32);
}

// Host: vertex-wise rounds 1 .. R of a layer's relaxation that take every
// item (rounds 0 .. R-1 write no flags); EMV_DENSE_ROUNDS overrides R = 6.
inline int layer_dense_rounds() {
    static const int v = [] {
        const char *e = getenv("EMV_DENSE_ROUNDS");
        return e ? (atoi(e) > 0 ? atoi(e) : 0) : 6;
    }();
    return v;
}

// (sum << 16) | Y  ->  (sum << k) | Y; all ones (no pair) stays all ones.
template <typename P>
__device__ __forceinline__ P key_packed(uint32_t key, int k) {
    return key == 0xFFFFFFFFu ? ~(P)0 : (((P)(key >> 16) << k) | (P)(key & 0xFFFFu));
}

// k_merge_v for 16-bit cells: one wave = 128 consecutive vertices (two per
// lane) x XG rows.  np is a multiple of 64, so in the last tile of an np
// that is not a multiple of 128 the upper half of the lanes has no vertices:
// they read the last pair and store nothing.
template <typename P, typename DT, int XG, int CW = kWave>
__global__ __launch_bounds__(kBlock) void k_merge_h2(const uint16_t *__restrict__ f, FLayout fl, int64_t np,
                                                     const uint32_t *__restrict__ xs, int nx,
                                                     int nsub, int64_t ntiles, int k,
                                                     DT *__restrict__ D, P *__restrict__ CODE,
                                                     int nxcd, int64_t per_xcd) {
    const int lane = threadIdx.x & (kWave - 1);
    const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    if ((int64_t)(blockIdx.x / nxcd) >= per_xcd) return;
    const int64_t item = xcd_item(nxcd, per_xcd) * (kBlock / kWave) + wid;
    const int64_t tile = item / nsub;
    const int sub = (int)(item - tile * nsub);
    if (tile >= ntiles) return;
    const int64_t v0 = tile * (2 * kWave);
    const int64_t v = v0 + 2 * lane;
    const bool live = v < np;
    const char *fb = reinterpret_cast<const char *>(f + fl.cell(0, v0));
    const uint32_t lb = 4u * (uint32_t)min((int64_t)lane, (np - v0) / 2 - 1);
    const DT INF = (DT)~(DT)0;
    uint32_t klo[XG], khi[XG];
#pragma unroll
    for (int jx = 0; jx < XG; jx++) {
        const int xi = sub * XG + jx;
        klo[jx] = khi[jx] = 0xFFFFFFFFu;
        if (xi < nx) {
            const uint32_t X = xs[xi];
            merge_outer_h2(fb, lb, 2 * fl.rcells, X, 0u, merge_steps(X), klo[jx], khi[jx]);
        }
    }
    if (!live) return;
    const int chunk = sub / (CW / XG);
    const int loff = (sub % (CW / XG)) * XG;
    DT *dd = D + ((int64_t)chunk * np + v) * CW + loff;
    P *cd = CODE + ((int64_t)chunk * np + v) * CW + loff;
#pragma unroll
    for (int jx = 0; jx < XG; jx++) {
        const P plo = key_packed<P>(klo[jx], k), phi = key_packed<P>(khi[jx], k);
        dd[jx] = klo[jx] == 0xFFFFFFFFu ? INF : (DT)(klo[jx] >> 16);
        cd[jx] = plo;
        dd[CW + jx] = khi[jx] == 0xFFFFFFFFu ? INF : (DT)(khi[jx] >> 16);
        cd[CW + jx] = phi;
    }
}

// k_merge_split for 16-bit cells, two vertices per lane (128-vertex tiles).
template <typename P>
__global__ __launch_bounds__(kBlock) void k_merge_split_h2(const uint16_t *__restrict__ f,
                                                           FLayout fl, int64_t np,
                                                           const uint32_t *__restrict__ xs,
                                                           int nx, int S, int64_t ntiles, int k,
                                                           P *__restrict__ M, int nxcd,
                                                           int64_t per_xcd) {
    const int lane = threadIdx.x & (kWave - 1);
    const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    if ((int64_t)(blockIdx.x / nxcd) >= per_xcd) return;
    const int64_t item = xcd_item(nxcd, per_xcd) * (kBlock / kWave) + wid;
    const int64_t per_tile = (int64_t)nx * S;
    const int64_t tile = item / per_tile;
    if (tile >= ntiles) return;
    const int rem = (int)(item - tile * per_tile);
    const int xi = rem / S;
    const int part = rem - xi * S;
    const uint32_t X = xs[xi];
    const uint32_t steps = merge_steps(X);
    const uint32_t o0 = (uint32_t)((uint64_t)steps * part / S);
    const uint32_t o1 = (uint32_t)((uint64_t)steps * (part + 1) / S);
    if (o0 >= o1) return;
    const int64_t v0 = tile * (2 * kWave);
    const int64_t v = v0 + 2 * lane;
    const bool live = v < np;
    const uint32_t lb = 4u * (uint32_t)min((int64_t)lane, (np - v0) / 2 - 1);
    uint32_t klo = 0xFFFFFFFFu, khi = 0xFFFFFFFFu;
    merge_outer_h2(reinterpret_cast<const char *>(f + fl.cell(0, v0)), lb, 2 * fl.rcells, X, o0,
                   o1, klo, khi);
    if (!live) return;
    atomicMin(&M[(int64_t)xi * np + v], key_packed<P>(klo, k));
    atomicMin(&M[(int64_t)xi * np + v + 1], key_packed<P>(khi, k));
}

// ------------------------------------------- merge by stars (16-bit cells)
// Every pair {Y, X\Y} of X has exactly one side without a chosen element a of
// X: with S = X \ {a}, the pairs of X are {T, (S\T) | a} over the nonempty
// T subset of S.  Rows X_i = S | a_i that share S therefore share every row T:
// a star of r rows loads each T once and its r partners, (1 + r) loads per r
// pairs instead of 2 r.  The packed key takes the side without X's top bit
// (the smaller set), as merge_outer does: T, unless top(X) is in T (a_i below
// top(S) and top(S) in T), then X ^ T.
#ifndef EMV_STAR_MAX
#define EMV_STAR_MAX 8
#endif
constexpr int kStarMax = EMV_STAR_MAX;  // rows per star (4 .. 8)
static_assert(kStarMax >= 4 && kStarMax <= 8, "stars of 4 to 8 rows");
// rows per star over the compact int32 copy with row bases (more registers a
// pair: 64-bit keys with the bases added; a 1024 x 1024 grid at k = 13 merged
// in 428 / 425 / 506 ms with stars of 3 / 4 / 6 rows)
#ifndef EMV_STAR_REL
#define EMV_STAR_REL 4
#endif
constexpr int kStarRel = EMV_STAR_REL;
static_assert(kStarRel >= 1 && kStarRel <= kStarMax, "relative stars of 1 to kStarMax rows");
// R clamped to kStarMax (to RMAX), so every case of a dispatch over star
// sizes compiles; stars of at most RMAX rows never reach the clamped cases
template <int R>
constexpr int star_r() { return R <= kStarMax ? R : kStarMax; }
template <int RMAX, int R>
constexpr int star_rw() { return R <= RMAX ? R : RMAX; }

struct MergeStar {
    uint32_t S;                 // the shared set, m - 1 elements
    uint32_t r;                 // rows in the star, 1 .. kStarMax
    uint32_t a[kStarMax];       // the extra element of each row (one bit)
    int32_t xi[kStarMax];       // the row's index in the layer batch
};

// Stars covering the batch's rows xs[0..nx) (all of m elements): each
// uncovered row, in order, takes the base set S = X without the element that
// leaves the most uncovered rows S | a in the batch, and up to kStarMax of
// them (X first, then by increasing a) form its star.
inline void build_stars(const uint32_t *xs, int nx, int k, std::vector<MergeStar> &out,
                        int rmax = kStarMax) {
    out.clear();
    std::vector<int32_t> idx((size_t)1 << k, -1);
    for (int i = 0; i < nx; i++) idx[xs[i]] = i;
    std::vector<char> covered((size_t)nx, 0);
    auto free_row = [&](uint32_t Y) { return idx[Y] >= 0 && !covered[(size_t)idx[Y]]; };
    for (int i = 0; i < nx; i++) {
        if (covered[(size_t)i]) continue;
        const uint32_t X = xs[i];
        uint32_t bestS = 0;
        int best = -1;
        for (uint32_t rest = X; rest; rest &= rest - 1) {
            const uint32_t S = X ^ (rest & (~rest + 1u));
            int cnt = 0;
            for (int b = 0; b < k; b++)
                if (!(S >> b & 1u) && free_row(S | (1u << b))) cnt++;
            if (cnt > best) {
                best = cnt;
                bestS = S;
            }
        }
        MergeStar st;
        memset(&st, 0, sizeof(st));
        st.S = bestS;
        st.a[0] = X ^ bestS;
        st.xi[0] = i;
        covered[(size_t)i] = 1;
        st.r = 1;
        for (int b = 0; b < k && st.r < (uint32_t)rmax; b++) {
            const uint32_t Y = bestS | (1u << b);
            if ((bestS >> b & 1u) || !free_row(Y)) continue;
            st.a[st.r] = 1u << b;
            st.xi[st.r] = idx[Y];
            covered[(size_t)idx[Y]] = 1;
            st.r++;
        }
        out.push_back(st);
    }
}

// Outer steps of a star: subsets Th of S's high part H (S without its three
// lowest bits L), as merge_steps.
__device__ __forceinline__ uint32_t star_steps(uint32_t S) {
    const int r = __popc(S);
    return r <= 3 ? 1u : (1u << (r - 3));
}

// The pairs of a star's rows over the outer steps [o0, o1), for the vertex
// pair at byte offset lb (4-aligned, < 256) from the wave-uniform byte base fb
// of a 128-vertex tile; klo / khi fold the (sum << 16) | Y keys of the low /
// high vertex of each row (all ones where a row has no pair yet).
// R = the star's rows, a compile-time count: the loads of all R rows of an
// outer step then issue before the first wait (a runtime count puts a branch
// and a wait between the rows, and a wave keeps 8 loads in flight instead
// of 8 (R + 1)).
template <int R>
__device__ __forceinline__ void star_h2_pairs_r(const char *__restrict__ fb, uint32_t lb,
                                                int64_t rb, uint32_t S,
                                                const uint32_t (&a)[kStarMax], uint32_t o0,
                                                uint32_t o1, uint32_t (&klo)[kStarMax],
                                                uint32_t (&khi)[kStarMax]) {
    constexpr int r = R;
    uint32_t hmask = 0xFFFF0000u;
    asm volatile("" : "+v"(hmask));
    // rb: bytes from a row's cell to the next row's
    const uint32_t topS = 1u << (31 - __clz(S));
    uint32_t L = 0, H = S;
    for (int q = 0; q < 3 && H; q++) {
        const uint32_t lo = H & (~H + 1u);
        L |= lo;
        H ^= lo;
    }
    const uint32_t nH = 1u << __popc(H);
#pragma unroll
    for (int i = 0; i < kStarMax; i++) klo[i] = khi[i] = 0xFFFFFFFFu;
    // the key's set for row i: T, or X_i ^ T when top(X_i) = top(S) lies in T
    auto yset = [&](int i, uint32_t T) -> uint32_t {
        return (a[i] < topS && (T & topS)) ? ((S | a[i]) ^ T) : T;
    };
    // |S| >= 4: L holds S's three lowest bits with a fourth above them, so
    // L <= 7 * 2^(k-4) and its row offsets fit the 32-bit lane offsets (the
    // host's merge_h2_ok bound); smaller S take the loop below
    if (__popc(S) >= 4) {
        uint32_t yl[8], va[8], vc[8];
        low_subsets(L, yl);
#pragma unroll
        for (int j = 0; j < 8; j++) {
            va[j] = lb + (uint32_t)((int64_t)yl[j] * rb);
            vc[j] = lb + (uint32_t)((int64_t)(L - yl[j]) * rb);
        }
        uint32_t Th = deposit_bits(nH - 1u - o0, H);
        const uint32_t oend = o1 < nH - 1u ? o1 : nH - 1u;  // Th == 0 last, apart
        for (uint32_t o = o0; o < oend; o++) {
            const __amdgpu_buffer_rsrc_t pt = row_rsrc(fb + (int64_t)Th * rb);
            uint32_t t[8];
#pragma unroll
            for (int j = 0; j < 8; j++) t[j] = ld_pair(pt, va[j]);
            // top(S) lies in H, so whether the key's set flips depends on Th
            // alone: the set of pair j is B_i ^ yl[j] with B_i per row and step
            const bool flip = (Th & topS) != 0u;
#pragma unroll
            for (int i = 0; i < kStarMax; i++) {
                if (i < r) {
                    const __amdgpu_buffer_rsrc_t pp =
                        row_rsrc(fb + (int64_t)((H ^ Th) | a[i]) * rb);
                    uint32_t pv[8];
#pragma unroll
                    for (int j = 0; j < 8; j++) pv[j] = ld_pair(pp, vc[j]);
                    const uint32_t B = (flip && a[i] < topS) ? ((S | a[i]) ^ Th) : Th;
#pragma unroll
                    for (int j = 0; j < 8; j++) {
                        const uint32_t sm = add_u16x2(t[j], pv[j]);
                        // wave-uniform set; the low half's key as a shift-add
                        // (the shifted sum has zero low bits)
                        const uint32_t Y = __builtin_amdgcn_readfirstlane(B ^ yl[j]);
                        klo[i] = min(klo[i], (sm << 16) + Y);
                        khi[i] = min(khi[i], (sm & hmask) | Y);
                    }
                }
            }
            Th = (Th - 1u) & H;
        }
        if (o1 == nH) {  // Th == 0: the seven nonempty subsets of L
            const __amdgpu_buffer_rsrc_t pt = row_rsrc(fb);
            uint32_t t[8];
#pragma unroll
            for (int j = 1; j < 8; j++) t[j] = ld_pair(pt, va[j]);
#pragma unroll
            for (int i = 0; i < kStarMax; i++) {
                if (i < r) {
                    const __amdgpu_buffer_rsrc_t pp = row_rsrc(fb + (int64_t)(H | a[i]) * rb);
                    uint32_t pv[8];
#pragma unroll
                    for (int j = 1; j < 8; j++) pv[j] = ld_pair(pp, vc[j]);
#pragma unroll
                    for (int j = 1; j < 8; j++) {
                        const uint32_t sm = add_u16x2(t[j], pv[j]);
                        const uint32_t Y = yset(i, yl[j]);
                        klo[i] = min(klo[i], (sm << 16) | Y);
                        khi[i] = min(khi[i], (sm & hmask) | Y);
                    }
                }
            }
        }
    } else {
        // |S| <= 3: one outer step over the nonempty subsets of S
        for (uint32_t T = S; T != 0; T = (T - 1u) & S) {
            const uint32_t t = ld_pair(row_rsrc(fb + (int64_t)T * rb), lb);
#pragma unroll
            for (int i = 0; i < kStarMax; i++) {
                if (i < r) {
                    const uint32_t sm =
                        add_u16x2(t, ld_pair(row_rsrc(fb + (int64_t)((S ^ T) | a[i]) * rb), lb));
                    const uint32_t Y = yset(i, T);
                    klo[i] = min(klo[i], (sm << 16) | Y);
                    khi[i] = min(khi[i], (sm & hmask) | Y);
                }
            }
        }
    }
}

__device__ __forceinline__ void star_h2_pairs(const char *__restrict__ fb, uint32_t lb,
                                              int64_t rb, uint32_t S, int r,
                                              const uint32_t (&a)[kStarMax], uint32_t o0,
                                              uint32_t o1, uint32_t (&klo)[kStarMax],
                                              uint32_t (&khi)[kStarMax]) {
    switch (r) {
    case 1: star_h2_pairs_r<1>(fb, lb, rb, S, a, o0, o1, klo, khi); break;
    case 2: star_h2_pairs_r<2>(fb, lb, rb, S, a, o0, o1, klo, khi); break;
    case 3: star_h2_pairs_r<3>(fb, lb, rb, S, a, o0, o1, klo, khi); break;
    case 4: star_h2_pairs_r<4>(fb, lb, rb, S, a, o0, o1, klo, khi); break;
    case 5: star_h2_pairs_r<star_r<5>()>(fb, lb, rb, S, a, o0, o1, klo, khi); break;
    case 6: star_h2_pairs_r<star_r<6>()>(fb, lb, rb, S, a, o0, o1, klo, khi); break;
    case 7: star_h2_pairs_r<star_r<7>()>(fb, lb, rb, S, a, o0, o1, klo, khi); break;
    default: star_h2_pairs_r<star_r<8>()>(fb, lb, rb, S, a, o0, o1, klo, khi); break;
    }
}

// Item = (tile, star, part), tile-major per XCD; parts fold into M with
// unsigned atomicMin like k_merge_split_h2.  Two vertices per lane.
template <typename P>
__global__ __launch_bounds__(kBlock) void k_merge_star_h2(
    const uint16_t *__restrict__ f, FLayout fl, int64_t np, const MergeStar *__restrict__ stars,
    int nstars,
    int parts, int64_t ntiles, int k, P *__restrict__ M, int nxcd, int64_t per_xcd) {
    const int lane = threadIdx.x & (kWave - 1);
    const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    if ((int64_t)(blockIdx.x / nxcd) >= per_xcd) return;
    const int64_t item = xcd_item(nxcd, per_xcd) * (kBlock / kWave) + wid;
    const int64_t per_tile = (int64_t)nstars * parts;
    const int64_t tile = item / per_tile;
    if (tile >= ntiles) return;
    const int rem = (int)(item - tile * per_tile);
    const int g = rem / parts;
    const int part = rem - g * parts;
    const uint32_t S = stars[g].S;
    const int r = (int)stars[g].r;
    uint32_t a[kStarMax];
#pragma unroll
    for (int i = 0; i < kStarMax; i++) a[i] = stars[g].a[i];
    const uint32_t steps = star_steps(S);
    const uint32_t o0 = (uint32_t)((uint64_t)steps * part / parts);
    const uint32_t o1 = (uint32_t)((uint64_t)steps * (part + 1) / parts);
    if (o0 >= o1) return;
    const int64_t v0 = tile * (2 * kWave);
    const int64_t v = v0 + 2 * lane;
    const bool live = v < np;
    const uint32_t lb = 4u * (uint32_t)min((int64_t)lane, (np - v0) / 2 - 1);
    const char *fb = reinterpret_cast<const char *>(f + fl.cell(0, v0));
    uint32_t klo[kStarMax], khi[kStarMax];
    star_h2_pairs(fb, lb, 2 * fl.rcells, S, r, a, o0, o1, klo, khi);
    if (!live) return;
#pragma unroll
    for (int i = 0; i < kStarMax; i++) {
        if (i < r) {
            P *acc = &M[(int64_t)stars[g].xi[i] * np + v];
            atomicMin(acc, key_packed<P>(klo[i], k));
            atomicMin(acc + 1, key_packed<P>(khi[i], k));
        }
    }
}

// ------------------------------------------ merge by grids (16-bit cells)
// A grid is the rows X_ij = S | a_i | b_j (i < p, j < q, p, q <= 3) of a layer
// batch: S has m - 2 elements, the a_i and b_j are distinct elements outside
// it.  Every pair {T, X_ij \ T} has one side T without b_j, so T runs over
// the nonempty U of S and over U | a_i (every U of S, the empty one too),
// with partners (S \ U) | a_i | b_j and (S \ U) | b_j.  For one U the grid
// loads f[U], the p rows f[U | a_i], the q rows f[(S \ U) | b_j] and the p q
// rows f[(S \ U) | a_i | b_j]: 1 + p + q + p q loads for 2 p q pairs (16 for
// 18 at 3 x 3), against the (1 + r) for r of a star of r rows.
struct MergeGrid {
    uint32_t S;
    uint32_t p, q;
    uint32_t a[3], b[3];
    int32_t xi[9];  // X_ij's index in the layer batch at 3 i + j
};

// Greedy grid cover of the batch's rows xs[0..nx) (all of m >= 4 elements):
// each uncovered row X, in order, tries its element pairs (a < b) as (a_0,
// b_0) with S = X \ {a, b}, grows A = {a_i} and B = {b_j} by the lowest
// elements outside S whose rows with every member of the other side are
// uncovered rows of the batch (A first, alternating, up to 3 each), and keeps
// the largest grid (the first 3 x 3 ends the search).
inline void build_grids(const uint32_t *xs, int nx, int k, std::vector<MergeGrid> &out) {
    out.clear();
    std::vector<int32_t> idx((size_t)1 << k, -1);
    for (int i = 0; i < nx; i++) idx[xs[i]] = i;
    std::vector<char> covered((size_t)nx, 0);
    auto free_row = [&](uint32_t Y) { return idx[Y] >= 0 && !covered[(size_t)idx[Y]]; };
    for (int i = 0; i < nx; i++) {
        if (covered[(size_t)i]) continue;
        const uint32_t X = xs[i];
        MergeGrid best;
        memset(&best, 0, sizeof(best));
        int best_sz = 0;
        for (uint32_t ra = X; ra && best_sz < 9; ra &= ra - 1) {
            const uint32_t a = ra & (~ra + 1u);
            for (uint32_t rb = ra & (ra - 1); rb && best_sz < 9; rb &= rb - 1) {
                const uint32_t b = rb & (~rb + 1u);
                const uint32_t S = X ^ a ^ b;
                uint32_t A[3] = {a, 0, 0}, B[3] = {b, 0, 0};
                int p = 1, q = 1;
                uint32_t used = a | b;
                bool grown = true;
                while (grown) {
                    grown = false;
                    if (p < 3)
                        for (int c = 0; c < k; c++) {
                            const uint32_t cb = 1u << c;
                            if ((S | used) & cb) continue;
                            bool ok = true;
                            for (int j = 0; j < q && ok; j++) ok = free_row(S | cb | B[j]);
                            if (ok) {
                                A[p++] = cb;
                                used |= cb;
                                grown = true;
                                break;
                            }
                        }
                    if (q < 3)
                        for (int c = 0; c < k; c++) {
                            const uint32_t cb = 1u << c;
                            if ((S | used) & cb) continue;
                            bool ok = true;
                            for (int j = 0; j < p && ok; j++) ok = free_row(S | A[j] | cb);
                            if (ok) {
                                B[q++] = cb;
                                used |= cb;
                                grown = true;
                                break;
                            }
                        }
                }
                if (p * q > best_sz) {
                    best_sz = p * q;
                    best.S = S;
                    best.p = (uint32_t)p;
                    best.q = (uint32_t)q;
                    for (int t = 0; t < 3; t++) {
                        best.a[t] = A[t];
                        best.b[t] = B[t];
                    }
                }
            }
        }
        for (int t = 0; t < 9; t++) best.xi[t] = -1;
        for (uint32_t ii = 0; ii < best.p; ii++)
            for (uint32_t jj = 0; jj < best.q; jj++) {
                const int32_t r = idx[best.S | best.a[ii] | best.b[jj]];
                best.xi[3 * ii + jj] = r;
                covered[(size_t)r] = 1;
            }
        out.push_back(best);
    }
}

// The pairs of a p x q grid over the outer steps [o0, o1) of U's high part
// (S without its two lowest elements L; U = Uh | ul over the four subsets ul
// of L), for the vertex pair at byte offset lb from the wave-uniform base fb
// of a 128-vertex tile, rows rb bytes apart; klo / khi[3 i + j] fold the
// (sum << 16) | Y keys of row X_ij's low / high vertex.  Y, the side without
// X_ij's top element: T, or X_ij ^ T when the top lies in T (a_i in T = U |
// a_i, or S's top in U), so Y = T ^ m with m = X_ij or 0 per step.
template <int P_, int Q_>
__device__ __forceinline__ void grid_h2_pairs(const char *__restrict__ fb, uint32_t lb, int64_t rb,
                                              uint32_t S, const uint32_t (&a)[3],
                                              const uint32_t (&b)[3], uint32_t o0, uint32_t o1,
                                              uint32_t (&klo)[9], uint32_t (&khi)[9]) {
    uint32_t hmask = 0xFFFF0000u;
    asm volatile("" : "+v"(hmask));
    const uint32_t l0 = S & (~S + 1u);
    const uint32_t l1 = (S ^ l0) & (~(S ^ l0) + 1u);
    const uint32_t L = l0 | l1;
    const uint32_t H = S ^ L;
    const uint32_t topS = 1u << (31 - __clz(S));  // in H: S has at least three elements
    const uint32_t nH = 1u << __popc(H);
    const uint32_t ul[4] = {0u, l0, l1, L};
    uint32_t va[4], vc[4];
#pragma unroll
    for (int u = 0; u < 4; u++) {
        va[u] = lb + (uint32_t)((int64_t)ul[u] * rb);
        vc[u] = lb + (uint32_t)((int64_t)(L ^ ul[u]) * rb);
    }
#pragma unroll
    for (int t = 0; t < 9; t++) klo[t] = khi[t] = 0xFFFFFFFFu;
    uint32_t Uh = deposit_bits(nH - 1u - o0, H);
    for (uint32_t o = o0; o < o1; o++) {
        const uint32_t Hc = H ^ Uh;  // (S \ U) = Hc | (L ^ ul)
        uint32_t t[4], ta[P_][4], pb[Q_][4], pab[P_][Q_][4];
        {
            const __amdgpu_buffer_rsrc_t r = row_rsrc(fb + (int64_t)Uh * rb);
#pragma unroll
            for (int u = 0; u < 4; u++) t[u] = ld_pair(r, va[u]);
        }
#pragma unroll
        for (int i = 0; i < P_; i++) {
            const __amdgpu_buffer_rsrc_t r = row_rsrc(fb + (int64_t)(Uh | a[i]) * rb);
#pragma unroll
            for (int u = 0; u < 4; u++) ta[i][u] = ld_pair(r, va[u]);
        }
#pragma unroll
        for (int j = 0; j < Q_; j++) {
            const __amdgpu_buffer_rsrc_t r = row_rsrc(fb + (int64_t)(Hc | b[j]) * rb);
#pragma unroll
            for (int u = 0; u < 4; u++) pb[j][u] = ld_pair(r, vc[u]);
        }
#pragma unroll
        for (int i = 0; i < P_; i++)
#pragma unroll
            for (int j = 0; j < Q_; j++) {
                const __amdgpu_buffer_rsrc_t r = row_rsrc(fb + (int64_t)(Hc | a[i] | b[j]) * rb);
#pragma unroll
                for (int u = 0; u < 4; u++) pab[i][j][u] = ld_pair(r, vc[u]);
            }
        const bool stop = (Uh & topS) != 0u;  // S's top in U
        const bool last = Uh == 0u;           // U = ul: U = 0 has no first pair
#pragma unroll
        for (int i = 0; i < P_; i++)
#pragma unroll
            for (int j = 0; j < Q_; j++) {
                const uint32_t X = S | a[i] | b[j];
                const uint32_t top = 1u << (31 - __clz(X));
                // first pair {U, X \ U}: flips when S's top is X's top and in U;
                // second {U | a_i, X \ (U | a_i)}: when a_i is X's top, or S's
                // top is and lies in U
                const uint32_t m1 = (top == topS && stop) ? X : 0u;
                const uint32_t m2 = (top == a[i] || (top == topS && stop)) ? X : 0u;
                const uint32_t B1 = Uh ^ m1, B2 = (Uh | a[i]) ^ m2;
#pragma unroll
                for (int u = 0; u < 4; u++) {
                    const uint32_t s2 = add_u16x2(ta[i][u], pb[j][u]);
                    const uint32_t Y2 = __builtin_amdgcn_readfirstlane(B2 ^ ul[u]);
                    klo[3 * i + j] = min(klo[3 * i + j], (s2 << 16) + Y2);
                    khi[3 * i + j] = min(khi[3 * i + j], (s2 & hmask) | Y2);
                    if (u > 0 || !last) {
                        const uint32_t s1 = add_u16x2(t[u], pab[i][j][u]);
                        const uint32_t Y1 = __builtin_amdgcn_readfirstlane(B1 ^ ul[u]);
                        klo[3 * i + j] = min(klo[3 * i + j], (s1 << 16) + Y1);
                        khi[3 * i + j] = min(khi[3 * i + j], (s1 & hmask) | Y1);
                    }
                }
            }
        Uh = (Uh - 1u) & H;
    }
}

// grid_h2_pairs for a grid of p x q rows, pq = 4 p + q.
__device__ __forceinline__ void grid_h2_dispatch(int pq, const char *__restrict__ fb, uint32_t lb,
                                                 int64_t rb, uint32_t S, const uint32_t (&a)[3],
                                                 const uint32_t (&b)[3], uint32_t o0, uint32_t o1,
                                                 uint32_t (&klo)[9], uint32_t (&khi)[9]) {
    switch (pq) {
    case 1 * 4 + 1: grid_h2_pairs<1, 1>(fb, lb, rb, S, a, b, o0, o1, klo, khi); break;
    case 1 * 4 + 2: grid_h2_pairs<1, 2>(fb, lb, rb, S, a, b, o0, o1, klo, khi); break;
    case 1 * 4 + 3: grid_h2_pairs<1, 3>(fb, lb, rb, S, a, b, o0, o1, klo, khi); break;
    case 2 * 4 + 1: grid_h2_pairs<2, 1>(fb, lb, rb, S, a, b, o0, o1, klo, khi); break;
    case 2 * 4 + 2: grid_h2_pairs<2, 2>(fb, lb, rb, S, a, b, o0, o1, klo, khi); break;
    case 2 * 4 + 3: grid_h2_pairs<2, 3>(fb, lb, rb, S, a, b, o0, o1, klo, khi); break;
    case 3 * 4 + 1: grid_h2_pairs<3, 1>(fb, lb, rb, S, a, b, o0, o1, klo, khi); break;
    case 3 * 4 + 2: grid_h2_pairs<3, 2>(fb, lb, rb, S, a, b, o0, o1, klo, khi); break;
    default: grid_h2_pairs<3, 3>(fb, lb, rb, S, a, b, o0, o1, klo, khi); break;
    }
}

// Outer steps of a grid: subsets of S without its two lowest elements.
__device__ __forceinline__ uint32_t grid_steps(uint32_t S) { return 1u << (__popc(S) - 2); }

// Item = (tile, grid, part), tile-major per XCD; parts fold into M with
// unsigned atomicMin as the stars do.  Two vertices per lane.
template <typename P>
__global__ __launch_bounds__(kBlock) void k_merge_grid_h2(
    const uint16_t *__restrict__ f, FLayout fl, int64_t np, const MergeGrid *__restrict__ grids,
    int ngrids, int parts, int64_t ntiles, int k, P *__restrict__ M, int nxcd, int64_t per_xcd) {
    const int lane = threadIdx.x & (kWave - 1);
    const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    if ((int64_t)(blockIdx.x / nxcd) >= per_xcd) return;
    const int64_t item = xcd_item(nxcd, per_xcd) * (kBlock / kWave) + wid;
    const int64_t per_tile = (int64_t)ngrids * parts;
    const int64_t tile = item / per_tile;
    if (tile >= ntiles) return;
    const int rem = (int)(item - tile * per_tile);
    const int g = rem / parts;
    const int part = rem - g * parts;
    const MergeGrid *gr = grids + g;
    const uint32_t S = gr->S;
    const int pq = (int)(gr->p * 4 + gr->q);
    uint32_t a[3], b[3];
#pragma unroll
    for (int i = 0; i < 3; i++) {
        a[i] = gr->a[i];
        b[i] = gr->b[i];
    }
    const uint32_t steps = grid_steps(S);
    const uint32_t o0 = (uint32_t)((uint64_t)steps * part / parts);
    const uint32_t o1 = (uint32_t)((uint64_t)steps * (part + 1) / parts);
    if (o0 >= o1) return;
    const int64_t v0 = tile * (2 * kWave);
    const int64_t v = v0 + 2 * lane;
    const uint32_t lb = 4u * (uint32_t)min((int64_t)lane, (np - v0) / 2 - 1);
    const char *fb = reinterpret_cast<const char *>(f + fl.cell(0, v0));
    const int64_t rb = 2 * fl.rcells;
    uint32_t klo[9], khi[9];
    grid_h2_dispatch(pq, fb, lb, rb, S, a, b, o0, o1, klo, khi);
    if (v >= np) return;
#pragma unroll
    for (int t = 0; t < 9; t++) {
        const int32_t xi = gr->xi[t];
        if (xi >= 0) {
            P *acc = &M[(int64_t)xi * np + v];
            atomicMin(acc, key_packed<P>(klo[t], k));
            atomicMin(acc + 1, key_packed<P>(khi[t], k));
        }
    }
}

// --------------------------------------------- merge by stars (32-bit cells)
// The stars of k_merge_star_h2 over an int32 table or over the compact int32
// copy of an int64 table (REL: a row holds f - fbase[row], and a pair's sum
// adds both rows' bases back), one vertex per lane on 64-vertex tiles (256 B
// per row).  The lane offsets of the low rows would not fit 32 bits with
// 4-byte cells, so every row is read from a scalar 64-bit base (the row's
// tile segment) at the lane's fixed offset.  The loop body is instantiated per
// star size R, so that the loads of all R + 1 rows of a step are in flight
// together.
template <typename P, bool REL, int R>
__device__ __forceinline__ void starw_pairs(const int32_t *__restrict__ fv0,
                                            const int64_t *__restrict__ fbase, int64_t rs,
                                            uint32_t S, const uint32_t *a, uint32_t o0,
                                            uint32_t o1, int k, P *acc) {
    const int lane = __lane_id();
    const uint32_t topS = 1u << (31 - __clz(S));
    uint32_t L = 0, H = S;
    for (int q = 0; q < 3 && H; q++) {
        const uint32_t lo = H & (~H + 1u);
        L |= lo;
        H ^= lo;
    }
    const uint32_t nH = 1u << __popc(H);
    // (sum of the two cells, with their rows' bases) << k | Y
    auto key = [&](int32_t x, int32_t y, int64_t b, uint32_t Y) -> P {
        P sum = (P)(uint32_t)x + (P)(uint32_t)y;
        if constexpr (REL) sum += (P)b;
        return (sum << k) | (P)Y;
    };
    if (__popc(S) >= 4) {
        uint32_t yl[8];
        low_subsets(L, yl);
        uint32_t Th = deposit_bits(nH - 1u - o0, H);
        const uint32_t oend = o1 < nH - 1u ? o1 : nH - 1u;  // Th == 0 last, apart
        for (uint32_t o = o0; o < oend; o++) {
            const int32_t *tb = fv0 + (int64_t)Th * rs;
            int32_t t[8], pv[R][8];
#pragma unroll
            for (int j = 0; j < 8; j++) t[j] = tb[(int64_t)yl[j] * rs + lane];
#pragma unroll
            for (int i = 0; i < R; i++) {
                const int32_t *pb = fv0 + (int64_t)((H ^ Th) | a[i]) * rs;
#pragma unroll
                for (int j = 0; j < 8; j++) pv[i][j] = pb[(int64_t)(L - yl[j]) * rs + lane];
            }
            const bool flip = (Th & topS) != 0u;
#pragma unroll
            for (int i = 0; i < R; i++) {
                const uint32_t B = (flip && a[i] < topS) ? ((S | a[i]) ^ Th) : Th;
                const uint32_t Pi = (H ^ Th) | a[i];
#pragma unroll
                for (int j = 0; j < 8; j++) {
                    int64_t b = 0;
                    if constexpr (REL) b = fbase[Th | yl[j]] + fbase[Pi | (L - yl[j])];
                    const P s2 = key(t[j], pv[i][j], b, B ^ yl[j]);
                    acc[i] = s2 < acc[i] ? s2 : acc[i];
                }
            }
            Th = (Th - 1u) & H;
        }
        if (o1 == nH) {  // Th == 0: the seven nonempty subsets of L (topS lies in H: Y = T)
            int32_t t[8], pv[R][8];
#pragma unroll
            for (int j = 1; j < 8; j++) t[j] = fv0[(int64_t)yl[j] * rs + lane];
#pragma unroll
            for (int i = 0; i < R; i++) {
                const int32_t *pb = fv0 + (int64_t)(H | a[i]) * rs;
#pragma unroll
                for (int j = 1; j < 8; j++) pv[i][j] = pb[(int64_t)(L - yl[j]) * rs + lane];
            }
#pragma unroll
            for (int i = 0; i < R; i++) {
#pragma unroll
                for (int j = 1; j < 8; j++) {
                    int64_t b = 0;
                    if constexpr (REL) b = fbase[yl[j]] + fbase[(H | a[i]) | (L - yl[j])];
                    const P s2 = key(t[j], pv[i][j], b, yl[j]);
                    acc[i] = s2 < acc[i] ? s2 : acc[i];
                }
            }
        }
    } else {
        // |S| <= 3: one outer step over the nonempty subsets of S
        for (uint32_t T = S; T != 0; T = (T - 1u) & S) {
            const int32_t t = fv0[(int64_t)T * rs + lane];
#pragma unroll
            for (int i = 0; i < R; i++) {
                const uint32_t Pr = (S ^ T) | a[i];
                int64_t b = 0;
                if constexpr (REL) b = fbase[T] + fbase[Pr];
                const uint32_t Y = (a[i] < topS && (T & topS)) ? ((S | a[i]) ^ T) : T;
                const P s2 = key(t, fv0[(int64_t)Pr * rs + lane], b, Y);
                acc[i] = s2 < acc[i] ? s2 : acc[i];
            }
        }
    }
}

// Item = (64-vertex tile, star, part), tile-major per XCD; parts fold into M
// with unsigned atomicMin.  np is a multiple of 64, so every lane's vertex
// lies inside the rows (padding columns feed only dropped results).
template <typename P, bool REL, int RMAX>
__global__ __launch_bounds__(kBlock) void k_merge_star_w(
    const int32_t *__restrict__ f, FLayout fl, const int64_t *__restrict__ fbase, int64_t np,
    const MergeStar *__restrict__ stars, int nstars, int parts, int64_t ntiles, int k,
    P *__restrict__ M, int nxcd, int64_t per_xcd) {
    const int lane = threadIdx.x & (kWave - 1);
    const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    if ((int64_t)(blockIdx.x / nxcd) >= per_xcd) return;
    const int64_t item = xcd_item(nxcd, per_xcd) * (kBlock / kWave) + wid;
    const int64_t per_tile = (int64_t)nstars * parts;
    const int64_t tile = item / per_tile;
    if (tile >= ntiles) return;
    const int rem = (int)(item - tile * per_tile);
    const int g = rem / parts;
    const int part = rem - g * parts;
    const uint32_t S = stars[g].S;
    const int r = (int)stars[g].r;
    uint32_t a[kStarMax];
#pragma unroll
    for (int i = 0; i < kStarMax; i++) a[i] = stars[g].a[i];
    const uint32_t steps = star_steps(S);
    const uint32_t o0 = (uint32_t)((uint64_t)steps * part / parts);
    const uint32_t o1 = (uint32_t)((uint64_t)steps * (part + 1) / parts);
    if (o0 >= o1) return;
    const int64_t v0 = tile * kWave;
    const int32_t *fv0 = f + fl.cell(0, v0);
    P acc[kStarMax];
#pragma unroll
    for (int i = 0; i < kStarMax; i++) acc[i] = ~(P)0;
    switch (r) {
    case 1: starw_pairs<P, REL, 1>(fv0, fbase, fl.rcells, S, a, o0, o1, k, acc); break;
    case 2: starw_pairs<P, REL, 2>(fv0, fbase, fl.rcells, S, a, o0, o1, k, acc); break;
    case 3: starw_pairs<P, REL, 3>(fv0, fbase, fl.rcells, S, a, o0, o1, k, acc); break;
    case 4: starw_pairs<P, REL, 4>(fv0, fbase, fl.rcells, S, a, o0, o1, k, acc); break;
    case 5: starw_pairs<P, REL, star_rw<RMAX, 5>()>(fv0, fbase, fl.rcells, S, a, o0, o1, k, acc); break;
    case 6: starw_pairs<P, REL, star_rw<RMAX, 6>()>(fv0, fbase, fl.rcells, S, a, o0, o1, k, acc); break;
    case 7: starw_pairs<P, REL, star_rw<RMAX, 7>()>(fv0, fbase, fl.rcells, S, a, o0, o1, k, acc); break;
    default: starw_pairs<P, REL, star_rw<RMAX, 8>()>(fv0, fbase, fl.rcells, S, a, o0, o1, k, acc); break;
    }
    const int64_t v = v0 + lane;
#pragma unroll
    for (int i = 0; i < kStarMax; i++)
        if (i < r) atomicMin(&M[(int64_t)stars[g].xi[i] * np + v], acc[i]);
}

// ------------------------------------------------------- cube merge (LDS)
// Layers of small subsets (16-bit cells, two vertices per lane): a workgroup
// takes one 128-vertex tile and one cube U of at most kCubeU elements.  It
// stages in LDS the tile's segment of every lower row T of U (1 <= |T| < m),
// at T's local mask (T's bits within U: an order-preserving relabelling), and
// then merges the layer rows X of U that the host assigned to the cube: the
// least packed key over the pairs {Y, X\Y}, Y a nonempty subset of X without
// its top element, both rows read from LDS.  A staged row is read from L2/HBM
// once per (tile, cube) and serves every pair of every assigned row that
// holds it (0.34-0.72 row loads per pair at k = 14 with the greedy cover,
// against 1.25-1.3 for the stars); the pairs run at LDS and VALU rate.
// Keys are (sum << 16) | Y with Y local (the order of local masks is the
// order of their global sets, so the least key names the same split); the
// winner becomes (sum << k) | Y with Y global, folded into M[xi][v] with an
// unsigned atomicMin (a row's pairs may be cut into parts over waves).
// Items = (tile, cube), tile-major per XCD, so that the cubes of a tile run
// together on one XCD and share its L2.
constexpr int kCubeU = 9;       // elements per cube: 2^9 local rows x 256 B = 128 KB of LDS
constexpr int kCubeLds = ((1 << kCubeU) * kWave + 2 * 128) * 4;  // + the row list (<= 126 rows)
constexpr int kCubeWaves = 16;  // waves per workgroup

// deposit_bits for a cube: the bits of t (t < 2^popc(U) <= 2^kCubeU) into U.
__device__ __forceinline__ uint32_t deposit_cube(uint32_t t, uint32_t U) {
    uint32_t r = 0;
#pragma unroll
    for (int q = 0; q < kCubeU; q++) {
        const uint32_t low = U & (~U + 1u);
        if (t >> q & 1u) r |= low;
        U ^= low;
    }
    return r;
}

struct MergeCube {
    uint32_t U;     // the cube (global mask)
    int32_t e0;     // its rows: entries e0 .. e0 + nx - 1 of the row list
    int32_t nx;
    int32_t parts;  // pair ranges per row
};

struct CubeRow {
    uint32_t Xl;  // the row X as a local mask within its cube
    int32_t xi;   // its index in the layer batch
};

// One persistent workgroup per CU walks the items of its XCD's range (tile-
// major), one item at a time: while it merges item i from LDS, the loads of
// item i + 1's rows and row list are in flight into registers; after a
// barrier they go to LDS and the next item starts.  Rows of the table that
// the items read are the lower layers only, which no launch of this layer
// writes.
// rowtab[cube * 2^kCubeU + 32 * w + q]: the table offset, in units of 128
// bytes, of the row that wave w stages as local row t = w + 16 q (its set
// deposited into the cube), or all ones for a local row the merges never read
// (m or more elements, or t >= 2^u).  Built by the host, so that staging
// costs each wave one scalar table read per row.
template <typename P>
__global__ __launch_bounds__(kCubeWaves * kWave) void k_merge_cube_h2(
    const uint16_t *__restrict__ f, FLayout fl, int64_t np, const MergeCube *__restrict__ cubes,
    const CubeRow *__restrict__ rows, const uint32_t *__restrict__ rowtab, int ncubes,
    int64_t ntiles, int k, P *__restrict__ M, int nxcd, int64_t per_xcd) {
    extern __shared__ uint32_t lds_rows[];  // [2^u][64] staged rows, then the row list
    const int lane = threadIdx.x & (kWave - 1);
    const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    // this workgroup's items: the XCD's range [x * per, (x + 1) * per), every
    // (gridDim / nxcd)-th from its slot
    const int64_t slots = gridDim.x / nxcd;
    const int64_t slot = blockIdx.x / nxcd;
    const int64_t first = (int64_t)(blockIdx.x % nxcd) * per_xcd;
    const int64_t total = ntiles * ncubes;
    constexpr int kPer = (1 << kCubeU) / kCubeWaves;  // staged rows per wave
    constexpr int kRowList = (1 << kCubeU) * kWave;   // word offset of the row list in LDS
    uint32_t staged[kPer];
    uint32_t cur_U = 0;
    int cur_nx = 0, cur_parts = 0;
    int64_t cur_tile = 0;
    CubeRow cur_row = {0u, 0};
    // loads of item `it` into registers: its rows (wave w: local rows w,
    // w + 16, ..., those the merges read) and, in waves 0 and 1, row list
    // entry 64 * wave + lane (< nx)
    auto prefetch = [&](int64_t it, uint32_t &U, int &nx, int &parts, int64_t &tile,
                        CubeRow &rl) {
        tile = it / ncubes;
        const int cb = (int)(it - tile * ncubes);
        U = cubes[cb].U;
        nx = cubes[cb].nx;
        parts = cubes[cb].parts;
        const int e0 = cubes[cb].e0;
        const int64_t v0 = tile * (2 * kWave);
        // lanes past the row read its last pair; their results are dropped
        const uint32_t lb = 4u * (uint32_t)min((int64_t)lane, (np - v0) / 2 - 1);
        const char *fb = reinterpret_cast<const char *>(f + fl.cell(0, v0)) + lb;
        const uint32_t *tab = rowtab + (int64_t)cb * (1 << kCubeU) + wid * kPer;
#pragma unroll
        for (int q = 0; q < kPer; q++) {
            const uint32_t off = tab[q];
            staged[q] = 0u;
            if (off != 0xFFFFFFFFu)
                staged[q] = *reinterpret_cast<const uint32_t *>(fb + (int64_t)off * 128);
        }
        const int ei = wid * kWave + lane;  // row list entries: waves 0 and 1
        if (wid < 2 && ei < nx) rl = rows[e0 + ei];
    };
    int64_t it = first + slot;
    const int64_t last = min(first + per_xcd, total);
    if (it >= last) return;
    prefetch(it, cur_U, cur_nx, cur_parts, cur_tile, cur_row);
    uint32_t hmask = 0xFFFF0000u;
    asm volatile("" : "+v"(hmask));
    for (; it < last; it += slots) {
        __syncthreads();  // the previous item's merges are done with LDS
        {
            // every local row is written (unread ones with a zero word)
#pragma unroll
            for (int q = 0; q < kPer; q++)
                lds_rows[(uint32_t)(wid + q * kCubeWaves) * kWave + lane] = staged[q];
            const int ei = wid * kWave + lane;
            if (wid < 2 && ei < cur_nx) {
                lds_rows[kRowList + 2 * ei] = cur_row.Xl;
                lds_rows[kRowList + 2 * ei + 1] = (uint32_t)cur_row.xi;
            }
        }
        __syncthreads();
        const uint32_t U = cur_U;
        const int nx = cur_nx, parts = cur_parts;
        const int64_t v0 = cur_tile * (2 * kWave);
        if (it + slots < last) prefetch(it + slots, cur_U, cur_nx, cur_parts, cur_tile, cur_row);
        const int ntask = nx * parts;
        for (int task = wid; task < ntask; task += kCubeWaves) {
            const int e = task / parts;
            const int part = task - e * parts;
            const uint32_t Xl = __builtin_amdgcn_readfirstlane(lds_rows[kRowList + 2 * e]);
            const int32_t xi = (int32_t)__builtin_amdgcn_readfirstlane(lds_rows[kRowList + 2 * e + 1]);
            const uint32_t Xp = Xl ^ (1u << (31 - __clz(Xl)));
            const uint32_t npairs = (1u << __popc(Xp)) - 1u;
            const uint32_t i0 = npairs * (uint32_t)part / (uint32_t)parts;
            const uint32_t i1 = npairs * (uint32_t)(part + 1) / (uint32_t)parts;
            if (i0 >= i1) continue;
            // the (i0 + 1)-th nonempty subset of Xp in increasing order, then
            // the next ones by (Y - Xp) & Xp
            uint32_t Y = deposit_bits(i0 + 1u, Xp);
            uint32_t klo = 0xFFFFFFFFu, khi = 0xFFFFFFFFu;
            uint32_t i = i0;
            for (; i + 4u <= i1; i += 4u) {
                uint32_t Ys[4];
                Ys[0] = Y;
#pragma unroll
                for (int q = 1; q < 4; q++) Ys[q] = (Ys[q - 1] - Xp) & Xp;
                Y = (Ys[3] - Xp) & Xp;
                uint32_t a[4], c[4];
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    a[q] = lds_rows[Ys[q] * kWave + lane];
                    c[q] = lds_rows[(Xl ^ Ys[q]) * kWave + lane];
                }
#pragma unroll
                for (int q = 0; q < 4; q += 2) {
                    const uint32_t s0 = add_u16x2(a[q], c[q]);
                    const uint32_t s1 = add_u16x2(a[q + 1], c[q + 1]);
                    klo = min(klo, min((s0 << 16) | Ys[q], (s1 << 16) | Ys[q + 1]));
                    khi = min(khi, min((s0 & hmask) | Ys[q], (s1 & hmask) | Ys[q + 1]));
                }
            }
            for (; i < i1; i++) {
                const uint32_t s0 =
                    add_u16x2(lds_rows[Y * kWave + lane], lds_rows[(Xl ^ Y) * kWave + lane]);
                klo = min(klo, (s0 << 16) | Y);
                khi = min(khi, (s0 & hmask) | Y);
                Y = (Y - Xp) & Xp;
            }
            const int64_t v = v0 + 2 * lane;
            if (v < np) {
                const P plo = ((P)(klo >> 16) << k) | (P)deposit_cube(klo & 0xFFFFu, U);
                const P phi = ((P)(khi >> 16) << k) | (P)deposit_cube(khi & 0xFFFFu, U);
                P *acc = &M[(int64_t)xi * np + v];
                atomicMin(acc, plo);
                atomicMin(acc + 1, phi);
            }
        }
    }
}

// Packed accumulators M[xi][v] -> vertex-major staging (label, packed word)
// through an LDS transpose of 64 rows x 64 vertices per workgroup; rows past
// nx are INF.  Workgroup = (group of 64 rows, vertex tile); a chunk of CW
// rows holds CW / 64 groups side by side.
template <typename P, typename DT, int CW = kWave>
__global__ __launch_bounds__(kBlock) void k_pack_v(const P *__restrict__ M, int64_t np, int nx,
                                                   int k, DT *__restrict__ D,
                                                   P *__restrict__ CODE, int64_t vtiles) {
    __shared__ P sm[kWave][kWave + 1];
    const int64_t grp64 = blockIdx.x / vtiles;
    const int64_t vt = blockIdx.x - grp64 * vtiles;
    const int64_t chunk = grp64 / (CW / kWave);
    const int64_t roff = (grp64 % (CW / kWave)) * kWave;  // the group's first row in its chunk
    const int t = threadIdx.x;
    const int lane = t & (kWave - 1);
    const int grp = t / kWave;
    const int64_t v0 = vt * kWave;
#pragma unroll
    for (int r = 0; r < kWave / (kBlock / kWave); r++) {
        const int xl = grp + r * (kBlock / kWave);
        const int64_t xi = grp64 * kWave + xl;
        sm[xl][lane] = (xi < nx) ? M[xi * np + v0 + lane] : ~(P)0;
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kWave / (kBlock / kWave); r++) {
        const int vv = grp + r * (kBlock / kWave);
        const P a = sm[lane][vv];
        const int64_t idx = (chunk * np + v0 + vv) * CW + roff + lane;
        const bool real = (grp64 * kWave + lane) < nx;
        D[idx] = real ? (DT)(a >> k) : ~(DT)0;
        CODE[idx] = real ? a : ~(P)0;
    }
}

// ----------------------------------------------------------------- records
// Record of one converged cell, lane = row of chunk c at vertex v (wave-
// uniform): the merge split Y when the label still equals the merged value
// (packed word pk = (sum << k) | Y), else 2^k + the smallest neighbour u with
// label(u) + w == label(v).  Neighbour entries are read as wave-uniform
// (scalar) loads, eight gathers in flight.
// Dc[u * stride + lane] holds the label of the lane's row at vertex u.
template <typename DT, typename P>
__device__ __forceinline__ int32_t cell_record(const DT *__restrict__ Dc, DT L, P pk, int k,
                                               bool real, int v, const int32_t *__restrict__ rp,
                                               const Edge *__restrict__ adj,
                                               int64_t stride = kWave) {
    const int lane = __lane_id();
    const bool edge = real && (P)L < (pk >> k);
    int32_t code = (int32_t)(pk & (((P)1 << k) - 1));
    if (__ballot(edge) == 0) return code;
    const DT INF = (DT)~(DT)0;
    uint32_t bu = 0xFFFFFFFFu;
    const int e0 = rp[v], e1 = rp[v + 1];
    for (int eb = e0; eb < e1; eb += 8) {
        int uu[8], ww[8];
        DT du[8];
#pragma unroll
        for (int q = 0; q < 8; q++) {
            const int e = min(eb + q, e1 - 1);
            uu[q] = adj[e].v;
            ww[q] = adj[e].w;
            du[q] = Dc[(int64_t)uu[q] * stride + lane];
        }
#pragma unroll
        for (int q = 0; q < 8; q++)
            if (eb + q < e1 && du[q] != INF && (DT)(du[q] + (DT)ww[q]) == L &&
                (uint32_t)uu[q] < bu)
                bu = (uint32_t)uu[q];
    }
    if (edge) code = (int32_t)((1u << k) + bu);
    return code;
}

// ----------------------------------------------------------------- unpack
// Vertex-major staging -> table rows f[X], b[X] through an LDS transpose of
// 64 vertices x 64 rows per workgroup; each wave decides the records of 16
// vertices (cell_record) on the way into LDS.  Item = (group of 64 rows,
// vertex tile); a chunk of CW rows holds CW / 64 groups side by side.
// rspan (relative labels, else null): row xi's labels are f - rspan[2 xi].
template <typename T, typename DT, typename P, int CW = kWave>
__global__ __launch_bounds__(kBlock) void k_unpack_v(const DT *__restrict__ D,
                                                     const P *__restrict__ CODE, int64_t np,
                                                     const uint32_t *__restrict__ xs, int nx,
                                                     int n, int k, const int32_t *__restrict__ rp,
                                                     const Edge *__restrict__ adj,
                                                     T *__restrict__ f, FLayout fl,
                                                     int32_t *__restrict__ b,
                                                     int nxcd, int64_t per_xcd, int64_t vtiles,
                                                     const unsigned long long *__restrict__ rspan,
                                                     int32_t *__restrict__ fcopy, FLayout fcl,
                                                     int64_t *__restrict__ fbase) {
    __shared__ DT sd[kWave][kWave + 1];
    __shared__ int32_t sc[kWave][kWave + 1];
    if ((int64_t)(blockIdx.x / nxcd) >= per_xcd) return;
    const int64_t item = xcd_item(nxcd, per_xcd);
    const int64_t g64 = item / vtiles;
    const int64_t vt = item - g64 * vtiles;
    if (g64 * kWave >= nx) return;
    const int64_t chunk = g64 / (CW / kWave);
    const int64_t roff = (g64 % (CW / kWave)) * kWave;  // the group's first row in its chunk
    const int t = threadIdx.x;
    const int lane = t & (kWave - 1);
    const int grp = __builtin_amdgcn_readfirstlane(t / kWave);
    const int64_t v0 = vt * kWave;
    const DT *Dc = D + chunk * np * CW + roff;
    const bool real = g64 * kWave + lane < nx;
    // relative labels: compare with the merged value less the row's base
    const P pbase = (rspan && real) ? ((P)rspan[2 * (g64 * kWave + lane)] << k) : (P)0;
    // the labels and packed words of kPre vertices are loaded together, then
    // their records decided one vertex at a time
    constexpr int kPre = 4;
#pragma unroll 1
    for (int r0 = 0; r0 < kWave / (kBlock / kWave); r0 += kPre) {
        DT Ls[kPre];
        P Ps[kPre];
#pragma unroll
        for (int u = 0; u < kPre; u++) {
            const int64_t idx =
                ((chunk * np) + v0 + grp + (r0 + u) * (kBlock / kWave)) * CW + roff + lane;
            Ls[u] = D[idx];
            Ps[u] = CODE[idx];
        }
#pragma unroll
        for (int u = 0; u < kPre; u++) {
            const int vv = grp + (r0 + u) * (kBlock / kWave);
            const int64_t v = v0 + vv;
            int32_t code = 0;
            if (v < n)
                code = cell_record<DT, P>(Dc, Ls[u], Ps[u] - pbase, k, real, (int)v, rp, adj, CW);
            sd[vv][lane] = Ls[u];
            sc[vv][lane] = code;
        }
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kWave / (kBlock / kWave); r++) {
        const int xl = grp + r * (kBlock / kWave);
        const int64_t xi = g64 * kWave + xl;
        const int64_t v = v0 + lane;
        if (xi < nx && v < n) {
            const int64_t X = xs[xi];
            const unsigned long long base = rspan ? rspan[2 * xi] : 0ull;
            f[fl.cell(X, v)] = (T)((unsigned long long)sd[lane][xl] + base);
            b[X * np + v] = sc[lane][xl];
            // compact copy (relative labels that fit int32): the relative
            // label itself, with the row's base
            if (fcopy) {
                fcopy[fcl.cell(X, v)] = (int32_t)sd[lane][xl];
                if (v == 0) fbase[X] = (int64_t)base;
            }
        }
    }
}

// ------------------------------------------- unpack, chunks of 128 16-bit rows
// k_unpack_v for 16-bit labels staged in chunks of 128 rows: a wave takes all
// 128 rows of a vertex, two per lane (the 32-bit word of rows 2l, 2l + 1), so
// the neighbour labels an edge record looks for are one 256-byte read per
// neighbour.  Item = (chunk, 64-vertex tile), one workgroup of kWideWaves
// waves, each deciding the records of 64 / kWideWaves vertices; then the
// workgroup writes the chunk's rows of the tile's vertices through LDS.
constexpr int kWideWaves = 8;

// Records of rows 2l (lo) and 2l + 1 (hi) of a converged word L at vertex v,
// the merged packed words p0, p1: the rule of cell_record, for two rows.
template <typename P>
__device__ __forceinline__ void cell_record2(const uint32_t *__restrict__ Dw, uint32_t L, P p0,
                                             P p1, int k, bool real0, bool real1, int v,
                                             const int32_t *__restrict__ rp,
                                             const Edge *__restrict__ adj, int32_t &code0,
                                             int32_t &code1) {
    const int lane = __lane_id();
    const uint32_t L0 = L & 0xFFFFu, L1 = L >> 16;
    const bool edge0 = real0 && (P)L0 < (p0 >> k);
    const bool edge1 = real1 && (P)L1 < (p1 >> k);
    code0 = (int32_t)(p0 & (((P)1 << k) - 1));
    code1 = (int32_t)(p1 & (((P)1 << k) - 1));
    if (__ballot(edge0 || edge1) == 0) return;
    uint32_t bu0 = 0xFFFFFFFFu, bu1 = 0xFFFFFFFFu;
    const int e0 = rp[v], e1 = rp[v + 1];
    for (int eb = e0; eb < e1; eb += 8) {
        int uu[8], ww[8];
        uint32_t du[8];
#pragma unroll
        for (int q = 0; q < 8; q++) {
            const int e = min(eb + q, e1 - 1);
            uu[q] = adj[e].v;
            ww[q] = adj[e].w;
            du[q] = Dw[(int64_t)uu[q] * kWave + lane];
        }
#pragma unroll
        for (int q = 0; q < 8; q++) {
            if (eb + q >= e1) continue;
            const uint32_t d0 = du[q] & 0xFFFFu, d1 = du[q] >> 16;
            // INF stays out: a finite label + w is below 0xFFFF
            if (d0 != 0xFFFFu && d0 + (uint32_t)ww[q] == L0 && (uint32_t)uu[q] < bu0) bu0 = (uint32_t)uu[q];
            if (d1 != 0xFFFFu && d1 + (uint32_t)ww[q] == L1 && (uint32_t)uu[q] < bu1) bu1 = (uint32_t)uu[q];
        }
    }
    if (edge0) code0 = (int32_t)((1u << k) + bu0);
    if (edge1) code1 = (int32_t)((1u << k) + bu1);
}

template <typename T, typename P>
__global__ __launch_bounds__(kWideWaves * kWave) void k_unpack_w(
    const uint32_t *__restrict__ D, const P *__restrict__ CODE, int64_t np,
    const uint32_t *__restrict__ xs, int nx, int n, int k, const int32_t *__restrict__ rp,
    const Edge *__restrict__ adj, T *__restrict__ f, FLayout fl, int32_t *__restrict__ b,
    int nxcd, int64_t per_xcd, int64_t vtiles) {
    constexpr int CW = 2 * kWave;
    // [vertex][row]; odd row strides (in 32-bit words) keep the column reads
    // of the write-out free of bank conflicts
    __shared__ uint32_t sd[kWave][kWave + 1];  // the 128 labels of a vertex as 64 words
    __shared__ int32_t sc[kWave][CW + 1];
    if ((int64_t)(blockIdx.x / nxcd) >= per_xcd) return;
    const int64_t item = xcd_item(nxcd, per_xcd);
    const int64_t chunk = item / vtiles;
    const int64_t vt = item - chunk * vtiles;
    if (chunk * CW >= nx) return;
    const int t = threadIdx.x;
    const int lane = t & (kWave - 1);
    const int wv = __builtin_amdgcn_readfirstlane(t / kWave);
    const int64_t v0 = vt * kWave;
    const uint32_t *Dw = D + chunk * np * kWave;
    const bool real0 = chunk * CW + 2 * lane < nx;
    const bool real1 = chunk * CW + 2 * lane + 1 < nx;
    constexpr int kPer = kWave / kWideWaves;  // vertices per wave
    constexpr int kPre = 4;                   // loaded together
#pragma unroll 1
    for (int r0 = 0; r0 < kPer; r0 += kPre) {
        uint32_t Ls[kPre];
        P P0[kPre], P1[kPre];
#pragma unroll
        for (int u = 0; u < kPre; u++) {
            const int64_t vv = v0 + wv + (r0 + u) * kWideWaves;
            Ls[u] = Dw[vv * kWave + lane];
            const P *cp = CODE + (chunk * np + vv) * CW + 2 * lane;
            P0[u] = cp[0];
            P1[u] = cp[1];
        }
#pragma unroll
        for (int u = 0; u < kPre; u++) {
            const int vl = wv + (r0 + u) * kWideWaves;
            const int64_t v = v0 + vl;
            int32_t c0 = 0, c1 = 0;
            if (v < n)
                cell_record2<P>(Dw, Ls[u], P0[u], P1[u], k, real0, real1, (int)v, rp, adj, c0, c1);
            sd[vl][lane] = Ls[u];
            sc[vl][2 * lane] = c0;
            sc[vl][2 * lane + 1] = c1;
        }
    }
    __syncthreads();
    // write-out: row xl of the chunk, lane = vertex
    const int64_t v = v0 + lane;
#pragma unroll 1
    for (int xl = wv; xl < CW; xl += kWideWaves) {
        const int64_t xi = chunk * CW + xl;
        if (xi < nx && v < n) {
            const int64_t X = xs[xi];
            const uint32_t w = sd[lane][xl >> 1];
            f[fl.cell(X, v)] = (T)((xl & 1) ? (w >> 16) : (w & 0xFFFFu));
            b[X * np + v] = sc[lane][xl];
        }
    }
}

// Compact copy of the singleton rows 2^t (base 0): int32 distances.
template <typename CT>
__global__ void k_copy_singletons(const int64_t *__restrict__ f, int64_t np, int n, int k,
                                  CT *__restrict__ fcopy, FLayout fcl) {
    const int t = blockIdx.y;
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < k && v < n) {
        const int64_t X = 1ll << t;
        fcopy[fcl.cell(X, v)] = (CT)f[X * np + v];
    }
}

// Per row xi of the staged batch: the minimum and maximum merged value over
// the real vertices, rspan[2 xi] and rspan[2 xi + 1] (start: all ones / 0).
// Relaxation only lowers labels and a relaxed label is at least the row's
// minimum merged value (f[X][v] = min over u of merged[u] + dist(u, v)), so
// every label of the row stays in [min, max].
template <typename P>
__global__ __launch_bounds__(kBlock) void k_row_span(const P *__restrict__ CODE, int64_t np,
                                                     int n, int nx, int k,
                                                     unsigned long long *__restrict__ rspan,
                                                     int64_t vper) {
    const int lane = threadIdx.x & (kWave - 1);
    const int64_t chunk = blockIdx.y;
    const int64_t xi = chunk * kWave + lane;
    if (xi >= nx) return;
    const int64_t v0 = ((int64_t)blockIdx.x * (kBlock / kWave) + threadIdx.x / kWave) * vper;
    const int64_t v1 = min(v0 + vper, (int64_t)n);
    unsigned long long lo = ~0ull, hi = 0ull;
    for (int64_t v = v0; v < v1; v++) {
        const unsigned long long x = (unsigned long long)(CODE[(chunk * np + v) * kWave + lane] >> k);
        lo = x < lo ? x : lo;
        hi = x > hi ? x : hi;
    }
    if (v0 < v1) {
        atomicMin(&rspan[2 * xi], lo);
        atomicMax(&rspan[2 * xi + 1], hi);
    }
}

template <typename U>
__global__ void k_fill_span(U *__restrict__ rspan, int64_t rows) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < rows) {
        rspan[2 * i] = ~(U)0;
        rspan[2 * i + 1] = (U)0;
    }
}

// Relative 32-bit labels from the merged values: merged - rspan[2 xi] for
// the real vertices of real rows, else INF.
template <typename P>
__global__ void k_rel_labels(const P *__restrict__ CODE, int64_t np, int n, int nx, int k,
                             const unsigned long long *__restrict__ rspan,
                             uint32_t *__restrict__ D, int64_t cells) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; i < cells; i += stride) {
        const int lane = (int)(i & (kWave - 1));
        const int64_t cv = i / kWave;  // chunk * np + v
        const int64_t chunk = cv / np;
        const int64_t v = cv - chunk * np;
        const int64_t xi = chunk * kWave + lane;
        uint32_t L = 0xFFFFFFFFu;
        if (xi < nx && v < n) L = (uint32_t)((unsigned long long)(CODE[i] >> k) - rspan[2 * xi]);
        D[i] = L;
    }
}

// Singleton rows: f[1<<t][v] = dist(kk[t], v) from chunk 0 lane t; record
// (kk[t], {t}) for every v (reader-el.c:2015-2027).  Also the per-row
// maximum distance (all ones if a vertex is unreached), which bounds the
// table: f[X][v] <= sum over t in X of dist(kk[t], v).
template <typename T, typename DT>
__global__ __launch_bounds__(kBlock) void k_unpack_singletons_v(
    const DT *__restrict__ D, int64_t np, const int32_t *__restrict__ kk, int k, int n,
    T *__restrict__ f, FLayout fl, int32_t *__restrict__ b, unsigned long long *__restrict__ rowmax) {
    const int t = blockIdx.y;
    const int v = blockIdx.x * kBlock + threadIdx.x;
    unsigned long long d = 0;
    if (v < n) {
        const DT x = D[(int64_t)v * kWave + t];
        d = (x == (DT)~(DT)0) ? ~0ull : (unsigned long long)x;
        if (f) {
            const int64_t X = 1ll << t;
            f[fl.cell(X, v)] = (T)d;
            b[X * np + v] = (v == kk[t]) ? (int32_t)X : (int32_t)(((int64_t)1 << k) + kk[t]);
        }
    }
    for (int off = 32; off > 0; off >>= 1) {
        unsigned long long o = __shfl_xor(d, off);
        d = o > d ? o : d;
    }
    if (__lane_id() == 0) atomicMax(&rowmax[t], d);
}

// Distances and predecessors of one shortest-path query (chunk 0, lane 0):
// the predecessor is the smallest neighbour u with d(u) + w == d(v); the
// source (distance 0, weights >= 1) and unreached vertices have none.
template <typename DT>
__global__ void k_unpack_sssp_v(const DT *__restrict__ D, int n, const int32_t *__restrict__ rp,
                                const Edge *__restrict__ adj, int64_t *__restrict__ d,
                                int64_t *__restrict__ pred) {
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n) return;
    const DT INF = (DT)~(DT)0;
    const DT dd = D[(int64_t)v * kWave];
    int64_t p = -1;
    if (dd != 0 && dd != INF) {
        for (int e = rp[v]; e < rp[v + 1]; e++) {
            const int u = adj[e].v;
            const DT du = D[(int64_t)u * kWave];
            if (du != INF && (uint64_t)du + (uint64_t)adj[e].w == (uint64_t)dd && (p < 0 || u < p))
                p = u;
        }
    }
    d[v] = (dd == INF) ? INT64_MAX : (int64_t)dd;
    pred[v] = p;
}

// -------------------------------------------------------------- traceback
// Depth-first walk of the records from (q, C) in the reference's order
// (reader-el.c:1907-1932): an edge record emits (v,u) and continues at
// (u, X); a merge record visits (v, Y) then (v, X\Y); Y == X stops.
// One lane walks; real trees are a few hundred steps.
[[maybe_unused]] static __global__ void k_traceback(const int32_t *__restrict__ b, int64_t np, int k, int64_t q,
                            int64_t C, int64_t *__restrict__ pairs, int64_t cap,
                            int64_t *__restrict__ stack, int64_t stcap,
                            int64_t *__restrict__ result) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const int64_t S = 1ll << k;
    int64_t sp = 0, npairs = 0;
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
            if (npairs >= cap) { result[0] = -1; return; }
            pairs[2 * npairs] = v;
            pairs[2 * npairs + 1] = u;
            npairs++;
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
    result[0] = npairs;
}

}  // namespace emv
