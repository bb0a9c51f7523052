// emv_batch_kernels.hpp -- gfx950 kernels of the multi-instance EMV engine.
//
// One engine step advances every active instance by one DP stage in the same
// launches: instances at their singleton stage contribute one chunk of k
// seeded shortest-path rows, instances at layer m contribute the chunks of
// their m-subsets.  Instances may sit on different graphs (batched solves)
// or share one resident graph (serving), and sit at different layers.
//
// The per-instance table keeps the single-instance layout in one arena
// region: costs f[tbase + X * np + v] (int32 cells, or 16-bit cells from the
// same base when the instance's bound allows), then the record codes
// b[tbase + 2^k * np + X * np + v] (int32).  The cell width is known only
// after the singleton stage (the bound): with k >= 3 that stage writes its
// rows as int32 at the end of the cost region (row t at cells
// (2^k - k + t) * np, clear of every singleton row and of 16-bit rows), and
// the first layer step places them (k_gplace) at the instance's width.
// Staging chunks keep the vertex-major layout, labels of type DT (16 or 32
// bits, per chunk group) at ((DT *)(D + 4 * dbase))[v * 64 + lane] -- dbase
// counts 4-byte cells -- and the merge's packed word (sum << k) | Y in
// CODE[dbase + v * 64 + lane], 32 or 64 bits wide (the step's P type).  Work items are located through prefix arrays
// over the descriptors (binary search), so instances of any size share one
// grid.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "emv_kernels.hpp"

namespace emv {

// One staging chunk: up to 64 rows of one instance.  Its packed words sit at
// CODE[dbase + v * 64 + lane]; its labels at ((DT *)(D + 4 * lbase))[v *
// lstride + lane].  lbase == dbase and lstride == 64 but for paired chunks
// (batched solves, 16-bit labels, vertex-wise rounds): chunks 2p and 2p + 1
// of an instance's layer keep their labels side by side, 128 rows per vertex
// in the first chunk's label region (lstride 128, the second at + 64 rows),
// so that the pair relaxes as one 128-row item whose neighbour reads are
// 256-byte blocks.
struct GChunk {
    int64_t dbase;  // cell offset of the chunk in the staging (np * 64 cells)
    int64_t fbase;  // offset of the chunk's activity flags (np words)
    int64_t tbase;  // element offset of the instance's table
    int64_t lbase;  // cell offset of the chunk's labels
    int32_t n, np;  // instance vertices, table row stride
    int32_t voff;   // first vertex of the instance's graph in the shared adjacency
    int32_t k;      // instance terminals; codes of edge records start at 2^k
    int32_t x0;     // first row of the chunk in the step's row list
    int32_t nx;     // rows in the chunk (lane j < nx is row x0 + j)
    int32_t kind;   // 0: layer rows, 1: singleton rows (row j = terminal j),
                    // 2: singleton rows of a group query (row j = group j)
    int32_t inst;   // engine slot of the instance
    int32_t b0;     // block relaxation: first block of the graph in the shared block arrays
    int32_t nblk;   //   and its blocks (0: vertex-wise relaxation)
    int32_t tbits;  // table cells: 16 or 32 bits
    int32_t lstride;  // labels per vertex: 64, or 128 for paired chunks
};

// One instance at a layer: its chunks are consecutive, its merge items are
// tile-major over all of them (so one tile's rows are read together).
struct GMerge {
    int32_t chunk0, nchunks;
    int32_t ntiles;  // np / 64
    int32_t nsub;    // nchunks * 64 / XG
    int32_t nrows;   // rows of the layer
    int32_t x0;      // first row in the step's row list
    int32_t k, np;
    int64_t tbase;   // the instance's table
    int64_t dbase0;  // staging cell of its first chunk; chunk q at dbase0 + q * np * 64
    int32_t S;       // split merge: parts per row (0: rows merged whole by k_gmerge)
    int32_t tbits;   // table cells: 16 or 32 bits (at the same tbase either way)
    int64_t mbase;   // split merge: accumulator of row xi at Macc[mbase + xi * np + v]
    int32_t h2;      // 16-bit cells merged two vertices per lane (128-vertex merge tiles)
    int32_t mtiles;  // merge tiles: np / 64, or ceil(np / 128) when h2
    int32_t lpair;   // chunks paired (an even count): labels as GChunk.lbase / lstride
    int32_t star0;   // star merge (nstars > 0): its stars in the step's star list
    int32_t nstars;
    int32_t pad_;
};

// Label base (cells) and labels per vertex of chunk q of a GMerge.
__device__ __forceinline__ int64_t merge_lbase(const GMerge &M, int64_t q) {
    const int64_t own = M.dbase0 + q * M.np * kWave;
    return M.lpair ? own - (q & 1) * M.np * kWave + (q & 1) * (kWave / 2) : own;
}

__device__ __forceinline__ int merge_lstride(const GMerge &M) {
    return M.lpair ? 2 * kWave : kWave;
}

template <typename DT>
__device__ __forceinline__ DT *chunk_labels(void *D, int64_t dbase) {
    return reinterpret_cast<DT *>(reinterpret_cast<char *>(D) + 4 * dbase);
}

__device__ __forceinline__ int find_seg(const int64_t *__restrict__ pre, int nseg, int64_t i) {
    // largest s with pre[s] <= i, pre[0] = 0, pre[nseg] = total
    int lo = 0, hi = nseg;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (pre[mid] <= i) lo = mid; else hi = mid;
    }
    return lo;
}

// The prefix array copied into LDS by the whole workgroup (when it has at
// most kSegCache entries), so that each wave's search costs LDS reads instead
// of dependent global loads.  Every thread of the workgroup must call it.
constexpr int kSegCache = 512;

__device__ __forceinline__ const int64_t *cache_pre(const int64_t *__restrict__ pre, int nseg,
                                                    int64_t *sh) {
    if (nseg + 1 > kSegCache) return pre;
    for (int i = threadIdx.x; i <= nseg; i += blockDim.x) sh[i] = pre[i];
    __syncthreads();
    return sh;
}

// ------------------------------------------------------ singleton seeding
// Every cell of every singleton chunk: INF label; then label 0 at the sources.
__global__ __launch_bounds__(kBlock) void k_gseed(const GChunk *__restrict__ ch,
                                                  const int64_t *__restrict__ cell_pre, int nch,
                                                  const int32_t *__restrict__ kk_all,
                                                  const int32_t *__restrict__ kk_off,
                                                  void *__restrict__ D) {
    __shared__ int64_t sh_pre[kSegCache];
    const int64_t *pre = cache_pre(cell_pre, nch, sh_pre);
    const int64_t total = pre[nch];
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    // a thread's cells increase, so its chunk is searched once and then
    // advanced across chunk boundaries
    int g = -1;
    int32_t src = -1;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += stride) {
        const int lane = (int)(i & (kWave - 1));  // chunk cell counts are multiples of 64
        if (g < 0 || pre[g + 1] <= i) {
            g = (g < 0) ? find_seg(pre, nch, i) : g;
            while (pre[g + 1] <= i) g++;
            const GChunk c = ch[g];
            src = lane < c.nx ? kk_all[kk_off[c.inst] + lane] : -1;
        }
        const int64_t off = i - pre[g];  // v * 64 + lane
        chunk_labels<uint32_t>(D, ch[g].dbase)[off] = ((int)(off / kWave) == src) ? 0u : ~0u;
    }
}

// Group queries: after k_gseed, label 0 at each (member, lane of its group),
// given as staging cells (the singleton chunks hold 32-bit labels).
__global__ __launch_bounds__(kBlock) void k_gseed_cells(const int64_t *__restrict__ cells,
                                                        int64_t count, void *__restrict__ D) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += stride)
        static_cast<uint32_t *>(D)[cells[i]] = 0u;
}

// ------------------------------------------------------------------ merge
// Same recurrence and packing as merge_one (reader-el.c:2068-2088), per
// instance: f rows of the instance's table, shift by the instance's k.
// A wave-uniform 64-bit value (computed from LDS or vector loads) moved to
// SGPRs, so that what derives from it is scalar too.
__device__ __forceinline__ int64_t wave_uniform(int64_t x) {
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)x);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)((uint64_t)x >> 32));
    return (int64_t)(((uint64_t)hi << 32) | lo);
}

// merge_outer_h2 on an h2 instance: the keys of vertices v0 + 2 * lane and
// v0 + 2 * lane + 1 (lanes past np read the last pair).
__device__ __forceinline__ void merge_cells_h2(const int32_t *__restrict__ f, const GMerge &M,
                                               int64_t v0, uint32_t X, uint32_t o0, uint32_t o1,
                                               uint32_t &klo, uint32_t &khi) {
    const uint32_t lb = 4u * (uint32_t)min((int64_t)__lane_id(), ((int64_t)M.np - v0) / 2 - 1);
    merge_outer_h2(reinterpret_cast<const char *>(f + M.tbase) + 2 * v0, lb, 2 * (int64_t)M.np, X,
                   o0, o1, klo, khi);
}

template <typename P, int XG, typename DT>
__global__ __launch_bounds__(kBlock) void k_gmerge(const int32_t *__restrict__ f,
                                                   const GMerge *__restrict__ mg,
                                                   const int64_t *__restrict__ item_pre, int nmg,
                                                   const uint32_t *__restrict__ xs,
                                                   void *__restrict__ D,
                                                   P *__restrict__ CODE, int nxcd,
                                                   int64_t per_xcd) {
    const int lane = threadIdx.x & (kWave - 1);
    // items dealt to XCDs in contiguous ranges: the waves of an XCD sweep the
    // rows of one tile together, and its lower-layer rows stay in that L2
    if ((int64_t)(blockIdx.x / nxcd) >= per_xcd) return;
    __shared__ int64_t sh_pre[kSegCache];
    const int64_t *pre = cache_pre(item_pre, nmg, sh_pre);
    const int64_t item = xcd_item(nxcd, per_xcd) * (kBlock / kWave) +
                         __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    if (item >= pre[nmg]) return;
    const int s = __builtin_amdgcn_readfirstlane(find_seg(pre, nmg, item));
    const GMerge M = mg[s];
    const int64_t local = wave_uniform(item - pre[s]);
    const int64_t tile = local / M.nsub;
    const int sub = (int)(local - tile * M.nsub);
    const int k = M.k;
    const int nrows = M.nrows;
    if (M.h2) {
        const int64_t v = tile * (2 * kWave) + 2 * lane;
        uint32_t klo[XG], khi[XG];
#pragma unroll
        for (int jx = 0; jx < XG; jx++) {
            const int xi = sub * XG + jx;
            klo[jx] = khi[jx] = 0xFFFFFFFFu;
            if (xi < nrows) {
                const uint32_t X = xs[M.x0 + xi];
                merge_cells_h2(f, M, tile * (2 * kWave), X, 0u, merge_steps(X), klo[jx], khi[jx]);
            }
        }
        if (v >= M.np) return;
        const int chunk = sub / (kWave / XG);
        const int loff = (sub % (kWave / XG)) * XG;
        const int64_t dbase = M.dbase0 + (int64_t)chunk * M.np * kWave;
        const int ls = merge_lstride(M);
        DT *dd = chunk_labels<DT>(D, merge_lbase(M, chunk)) + v * ls + loff;
        P *cd = CODE + dbase + v * kWave + loff;
        const DT INF = (DT)~(DT)0;
#pragma unroll
        for (int jx = 0; jx < XG; jx++) {
            dd[jx] = klo[jx] == 0xFFFFFFFFu ? INF : (DT)(klo[jx] >> 16);
            cd[jx] = key_packed<P>(klo[jx], k);
            dd[ls + jx] = khi[jx] == 0xFFFFFFFFu ? INF : (DT)(khi[jx] >> 16);
            cd[kWave + jx] = key_packed<P>(khi[jx], k);
        }
        return;
    }
    const int64_t v = tile * kWave + lane;
    DT dv[XG];
    P pv[XG];
#pragma unroll
    for (int jx = 0; jx < XG; jx++) {
        const int xi = sub * XG + jx;
        dv[jx] = (DT)~(DT)0;
        pv[jx] = ~(P)0;
        if (xi < nrows) {
            const uint32_t X = xs[M.x0 + xi];
            const P acc = merge_cells<P>(f, M, v, X, 0u, merge_steps(X));
            dv[jx] = (DT)(acc >> k);
            pv[jx] = acc;
        }
    }
    const int chunk = sub / (kWave / XG);
    const int loff = (sub % (kWave / XG)) * XG;
    const int64_t dbase = M.dbase0 + (int64_t)chunk * M.np * kWave;
    DT *dd = chunk_labels<DT>(D, merge_lbase(M, chunk)) + v * merge_lstride(M) + loff;
    P *cd = CODE + dbase + v * kWave + loff;
#pragma unroll
    for (int jx = 0; jx < XG; jx++) {
        dd[jx] = dv[jx];
        cd[jx] = pv[jx];
    }
}

// Split merge of the instances whose rows have many pairs (the upper
// layers): as k_merge_split, item = (tile, row, part) numbered tile-major per
// instance, each part folded into the row-major accumulator with an
// unsigned atomicMin (the minimum does not depend on the order).  Macc starts
// at all ones; k_gpack moves it to the staging.
template <typename P>
__global__ __launch_bounds__(kBlock) void k_gmerge_split(
    const int32_t *__restrict__ f, const GMerge *__restrict__ mg,
    const int64_t *__restrict__ item_pre, int nmg, const uint32_t *__restrict__ xs,
    P *__restrict__ Macc, int nxcd, int64_t per_xcd) {
    const int lane = threadIdx.x & (kWave - 1);
    if ((int64_t)(blockIdx.x / nxcd) >= per_xcd) return;
    __shared__ int64_t sh_pre[kSegCache];
    const int64_t *pre = cache_pre(item_pre, nmg, sh_pre);
    const int64_t item = xcd_item(nxcd, per_xcd) * (kBlock / kWave) +
                         __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    if (item >= pre[nmg]) return;
    const int s = __builtin_amdgcn_readfirstlane(find_seg(pre, nmg, item));
    const GMerge M = mg[s];
    const int64_t local = wave_uniform(item - pre[s]);
    const int64_t per_tile = (int64_t)M.nrows * M.S;
    const int64_t tile = local / per_tile;
    const int rem = (int)(local - tile * per_tile);
    const int xi = rem / M.S;
    const int part = rem - xi * M.S;
    const uint32_t X = xs[M.x0 + xi];
    const uint32_t steps = merge_steps(X);
    const uint32_t o0 = (uint32_t)((uint64_t)steps * part / M.S);
    const uint32_t o1 = (uint32_t)((uint64_t)steps * (part + 1) / M.S);
    if (o0 >= o1) return;
    if (M.h2) {
        const int64_t v = tile * (2 * kWave) + 2 * lane;
        uint32_t klo = 0xFFFFFFFFu, khi = 0xFFFFFFFFu;
        merge_cells_h2(f, M, tile * (2 * kWave), X, o0, o1, klo, khi);
        if (v >= M.np) return;
        P *acc = &Macc[M.mbase + (int64_t)xi * M.np + v];
        atomicMin(acc, key_packed<P>(klo, M.k));
        atomicMin(acc + 1, key_packed<P>(khi, M.k));
        return;
    }
    const int64_t v = tile * kWave + lane;
    const P acc = merge_cells<P>(f, M, v, X, o0, o1);
    atomicMin(&Macc[M.mbase + (int64_t)xi * M.np + v], acc);
}

// Star merges of the 16-bit instances whose layer rows have many pairs: the
// stars of k_merge_star_h2 per instance (its stars are stars[M.star0 ..
// M.star0 + M.nstars), row indices local to the instance's layer), item =
// (tile, star, part) numbered tile-major per instance, M.S parts per star,
// each folded into the instance's accumulators with an unsigned atomicMin
// (k_gpack moves them to the staging, as for the split merges).
template <typename P>
__global__ __launch_bounds__(kBlock) void k_gmerge_star(
    const int32_t *__restrict__ f, const GMerge *__restrict__ mg,
    const int64_t *__restrict__ item_pre, int nmg, const MergeStar *__restrict__ stars,
    P *__restrict__ Macc, int nxcd, int64_t per_xcd) {
    const int lane = threadIdx.x & (kWave - 1);
    if ((int64_t)(blockIdx.x / nxcd) >= per_xcd) return;
    __shared__ int64_t sh_pre[kSegCache];
    const int64_t *pre = cache_pre(item_pre, nmg, sh_pre);
    const int64_t item = xcd_item(nxcd, per_xcd) * (kBlock / kWave) +
                         __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    if (item >= pre[nmg]) return;
    const int s = __builtin_amdgcn_readfirstlane(find_seg(pre, nmg, item));
    const GMerge M = mg[s];
    const int64_t local = wave_uniform(item - pre[s]);
    const int64_t per_tile = (int64_t)M.nstars * M.S;
    const int64_t tile = local / per_tile;
    const int rem = (int)(local - tile * per_tile);
    const int g = rem / M.S;
    const int part = rem - g * M.S;
    const MergeStar *st = stars + M.star0 + g;
    const uint32_t S = st->S;
    const int r = (int)st->r;
    uint32_t a[kStarMax];
#pragma unroll
    for (int i = 0; i < kStarMax; i++) a[i] = st->a[i];
    const uint32_t steps = star_steps(S);
    const uint32_t o0 = (uint32_t)((uint64_t)steps * part / M.S);
    const uint32_t o1 = (uint32_t)((uint64_t)steps * (part + 1) / M.S);
    if (o0 >= o1) return;
    const int64_t v0 = tile * (2 * kWave);
    const int64_t v = v0 + 2 * lane;
    // lanes past np read the last pair; their results are dropped
    const uint32_t lb = 4u * (uint32_t)min((int64_t)lane, ((int64_t)M.np - v0) / 2 - 1);
    const char *fb = reinterpret_cast<const char *>(f + M.tbase) + 2 * v0;
    uint32_t klo[kStarMax], khi[kStarMax];
    star_h2_pairs(fb, lb, 2 * (int64_t)M.np, S, r, a, o0, o1, klo, khi);
    if (v >= M.np) return;
#pragma unroll
    for (int i = 0; i < kStarMax; i++) {
        if (i < r) {
            P *acc = &Macc[M.mbase + (int64_t)st->xi[i] * M.np + v];
            atomicMin(acc, key_packed<P>(klo[i], M.k));
            atomicMin(acc + 1, key_packed<P>(khi[i], M.k));
        }
    }
}

// Accumulators of the split instances -> vertex-major staging (label, packed
// word) through an LDS transpose; workgroup = (instance, chunk, 64-vertex
// tile) numbered by blk_pre.  Lanes past the instance's rows are INF.
template <typename P, typename DT>
__global__ __launch_bounds__(kBlock) void k_gpack(const P *__restrict__ Macc,
                                                  const GMerge *__restrict__ mg,
                                                  const int64_t *__restrict__ blk_pre, int nmg,
                                                  void *__restrict__ D,
                                                  P *__restrict__ CODE) {
    __shared__ P sm[kWave][kWave + 1];
    const int64_t b = blockIdx.x;
    if (b >= blk_pre[nmg]) return;
    const int s = find_seg(blk_pre, nmg, b);
    const GMerge M = mg[s];
    const int64_t local = b - blk_pre[s];
    const int64_t chunk = local / M.ntiles;
    const int64_t v0 = (local - chunk * M.ntiles) * kWave;
    const int lane = threadIdx.x & (kWave - 1);
    const int grp = threadIdx.x / kWave;
#pragma unroll
    for (int r = 0; r < kWave / (kBlock / kWave); r++) {
        const int xl = grp + r * (kBlock / kWave);
        const int64_t xi = chunk * kWave + xl;
        sm[xl][lane] = (xi < M.nrows) ? Macc[M.mbase + xi * M.np + v0 + lane] : ~(P)0;
    }
    __syncthreads();
    const int64_t dbase = M.dbase0 + chunk * M.np * kWave;
    DT *Dd = chunk_labels<DT>(D, merge_lbase(M, chunk));
    const int ls = merge_lstride(M);
    const bool real = chunk * kWave + lane < M.nrows;
#pragma unroll
    for (int r = 0; r < kWave / (kBlock / kWave); r++) {
        const int vv = grp + r * (kBlock / kWave);
        const P a = sm[lane][vv];
        const int64_t idx = (v0 + vv) * kWave + lane;
        Dd[(v0 + vv) * ls + lane] = real ? (DT)(a >> M.k) : (DT)~(DT)0;
        CODE[dbase + idx] = real ? a : ~(P)0;
    }
}

// All pairs of row X at the instance's table cells (16 or 32 bits), from
// parts o0 .. o1 of the outer steps.
template <typename P>
__device__ __forceinline__ P merge_cells(const int32_t *__restrict__ f, const GMerge &M, int64_t v,
                                         uint32_t X, uint32_t o0, uint32_t o1) {
    if (M.tbits == 16)
        return merge_outer<uint16_t, P>(reinterpret_cast<const uint16_t *>(f + M.tbase) + v,
                                        (int64_t)M.np, X, M.k, o0, o1);
    return merge_outer<int32_t, P>(f + M.tbase + v, (int64_t)M.np, X, M.k, o0, o1);
}

// Singleton rows of instances at their first layer, from their int32 copies
// at the end of the cost region to rows 2^t at the instance's cell width, and
// row 0 (INF).  Workgroup = (instance, row r): r < k moves singleton row r,
// r == k fills row 0.  For k >= 3 the copies lie above every place written
// here, at either width, so the workgroups need no order.
struct GPlace {
    int64_t tbase;
    int32_t np, k, tbits, pad_;
};

__global__ __launch_bounds__(kBlock) void k_gplace(int32_t *__restrict__ f,
                                                   const GPlace *__restrict__ pl,
                                                   const int64_t *__restrict__ row_pre, int npl) {
    __shared__ int64_t sh_pre[kSegCache];
    const int64_t *pre = cache_pre(row_pre, npl, sh_pre);
    const int64_t blk = blockIdx.x;
    if (blk >= pre[npl]) return;
    const int s = find_seg(pre, npl, blk);
    const GPlace G = pl[s];
    const int r = (int)(blk - pre[s]);
    const int64_t S2 = 1ll << G.k;
    int32_t *fr = f + G.tbase;
    const int32_t *copy = fr + (S2 - G.k + r) * G.np;
    const int64_t X = r < G.k ? (1ll << r) : 0;
    for (int64_t v = threadIdx.x; v < G.np; v += blockDim.x) {
        const int32_t x = r < G.k ? copy[v] : INT32_MAX;
        if (G.tbits == 16)
            reinterpret_cast<uint16_t *>(fr)[X * G.np + v] =
                (x == INT32_MAX) ? (uint16_t)0xFFFF : (uint16_t)x;
        else
            fr[X * G.np + v] = x;
    }
}

// ------------------------------------------------------------ lane packing
// Partial chunks of instances on one graph share relaxation chunks: a packed
// chunk ("bin") at dbase takes, in lane j, row src[j].lane of the staging
// chunk at src[j].dbase (src[j].dbase < 0: an unused lane, INF).  Rows relax
// independently over the same graph, so a bin relaxes exactly as its source
// chunks would.  k_glanes<DT, true> copies the rows in before the rounds,
// k_glanes<DT, false> copies them back after.  Workgroup = (bin, 64-vertex
// tile), numbered by blk_pre.
struct GLane {
    int64_t dbase;
    int32_t lane;
    int32_t pad_;
};

struct GBin {
    int64_t dbase;  // the bin's staging cells
    int32_t np;
    int32_t map0;   // its 64 GLane entries
};

template <typename DT, bool IN>
__global__ __launch_bounds__(kBlock) void k_glanes(void *__restrict__ D,
                                                   const GBin *__restrict__ bins,
                                                   const int64_t *__restrict__ blk_pre, int nbin,
                                                   const GLane *__restrict__ lanes) {
    const int64_t b = blockIdx.x;
    if (b >= blk_pre[nbin]) return;
    const int s = find_seg(blk_pre, nbin, b);
    const GBin B = bins[s];
    const int64_t v0 = (b - blk_pre[s]) * kWave;
    const int lane = threadIdx.x & (kWave - 1);
    const int grp = threadIdx.x / kWave;
    const GLane src = lanes[B.map0 + lane];
    DT *Db = chunk_labels<DT>(D, B.dbase);
    DT *Ds = src.dbase >= 0 ? chunk_labels<DT>(D, src.dbase) + src.lane : nullptr;
#pragma unroll 4
    for (int r = 0; r < kWave / (kBlock / kWave); r++) {
        const int64_t v = v0 + grp + r * (kBlock / kWave);
        if constexpr (IN)
            Db[v * kWave + lane] = Ds ? Ds[v * kWave] : (DT)~(DT)0;
        else if (Ds)
            Ds[v * kWave] = Db[v * kWave + lane];
    }
}

// Flags neighbour nb (< 0: none) of a changed item of chunk c and appends
// its item id (item_pre of the chunk + nb) while the list has room.
__device__ __forceinline__ void gflag(const GChunk &c, int64_t base_item, int nb,
                                      uint32_t *__restrict__ Fnext, uint32_t *__restrict__ Qnext,
                                      uint32_t *__restrict__ cnext, uint32_t qcap, bool list_open) {
    if (!list_open) {
        // flag scan only: a plain store, no read of the flag (list_open is
        // wave-uniform)
        if (nb >= 0) Fnext[c.fbase + nb] = 1;
        return;
    }
    bool want = false;
    uint32_t ent = 0;
    if (nb >= 0) {
        uint32_t *fp = Fnext + c.fbase + nb;
        if (*fp == 0) {
            *fp = 1;
            want = true;
            ent = (uint32_t)(base_item + nb);
        }
    }
    wave_append(want, ent, Qnext, cnext, qcap);
}

// ------------------------------------------------------------ relaxation
// The pull round of k_relax_v (worklist or flag scan, same rules) over the
// items of many chunks: item i of chunk g is vertex i - item_pre[g] of that
// chunk's instance; blk_chunk[b] names the chunk of item 64 * b, so a wave
// finds its items' chunks with one load.  Adjacency is the shared padded fixed-degree copy
// (DMAX > 0, neighbour ids local to the instance's graph) or the shared CSR.
template <typename DT, int DMAX, bool PK = false>
__global__ __launch_bounds__(kBlock) void k_grelax(
    void *__restrict__ D, const GChunk *__restrict__ ch,
    const int64_t *__restrict__ item_pre, const int32_t *__restrict__ blk_chunk, int nch, const int32_t *__restrict__ rp,
    const Edge *__restrict__ adj, const Edge *__restrict__ ell, int dense0,
    uint32_t *__restrict__ Fcur, uint32_t *__restrict__ Fnext, const uint32_t *__restrict__ Qcur,
    uint32_t *__restrict__ Qnext, uint32_t qcap, uint32_t *__restrict__ any,
    uint32_t *__restrict__ cnt, int r, int dense_rounds) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        any[(r + 2) & 3] = 0;
        cnt[(r + 2) & 3] = 0;
    }
    // dense stretch, as k_relax_v: rounds 1 .. dense_rounds take every item,
    // rounds 0 .. dense_rounds-1 write no flags
    const bool dense = dense0 || (r >= 1 && r <= dense_rounds);
    const bool noflags = r < dense_rounds;
    if (!dense && any[r & 3] == 0) return;
    const uint32_t qn = dense ? 0u : cnt[r & 3];
    const bool use_list = !dense && qn < qcap;
    uint32_t *cnext = &cnt[(r + 1) & 3];
    const int lane = threadIdx.x & (kWave - 1);
    const int64_t wave0 = (int64_t)blockIdx.x * (kBlock / kWave) + threadIdx.x / kWave;
    const int64_t wstride = (int64_t)gridDim.x * (kBlock / kWave);
    const int64_t nitems = item_pre[nch];
    const int64_t span = use_list ? (int64_t)qn : nitems;
    const int64_t nblk = (span + kWave - 1) / kWave;
    const DT INF = (DT)~(DT)0;
    bool raised = false;
    for (int64_t blk = wave0; blk < nblk; blk += wstride) {
        const int64_t j = blk * kWave + lane;
        int64_t mi = -1;
        if (j < span) mi = use_list ? (int64_t)Qcur[j] : j;
        int mg = 0, mv = 0;
        bool fl = false;
        if (mi >= 0) {
            if (use_list) {
                mg = find_seg(item_pre, nch, mi);
            } else {
                // the block's chunk from the table; a chunk boundary inside
                // the block moves the lanes past it on
                mg = blk_chunk[blk];
                while (item_pre[mg + 1] <= mi) mg++;
            }
            mv = (int)(mi - item_pre[mg]);
            const int64_t mf = ch[mg].fbase + mv;
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
        const bool list_open = *(volatile uint32_t *)cnext < qcap;
        int me0 = 0, me1 = 0;
        Edge e[DMAX > 0 ? DMAX : 1];
        if constexpr (DMAX > 0) {
#pragma unroll
            for (int q = 0; q < DMAX; q++) e[q] = Edge{-1, 0};
            if (fl) {
                const int64_t gv = (int64_t)ch[mg].voff + mv;
#pragma unroll
                for (int q = 0; q < DMAX; q++) e[q] = ell[gv * DMAX + q];
            }
        } else {
            if (fl) {
                const int64_t gv = (int64_t)ch[mg].voff + mv;
                me0 = rp[gv];
                me1 = rp[gv + 1];
            }
        }
        if constexpr (DMAX > 0) {
            // two items per step: 2 * DMAX neighbour gathers in flight
            while (act) {
                const int t1 = __ffsll((unsigned long long)act) - 1;
                act &= act - 1;
                const bool two = act != 0;
                const int t2 = two ? __ffsll((unsigned long long)act) - 1 : t1;
                if (two) act &= act - 1;
                const int ga = __builtin_amdgcn_readlane(mg, t1);
                const int gb = __builtin_amdgcn_readlane(mg, t2);
                const int va = __builtin_amdgcn_readlane(mv, t1);
                const int vb = __builtin_amdgcn_readlane(mv, t2);
                const GChunk c1 = ch[ga];
                const GChunk c2 = ch[gb];
                DT *Da = chunk_labels<DT>(D, c1.lbase);
                DT *Db = chunk_labels<DT>(D, c2.lbase);
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
                if (noflags) {
                    if (__ballot(chga | chgb) != 0) raised = true;
                } else if (__ballot(chga | chgb) != 0) {
                    // flag only the neighbours that a dropped row now
                    // improves (as k_relax_v); lane q < DMAX holds neighbour q
                    int na = -1, nb = -1;
#pragma unroll
                    for (int q = 0; q < DMAX; q++) {
                        const bool ia2 = chga && lbl_improves<DT, PK>(besta, cura, wa[q], da[q]);
                        const bool ib2 = chgb && lbl_improves<DT, PK>(bestb, curb, wb[q], db[q]);
                        const bool fa = __ballot(ia2) != 0, fb = __ballot(ib2) != 0;
                        if (lane == q) {
                            na = fa ? ua[q] : -1;
                            nb = fb ? ub[q] : -1;
                        }
                    }
                    const bool anya = __ballot(na >= 0) != 0, anyb = __ballot(nb >= 0) != 0;
                    if (anya) gflag(c1, item_pre[ga], na, Fnext, Qnext, cnext, qcap, list_open);
                    if (anyb) gflag(c2, item_pre[gb], nb, Fnext, Qnext, cnext, qcap, list_open);
                    if (anya || anyb) raised = true;
                }
            }
        } else {
            while (act) {
                const int t = __ffsll((unsigned long long)act) - 1;
                act &= act - 1;
                const int g = __builtin_amdgcn_readlane(mg, t);
                const int v = __builtin_amdgcn_readlane(mv, t);
                const GChunk c = ch[g];
                DT *Dc = chunk_labels<DT>(D, c.lbase);
                const int64_t idx = (int64_t)v * kWave + lane;
                const DT cur = Dc[idx];
                DT best = cur;
                const int e0 = __builtin_amdgcn_readlane(me0, t);
                const int e1 = __builtin_amdgcn_readlane(me1, t);
                for (int eb = e0; eb < e1; eb += kWave) {
                    const int cnt_e = min(kWave, e1 - eb);
                    Edge my = {0, 0};
                    if (lane < cnt_e) my = adj[eb + lane];
                    for (int tt = 0; tt < cnt_e; tt++) {
                        const int u = __builtin_amdgcn_readlane(my.v, tt);
                        const int w = __builtin_amdgcn_readlane(my.w, tt);
                        best = relax_lbl<DT, PK>(Dc[(int64_t)u * kWave + lane], u, w, INF, best);
                    }
                }
                const bool dchg = lbl_dropped<DT, PK>(best, cur);
                if (dchg) Dc[idx] = best;
                if (__ballot(dchg) != 0) {
                    if (!noflags) {
                        const int64_t base_item = item_pre[g];
                        for (int eb = e0; eb < e1; eb += kWave) {
                            const int e2 = eb + lane;
                            gflag(c, base_item, e2 < e1 ? adj[e2].v : -1, Fnext, Qnext, cnext,
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

// The block-local round of k_relax_blk over the blocks of many chunks: item i
// of chunk g (item_pre[g] <= i < item_pre[g+1]) is block i - item_pre[g] of
// that chunk's graph, whose slots start at boff[b0 + block].  The block
// arrays of all graphs share one slot and adjacency numbering (bvert,
// brp/bext, badj); vblk and the outside entries of badj hold vertex ids local
// to the graph.  The flags of item i are Fcur / Fnext[F0 + i], chunk g's
// from F0 + item_pre[g] = its fbase.  Dm / bdoff: the in-block distance
// matrices of all block graphs (Dm null: sweeps only).
template <typename DT, typename DW>
__global__ __launch_bounds__(kBT) __attribute__((amdgpu_waves_per_eu(kBlkWaves<DT>()))) void k_gblk(
    void *__restrict__ D, const GChunk *__restrict__ ch, const int64_t *__restrict__ item_pre,
    int nch, const int32_t *__restrict__ boff, const int32_t *__restrict__ bvert,
    const int32_t *__restrict__ brp, const int32_t *__restrict__ bext,
    const Edge *__restrict__ badj, const int32_t *__restrict__ vblk, const DW *__restrict__ Dm,
    const int64_t *__restrict__ bdoff, int64_t F0, int dense, uint32_t *__restrict__ Fcur,
    uint32_t *__restrict__ Fnext, uint32_t *__restrict__ any, int r) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    DT *L = reinterpret_cast<DT *>(lds_raw);  // L[slot * 64 + row]
    uint32_t *s_misc = reinterpret_cast<uint32_t *>(lds_raw + sizeof(DT) * kBV * kWave);
    uint32_t *s_flag = s_misc;     // 3 rotating "sweep lowered something" words
    uint32_t *s_cnt = s_misc + 4;  // flagged items of the current piece
    uint8_t *s_list = reinterpret_cast<uint8_t *>(s_misc + 8);        // 256 entries
    uint8_t *s_dirty = reinterpret_cast<uint8_t *>(s_misc + 8 + 64);  // 2 x kBV marks
    uint64_t *s_chg = reinterpret_cast<uint64_t *>(s_misc + kBMiscChg);  // kBW masks
    const int tid = threadIdx.x;
    const int lane = tid & (kWave - 1);
    if (blockIdx.x == 0 && tid == 0) any[(r + 2) & 3] = 0;
    if (!dense && any[r & 3] == 0) return;
    const int64_t nitems = item_pre[nch];
    const int64_t G = gridDim.x;
    const int64_t mine = (nitems > (int64_t)blockIdx.x) ? (nitems - 1 - blockIdx.x) / G + 1 : 0;
    for (int i = tid; i < 2 * kBV; i += kBT) s_dirty[i] = 0;
    bool raised = false;
    for (int64_t p0 = 0; p0 < mine; p0 += kBPiece) {
        if (tid == 0) *s_cnt = 0;
        __syncthreads();
        if (tid < kBPiece && p0 + tid < mine) {
            const int64_t it = blockIdx.x + G * (p0 + tid);
            bool f = dense != 0;
            if (!dense && Fcur[F0 + it] != 0) {
                f = true;
                Fcur[F0 + it] = 0;
            }
            if (f) s_list[atomicAdd(s_cnt, 1u)] = (uint8_t)tid;
        }
        __syncthreads();
        const int nl = (int)*s_cnt;
        for (int qi = 0; qi < nl; qi++) {
            const int64_t item = blockIdx.x + G * (p0 + s_list[qi]);
            const int g = find_seg(item_pre, nch, item);
            const GChunk c = ch[g];
            const int b = c.b0 + (int)(item - item_pre[g]);
            blk_visit<DT, DW>(chunk_labels<DT>(D, c.lbase), boff[b], boff[b + 1] - boff[b],
                              bvert, brp, bext, badj, vblk + c.voff, Dm ? Dm + bdoff[b] : nullptr,
                              Fnext + c.fbase, dense, L, s_flag, s_dirty, s_chg, nullptr, raised);
            __syncthreads();  // LDS is reused by the next item
        }
        __syncthreads();  // s_cnt is reset by the next piece
    }
    if (raised && lane == 0) any[(r + 1) & 3] = 1;
}

// ----------------------------------------------------------------- unpack
// Staging chunks -> table rows of their instances (LDS transpose of 64
// vertices x 64 rows per workgroup).  Layer chunks decide each record from
// the converged labels (cell_record over the instance's adjacency).
// Singleton chunks write f[1<<t][v] and the record (kk[t], {t})
// (reader-el.c:2015-2027), and fold the per-row maximum into
// rowmax[inst * 64 + t].
template <typename DT, typename P>
__global__ __launch_bounds__(kBlock) void k_gunpack(
    const void *__restrict__ D, const P *__restrict__ CODE,
    const GChunk *__restrict__ ch, const int64_t *__restrict__ blk_pre, int nch,
    const uint32_t *__restrict__ xs, const int32_t *__restrict__ kk_all,
    const int32_t *__restrict__ kk_off, const int32_t *__restrict__ rp,
    const Edge *__restrict__ adj, int32_t *__restrict__ f, int32_t *__restrict__ b,
    unsigned long long *__restrict__ rowmax) {
    __shared__ DT sd[kWave][kWave + 1];  // label type: 16-bit labels halve this LDS
    __shared__ int32_t sc[kWave][kWave + 1];
    const int64_t item = blockIdx.x;
    if (item >= blk_pre[nch]) return;
    const int g = find_seg(blk_pre, nch, item);
    const GChunk c = ch[g];
    const int64_t v0 = (item - blk_pre[g]) * kWave;
    const int t = threadIdx.x;
    const int lane = t & (kWave - 1);
    const int grp = __builtin_amdgcn_readfirstlane(t / kWave);
    const DT *Dc = chunk_labels<DT>(const_cast<void *>(D), c.lbase);
    const int64_t ls = c.lstride;
    const bool real = lane < c.nx;
    // the labels and packed words of kPre vertices are loaded together, then
    // their records decided one vertex at a time
    constexpr int kPre = 4;
#pragma unroll 1
    for (int r0 = 0; r0 < kWave / (kBlock / kWave); r0 += kPre) {
        DT Ls[kPre];
        P Ps[kPre];
#pragma unroll
        for (int u = 0; u < kPre; u++) {
            const int64_t vu = v0 + grp + (r0 + u) * (kBlock / kWave);
            Ls[u] = Dc[vu * ls + lane];
            Ps[u] = c.kind == 0 ? CODE[c.dbase + vu * kWave + lane] : (P)0;
        }
#pragma unroll
        for (int u = 0; u < kPre; u++) {
            const int vv = grp + (r0 + u) * (kBlock / kWave);
            const int64_t v = v0 + vv;
            int32_t code = 0;
            if (c.kind == 0 && v < c.n)
                code = cell_record<DT, P>(Dc, Ls[u], Ps[u], c.k, real, c.voff + (int)v, rp, adj, ls);
            else if (c.kind == 2 && v < c.n)
                // group rows: {j} at the members (label 0: weights are >= 1),
                // elsewhere the edge from the smallest neighbour on a
                // shortest path (the merge word's sum is out of reach)
                code = cell_record<DT, uint64_t>(
                    Dc, Ls[u], Ls[u] == (DT)0 ? (uint64_t)(1ull << lane) : ~0ull, c.k, real,
                    c.voff + (int)v, rp, adj, ls);
            sd[vv][lane] = Ls[u];
            sc[vv][lane] = code;
        }
    }
    __syncthreads();
    const int64_t S2 = 1ll << c.k;
    const int64_t bbase = c.tbase + S2 * c.np;  // records follow the costs
#pragma unroll
    for (int r = 0; r < kWave / (kBlock / kWave); r++) {
        const int xl = grp + r * (kBlock / kWave);
        const int64_t v = v0 + lane;
        unsigned long long dmax = 0;
        if (xl < c.nx && v < c.n) {
            const DT dl = sd[lane][xl];
            const uint32_t d = (dl == (DT)~(DT)0) ? ~0u : (uint32_t)dl;
            uint16_t *f16 = reinterpret_cast<uint16_t *>(f + c.tbase);
            if (c.kind == 0) {
                const int64_t X = xs[c.x0 + xl];
                if (c.tbits == 16)
                    f16[X * c.np + v] = (uint16_t)d;
                else
                    f[c.tbase + X * c.np + v] = (int32_t)d;
                b[bbase + X * c.np + v] = sc[lane][xl];
            } else {
                const int64_t X = 1ll << xl;
                const int32_t src = kk_all[kk_off[c.inst] + xl];
                const int64_t row = c.k >= 3 ? S2 - c.k + xl : X;  // staged copy, or in place
                f[c.tbase + row * c.np + v] = (d == ~0u) ? INT32_MAX : (int32_t)d;
                if (c.kind == 1)
                    b[bbase + X * c.np + v] = (v == src) ? (int32_t)X : (int32_t)(S2 + src);
                else
                    b[bbase + X * c.np + v] = sc[lane][xl];
                dmax = d;
            }
        }
        if (c.kind != 0) {
            for (int off = 32; off > 0; off >>= 1) {
                const unsigned long long o = __shfl_xor(dmax, off);
                dmax = o > dmax ? o : dmax;
            }
            if (lane == 0 && xl < c.nx) atomicMax(&rowmax[(int64_t)c.inst * kWave + xl], dmax);
        }
    }
}

// Per-step convergence word for graph replays: out[slot] = any[r & 3].
__global__ void k_copy_word(const uint32_t *__restrict__ any, int r, uint32_t *__restrict__ out,
                            int slot) {
    if (threadIdx.x == 0 && blockIdx.x == 0) out[slot] = any[r & 3];
}

// Answer cells f[C][q] of finished instances.
__global__ void k_gather_costs(const int32_t *__restrict__ f, const int64_t *__restrict__ cell,
                               int n, int64_t *__restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        const int32_t x = f[cell[i]];
        out[i] = (x == INT32_MAX) ? INT64_MAX : (int64_t)x;
    }
}

}  // namespace emv
