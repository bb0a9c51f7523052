// emv_engine.hip -- multi-instance EMV engine: batched solves and serving.
//
// The engine keeps graphs resident (one shared adjacency), owns a device
// arena for the per-instance tables and advances every active instance by
// one DP stage per step (reader-el.c:1988-2134 per instance): the singleton
// shortest paths for newly admitted instances, the merge + relaxation of
// layer m for the others.  All instances of a step share one merge launch,
// one set of relaxation rounds and one unpack launch.
//
//   batch  : many instances, possibly on different graphs, admitted at once
//            and stepped in lockstep; steps 2..kmax can be captured in a
//            hipGraph with a fixed number of relaxation rounds per step,
//            checked after the replay (a step that had not converged makes
//            the solve run again eagerly).
//   server : one resident graph, requests admitted by their table footprint
//            2^k * np * 8 bytes against the arena as they arrive, requests at
//            different layers stepped together (continuous batching).
#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <deque>
#include <map>
#include <new>
#include <vector>

#include "emv_batch_kernels.hpp"
#include "emv_blocks.hpp"
#include "emv_certify.hpp"
#include "emv_engine.h"
#include "emv_hip.h"

using namespace emv;

namespace {

double now_ms() {
    return std::chrono::duration<double, std::milli>(
               std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

int bits_for(uint64_t x) {
    int b = 0;
    while (b < 64 && (x >> b) != 0) b++;
    return b;
}

int64_t row_stride(int64_t n) {
    int64_t np = (n + kWave - 1) / kWave * kWave;
    if ((np / kWave) % 2 == 0) np += kWave;
    return np;
}

template <typename T>
void dfree(T *&p) {
    if (p) (void)hipFree((void *)p);
    p = nullptr;
}

// First-fit allocator over [0, cap) cells with coalescing frees.
struct Arena {
    int64_t cap = 0;
    std::map<int64_t, int64_t> free_;  // offset -> length
    void reset(int64_t c) {
        cap = c;
        free_.clear();
        if (c > 0) free_[0] = c;
    }
    int64_t alloc(int64_t len) {
        len = (len + 255) / 256 * 256;
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second >= len) {
                const int64_t off = it->first, rest = it->second - len;
                free_.erase(it);
                if (rest > 0) free_[off + len] = rest;
                return off;
            }
        }
        return -1;
    }
    void release(int64_t off, int64_t len) {
        len = (len + 255) / 256 * 256;
        auto it = free_.emplace(off, len).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
    int64_t largest() const {
        int64_t m = 0;
        for (auto &kv : free_) m = std::max(m, kv.second);
        return m;
    }
};

struct GraphRec {
    int64_t voff, n, np, maxw;
    int64_t wsum;  // sum of the edge weights (a group query's M = 1 + wsum)
    int maxdeg;
    int32_t b0 = 0, nblk = 0;  // blocks of block-local relaxation (nblk 0: vertex-wise)
};

enum { ST_QUEUED = 0, ST_RUNNING = 1, ST_DONE = 2 };

struct Inst {
    int64_t ticket = 0;
    int graph = 0;
    int k = 0;
    std::vector<int32_t> kk;
    // group query: groups[j] = the members of group j (k = groups); the
    // singleton row {j} holds the distance to the nearest member
    std::vector<std::vector<int32_t>> groups;
    int64_t root = -1;          // traceback start: kk[k-1], or the nearest member of the last group
    int64_t reduced_cost = 0;   // group query: cost + k * M (the reduced instance's cost)
    uint32_t flags = 0;
    int state = ST_QUEUED;
    int stage = 1;  // next stage: 1 = singletons, m >= 2 = layer m
    int pbits = 32;
    int dbits = 32;  // relaxation label width of the layer stages (16 when the bound allows)
    int tbits = 32;  // table cell width from layer 2 on: 16 when the bound B < 0xFFFF
    int status = EMV_OK;
    int64_t abase = -1, acells = 0;  // arena region: f then b, S * np cells each
    int64_t cost = 0;
    std::vector<int64_t> pairs;
    double t_submit = 0, t_admit = 0, t_done = 0;
    int slot = -1;  // row of the kk table while running
};

}  // namespace

namespace {
struct StepPlan;
}  // namespace

struct emv_engine {
    int device = 0;
    hipStream_t stream = nullptr;
    int num_cu = 256;
    char err[512] = {0};
    // resident graphs
    std::vector<GraphRec> graphs;
    int32_t *d_rp = nullptr;
    Edge *d_adj = nullptr;
    Edge *d_ell = nullptr;
    int ell_deg = 0;
    int64_t gvert = 0;
    // blocks of the graphs that take block-local relaxation, in one numbering
    int32_t *d_boff = nullptr, *d_bvert = nullptr, *d_brp = nullptr, *d_bext = nullptr;
    int32_t *d_vblk = nullptr;  // gvert entries: block of each vertex within its graph
    Edge *d_badj = nullptr;
    // in-block distances from border slots, all block graphs (bdw 16 / 32 bits; 0: sweeps only)
    void *d_bdist = nullptr;
    int64_t *d_bdoff = nullptr;
    int bdw = 0;
    size_t lds_cu = 65536;      // LDS per CU
    // table arena (int32 cells)
    int32_t *d_arena = nullptr;
    Arena arena;
    // instances
    std::vector<Inst> inst;
    std::deque<int64_t> queue;  // tickets waiting for admission
    int64_t next_ticket = 0;
    int max_slots = 0;
    std::vector<int> free_slots;
    int32_t *d_kk = nullptr;     // max_slots * 64 terminals
    int32_t *d_kkoff = nullptr;  // slot -> offset (slot * 64)
    unsigned long long *d_rowmax = nullptr;  // max_slots * 64
    // staging
    uint32_t *d_D = nullptr;
    uint64_t *d_code = nullptr;  // merge's packed words (a step with 32-bit words uses half)
    int64_t stage_cells = 0;
    uint32_t *d_flag[2] = {nullptr, nullptr};
    int64_t flag_bytes = 0;  // flag words held per array
    void *d_macc = nullptr;  // split-merge accumulators (packed words)
    int64_t macc_bytes = 0;
    bool no_split = false;   // EMV_NO_SPLIT: every layer through k_gmerge
    bool no_pack = false;    // EMV_NO_PACK: every staging chunk relaxes in place
    bool pair_chunks = false;  // batched solves: 16-bit vertex-wise chunks relax in pairs
    bool star_merge = true;    // split 16-bit layers merge in stars (EMV_MERGE_STAR=0: by rows)
    uint32_t star_pairs = 63;  // and 16-bit rows of this many pairs or more (EMV_STAR_PAIRS)
    int split_items = 32;    // split merge: items per tile aimed at (EMV_SPLIT_ITEMS)
    uint32_t *d_any = nullptr;
    uint32_t *d_cnt = nullptr;   // 4 rotating worklist counters
    uint32_t *d_q[2] = {nullptr, nullptr};
    uint32_t qcap = 0;  // worklist rounds off unless EMV_QCAP sets a capacity
    uint32_t *h_pin = nullptr;
    // per-step descriptor buffer (eager steps reuse one; graph steps own one each)
    std::vector<void *> desc_bufs;
    std::vector<int64_t> desc_caps;
    // the last batch replay, kept while its key (descriptors, rounds, buffers,
    // launch shape) repeats, so that a repeated solve relaunches it
    hipGraph_t ggraph = nullptr;
    hipGraphExec_t gexec = nullptr;
    std::vector<char> gkey;
    uint32_t *d_conv = nullptr;  // per replayed step: 1 if it ran out of rounds
    size_t conv_cap = 0;
    std::vector<StepPlan> gplans;     // plans of the replayed steps, and the instance
    std::vector<int64_t> gplan_sig;   // states and settings they were built from
    // tunables
    emv_engine_params prm;
    // counters
    int64_t steps = 0, relax_rounds = 0;
    double relax_ms_acc = 0, merge_ms_acc = 0;
    std::vector<int> rounds_per_step;  // eager: rounds used by step s (batch)
    bool server = false;         // created by emv_server_create (requests may be cancelled)
    int64_t fail_alloc = -1;     // test hook: step allocations left before one fails (-1: off)
};

namespace {

void eset(emv_engine *e, const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(e->err, sizeof(e->err), fmt, ap);
    va_end(ap);
}

#define EHIP(e, expr)                                                                      \
    do {                                                                                   \
        hipError_t _x = (expr);                                                            \
        if (_x != hipSuccess) {                                                            \
            eset((e), "%s failed: %s (%s:%d)", #expr, hipGetErrorString(_x), __FILE__,     \
                 __LINE__);                                                                \
            return (_x == hipErrorOutOfMemory || _x == hipErrorMemoryAllocation)           \
                       ? EMV_ERR_NOMEM                                                     \
                       : EMV_ERR_HIP;                                                      \
        }                                                                                  \
    } while (0)

#define ECHK(expr)                   \
    do {                             \
        int _r = (expr);             \
        if (_r != EMV_OK) return _r; \
    } while (0)

// Device allocation of a step's scratch (staging, flags, merge accumulators,
// descriptors).  With the test hook armed (emv_engine_fail_alloc) the
// allocation it names fails as if the device were out of memory, without
// calling hipMalloc.  Every such allocation happens before the step launches
// anything, so a failed step leaves each request at its stage.
hipError_t step_malloc(emv_engine *e, void **p, size_t bytes) {
    if (e->fail_alloc == 0) {
        e->fail_alloc = -1;
        *p = nullptr;
        return hipErrorOutOfMemory;
    }
    if (e->fail_alloc > 0) e->fail_alloc--;
    return hipMalloc(p, bytes);
}

// The m-subsets of k bits in increasing numeric order (reader-el.c:2040-2046).
void subsets_of_size(int k, int m, std::vector<uint32_t> &out) {
    for (uint64_t X = (1ull << m) - 1; X < (1ull << k);) {
        out.push_back((uint32_t)X);
        const uint64_t lo = X & (~X + 1);
        const uint64_t r = X + lo;
        X = (((r ^ X) >> 2) / lo) | r;
    }
}

template <typename T>
int upload_vec(emv_engine *e, T *&d, const std::vector<T> &h) {
    EHIP(e, hipMalloc((void **)&d, sizeof(T) * std::max<size_t>(1, h.size())));
    if (!h.empty()) EHIP(e, hipMemcpy(d, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice));
    return EMV_OK;
}

// Cuts the graphs that suit block-local relaxation into blocks (the rule of
// the single-instance path) and uploads the blocks of all of them in one
// numbering: graph g's blocks are b0 .. b0+nblk-1, their slots follow on
// from the previous graph's, brp / bext index the shared badj.
int upload_blocks(emv_engine *e, const std::vector<int32_t> &rp, const std::vector<Edge> &adj) {
    const int mode = blocks_mode();
    if (mode == 0) return EMV_OK;
    std::vector<int32_t> boff(1, 0), bvert, brp, bext, vblk((size_t)e->gvert, -1);
    std::vector<Edge> badj;
    std::vector<int64_t> bdoff(1, 0);
    std::vector<uint64_t> bdist;
    uint64_t dmax = 0;
    const bool closure = closure_mode() != 0;
    int32_t nb = 0;
    for (GraphRec &G : e->graphs) {
        if (G.n < 2 || rp[G.voff + G.n] == rp[G.voff]) continue;
        BlockPart P;
        const double inside = build_blocks(G.n, rp.data() + G.voff, adj.data(), P);
        if (!blocks_wanted(mode, G.n, inside)) continue;
        const int32_t slot0 = (int32_t)bvert.size(), e0 = (int32_t)badj.size();
        G.b0 = nb;
        G.nblk = P.nblk;
        for (int i = 1; i <= P.nblk; i++) boff.push_back(slot0 + P.boff[i]);
        bvert.insert(bvert.end(), P.bvert.begin(), P.bvert.end());
        for (int64_t s = 0; s < G.n; s++) {
            brp.push_back(e0 + P.brp[s]);
            bext.push_back(e0 + P.bext[s]);
        }
        badj.insert(badj.end(), P.badj.begin(), P.badj.end());
        for (int64_t u = 0; u < G.n; u++) vblk[(size_t)(G.voff + u)] = P.vblk[u];
        nb += P.nblk;
        if (closure) {
            BlockDist B;
            block_distances(P, B);
            const int64_t d0 = (int64_t)bdist.size();
            for (int i = 1; i <= P.nblk; i++) bdoff.push_back(d0 + B.bdoff[i]);
            bdist.insert(bdist.end(), B.d.begin(), B.d.end());
            dmax = std::max(dmax, B.dmax);
        }
    }
    if (nb == 0) return EMV_OK;
    brp.push_back((int32_t)badj.size());
    e->bdw = closure ? dist_width(dmax) : 0;
    if (e->bdw == 16) {
        uint16_t *p = nullptr;
        ECHK(upload_vec(e, p, narrow<uint16_t>(bdist)));
        e->d_bdist = p;
    } else if (e->bdw == 32) {
        uint32_t *p = nullptr;
        ECHK(upload_vec(e, p, narrow<uint32_t>(bdist)));
        e->d_bdist = p;
    }
    if (e->bdw) ECHK(upload_vec(e, e->d_bdoff, bdoff));
    ECHK(upload_vec(e, e->d_boff, boff));
    ECHK(upload_vec(e, e->d_bvert, bvert));
    ECHK(upload_vec(e, e->d_brp, brp));
    ECHK(upload_vec(e, e->d_bext, bext));
    ECHK(upload_vec(e, e->d_vblk, vblk));
    ECHK(upload_vec(e, e->d_badj, badj));
    return EMV_OK;
}

int upload_graphs(emv_engine *e, const std::vector<std::vector<int64_t>> &edges,
                  const std::vector<int64_t> &ns) {
    const size_t G = ns.size();
    std::vector<int32_t> rp(1, 0);
    std::vector<Edge> adj;
    int64_t voff = 0;
    int maxdeg_all = 0;
    for (size_t g = 0; g < G; g++) {
        const int64_t n = ns[g];
        const std::vector<int64_t> &E = edges[g];
        const int64_t m = (int64_t)E.size() / 3;
        std::vector<int32_t> deg(n, 0);
        int64_t maxw = 0, wsum = 0;
        for (int64_t j = 0; j < m; j++) {
            const int64_t u = E[3 * j], v = E[3 * j + 1], w = E[3 * j + 2];
            if (u < 0 || v < 0 || u >= n || v >= n) {
                eset(e, "graph %zu: edge %lld endpoint out of range", g, (long long)j);
                return EMV_ERR_ARG;
            }
            if (w < 1 || w > INT32_MAX) {
                eset(e, "graph %zu: weight %lld outside [1, 2^31-1]", g, (long long)w);
                return EMV_ERR_RANGE;
            }
            maxw = std::max(maxw, w);
            wsum += w;
            deg[u]++;
            deg[v]++;
        }
        const int64_t e0 = (int64_t)adj.size();
        std::vector<int64_t> fill(n);
        int64_t run = e0;
        int maxdeg = 0;
        for (int64_t u = 0; u < n; u++) {
            fill[u] = run;
            run += deg[u];
            maxdeg = std::max(maxdeg, deg[u]);
            rp.push_back((int32_t)run);
        }
        if (run >= INT32_MAX) {
            eset(e, "adjacency exceeds 2^31 entries");
            return EMV_ERR_RANGE;
        }
        adj.resize(run);
        for (int64_t j = 0; j < m; j++) {
            const int32_t u = (int32_t)E[3 * j], v = (int32_t)E[3 * j + 1];
            const int32_t w = (int32_t)E[3 * j + 2];
            adj[fill[u]++] = Edge{v, w};
            adj[fill[v]++] = Edge{u, w};
        }
        GraphRec gr;
        gr.voff = voff;
        gr.n = n;
        gr.np = row_stride(n);
        gr.maxw = maxw;
        gr.wsum = wsum;
        gr.maxdeg = maxdeg;
        e->graphs.push_back(gr);
        voff += n;
        maxdeg_all = std::max(maxdeg_all, maxdeg);
    }
    e->gvert = voff;
    EHIP(e, hipMalloc((void **)&e->d_rp, sizeof(int32_t) * rp.size()));
    EHIP(e, hipMemcpy(e->d_rp, rp.data(), sizeof(int32_t) * rp.size(), hipMemcpyHostToDevice));
    EHIP(e, hipMalloc((void **)&e->d_adj, sizeof(Edge) * std::max<size_t>(adj.size(), 1)));
    if (!adj.empty())
        EHIP(e, hipMemcpy(e->d_adj, adj.data(), sizeof(Edge) * adj.size(), hipMemcpyHostToDevice));
    int ell = 0;
    for (int d : {4, 8, 16})
        if (maxdeg_all <= d) {
            ell = d;
            break;
        }
    if (ell && (int64_t)ell * voff > 2 * std::max<int64_t>((int64_t)adj.size(), 1)) ell = 0;
    if (getenv("EMV_NO_ELL")) ell = 0;
    if (ell) {
        std::vector<Edge> el((size_t)(voff * ell), Edge{-1, 0});
        for (int64_t gv = 0; gv < voff; gv++)
            for (int32_t i = rp[gv]; i < rp[gv + 1]; i++)
                el[(size_t)(gv * ell + (i - rp[gv]))] = adj[i];
        EHIP(e, hipMalloc((void **)&e->d_ell, sizeof(Edge) * el.size()));
        EHIP(e, hipMemcpy(e->d_ell, el.data(), sizeof(Edge) * el.size(), hipMemcpyHostToDevice));
    }
    e->ell_deg = ell;
    return upload_blocks(e, rp, adj);
}

int ensure_slots(emv_engine *e, int slots) {
    if (slots <= e->max_slots) return EMV_OK;
    // grows only while nothing is running (create time)
    dfree(e->d_kk);
    dfree(e->d_kkoff);
    dfree(e->d_rowmax);
    EHIP(e, hipMalloc((void **)&e->d_kk, sizeof(int32_t) * slots * kWave));
    EHIP(e, hipMalloc((void **)&e->d_kkoff, sizeof(int32_t) * slots));
    EHIP(e, hipMalloc((void **)&e->d_rowmax, sizeof(unsigned long long) * slots * kWave));
    std::vector<int32_t> off(slots);
    for (int s = 0; s < slots; s++) off[s] = s * kWave;
    EHIP(e, hipMemcpy(e->d_kkoff, off.data(), sizeof(int32_t) * slots, hipMemcpyHostToDevice));
    for (int s = e->max_slots; s < slots; s++) e->free_slots.push_back(s);
    e->max_slots = slots;
    return EMV_OK;
}

// Clears the item flags.
void clear_flags(emv_engine *e) {
    const int64_t c = e->flag_bytes;
    const int blocks = (int)std::min<int64_t>((c + kBlock - 1) / kBlock, 65536);
    for (int i = 0; i < 2; i++)
        hipLaunchKernelGGL(k_fill<uint32_t>, dim3(blocks), dim3(kBlock), 0, e->stream,
                           e->d_flag[i], c, 0u);
}

// Split-merge accumulators of a step: bytes = cells * (P bits / 8).
int ensure_macc(emv_engine *e, int64_t bytes) {
    if (bytes <= e->macc_bytes) return EMV_OK;
    dfree(e->d_macc);
    e->macc_bytes = 0;
    EHIP(e, step_malloc(e, &e->d_macc, (size_t)bytes));
    e->macc_bytes = bytes;
    return EMV_OK;
}

int ensure_stage(emv_engine *e, int64_t cells, int64_t flags) {
    if (cells > e->stage_cells) {
        dfree(e->d_D);
        dfree(e->d_code);
        e->stage_cells = 0;
        const int64_t c = cells + cells / 4;
        EHIP(e, step_malloc(e, (void **)&e->d_D, sizeof(uint32_t) * c));
        EHIP(e, step_malloc(e, (void **)&e->d_code, sizeof(uint64_t) * c));
        e->stage_cells = c;
    }
    if (flags > e->flag_bytes) {
        dfree(e->d_flag[0]);
        dfree(e->d_flag[1]);
        e->flag_bytes = 0;
        const int64_t c = flags + flags / 4;
        EHIP(e, step_malloc(e, (void **)&e->d_flag[0], sizeof(uint32_t) * c));
        EHIP(e, step_malloc(e, (void **)&e->d_flag[1], sizeof(uint32_t) * c));
        e->flag_bytes = c;
        clear_flags(e);
        EHIP(e, hipGetLastError());
    }
    return EMV_OK;
}

// Host-side plan of one step, packed into one descriptor upload.  The
// step's chunks form two groups by relaxation label width: group 0 holds the
// 32-bit chunks (every singleton stage, and the layer stages whose bound
// needs 32 bits), group 1 the 16-bit ones; each group's kernels get the
// group's slice of the chunk table and prefix arrays of their own.  Chunk
// dbase counts cells of 4 bytes in the staging; a 16-bit chunk uses the
// first half of its region.
struct StepGroup {
    bool paired = false;   // chunks relax in pairs of 128 rows (packed 16-bit labels)
    int ch0 = 0, nch = 0;  // chunk slice [ch0, ch0 + nch)
    int mg0 = 0, nmg = 0;  // merge slice (GMerge.chunk0 is relative to ch0)
    int ms0 = 0, nms = 0;  // split-merge slice, right after the merge slice: rows cut into
    int nss = 0;           // parts, then the last nss merging in stars
    std::vector<int64_t> star_pre;
    size_t off_stp = 0;
    int rch0 = 0, nrch = 0;  // relaxation chunks (staging chunks in place, or packed bins)
    int bin0 = 0, nbin = 0;  // packed bins
    std::vector<int64_t> relax_pre, merge_pre, unpack_pre, split_pre, pack_pre, bin_pre;
    std::vector<int32_t> relax_blk;  // relaxation chunk (relative) of relaxation item 64 * b
    size_t off_rp = 0, off_mp = 0, off_up = 0, off_rb = 0, off_sp = 0, off_pp = 0, off_bp = 0;
};

struct StepPlan {
    std::vector<GChunk> ch;
    std::vector<GPlace> places;      // instances at their first layer: singleton rows to place
    std::vector<int64_t> place_pre;  // rows (k + 1) per place
    std::vector<GChunk> rch;  // relaxation chunks of all groups
    std::vector<GBin> bins;
    std::vector<GLane> lanes;
    std::vector<GMerge> mg;
    std::vector<MergeStar> stars;  // star merges' stars (rows local to their instance's layer)
    StepGroup grp[4];  // label width (u32, u16) x relaxation (vertex-wise, block-local)
    std::vector<int64_t> seed_pre;
    std::vector<uint32_t> xs;
    std::vector<int> seed_chunks;  // indices into ch
    std::vector<int64_t> member_cells;  // group queries: staging cells seeded with 0 (a member, its lane)
    std::vector<int64_t> insts;    // tickets stepped
    int xg = 1;
    int pbits = 32;
    int dense_rounds = 0;  // vertex-wise rounds taking every item (0: a singleton-only step)
    int64_t cells = 0, flagbytes = 0;
    int64_t macc_cells = 0;  // split-merge accumulators of the step
    // device layout (byte offsets into the descriptor buffer)
    size_t off_ch = 0, off_mg = 0, off_sp = 0, off_xs = 0, off_seedch = 0, bytes = 0;
    size_t off_rch = 0, off_bins = 0, off_lanes = 0, off_pl = 0, off_plp = 0, off_mc = 0;
    size_t off_stars = 0;
    std::vector<char> blob;
    std::vector<GChunk> seed_ch;
};

// Step group of an instance: bit 0 its label width (set: 16-bit layer labels),
// bit 1 block-local relaxation on its graph.
int inst_group(const emv_engine *e, const Inst &I) {
    return ((I.stage >= 2 && I.dbits == 16) ? 1 : 0) + (e->graphs[I.graph].nblk > 0 ? 2 : 0);
}

bool group_blocks(int gsel) { return (gsel & 2) != 0; }

// Split merge: items per tile aimed at (the params, else EMV_SPLIT_ITEMS / 32).
int split_items(const emv_engine *e) {
    return e->prm.split_items > 0 ? e->prm.split_items : e->split_items;
}

void build_plan(emv_engine *e, const std::vector<int64_t> &tickets, StepPlan &P) {
    P = StepPlan();
    int64_t dcur = 0, fcur = 0;
    int maxrows = 0, maxstage = 0;
    std::vector<uint32_t> rows;
    std::vector<GMerge> split, starsplit;
    std::vector<MergeStar> tmp_stars;
    std::vector<GChunk> pairs[4];  // per group: the relaxation items of paired chunks
    for (int gsel = 0; gsel < 4; gsel++) {
        StepGroup &SG = P.grp[gsel];
        SG.ch0 = (int)P.ch.size();
        SG.mg0 = (int)P.mg.size();
        // 16-bit labels relaxed vertex-wise: chunks in pairs when the engine
        // pairs (batched solves: no lane packing across instances)
        SG.paired = e->pair_chunks && gsel == 1;
        split.clear();
        starsplit.clear();
        for (int64_t t : tickets) {
            Inst &I = e->inst[t];
            if (inst_group(e, I) != gsel) continue;
            const GraphRec &G = e->graphs[I.graph];
            P.insts.push_back(t);
            GChunk c;
            memset(&c, 0, sizeof(c));  // padding too: descriptors are compared bytewise
            c.tbase = I.abase;
            c.n = (int32_t)G.n;
            c.np = (int32_t)G.np;
            c.voff = (int32_t)G.voff;
            c.k = I.k;
            c.inst = I.slot;
            c.b0 = G.b0;
            c.nblk = G.nblk;
            c.tbits = I.tbits;
            // activity flags: one per vertex, or one per block
            const int64_t nflags = G.nblk > 0 ? G.nblk : G.np;
            c.lstride = kWave;
            if (I.stage == 1) {
                c.dbase = dcur;
                c.lbase = dcur;
                c.fbase = fcur;
                c.x0 = 0;
                c.nx = I.k;
                c.kind = I.groups.empty() ? 1 : 2;
                // group query: lane j starts at 0 on every member of group j
                for (int j = 0; j < (int)I.groups.size(); j++)
                    for (int32_t v : I.groups[(size_t)j])
                        P.member_cells.push_back(dcur + (int64_t)v * kWave + j);
                P.seed_chunks.push_back((int)P.ch.size());
                P.ch.push_back(c);
                dcur += G.np * kWave;
                fcur += nflags;
                continue;
            }
            if (I.stage == 2 && I.k >= 3) {
                GPlace pl;
                memset(&pl, 0, sizeof(pl));
                pl.tbase = I.abase;
                pl.np = (int32_t)G.np;
                pl.k = I.k;
                pl.tbits = I.tbits;
                P.places.push_back(pl);
            }
            rows.clear();
            subsets_of_size(I.k, I.stage, rows);
            const int nrows = (int)rows.size();
            // rows of >= kSplitPairs pairs take the split merge, and so do 16-bit
            // rows of >= star_pairs pairs, merged in stars
            const uint32_t row_pairs = (1u << (I.stage - 1)) - 1u;
            const bool h2 = !(I.flags & EMV_MERGE_1V) && merge_h2_ok(I.tbits, I.k, G.np);
            const bool stars = h2 && e->star_merge && row_pairs >= e->star_pairs;
            const bool splits = (row_pairs >= kSplitPairs || stars) && !e->no_split;
            if (!splits) {
                maxrows = std::max(maxrows, nrows);
                maxstage = std::max(maxstage, I.stage);
            }
            if (I.pbits == 64) P.pbits = 64;
            GMerge M;
            memset(&M, 0, sizeof(M));
            M.chunk0 = (int)P.ch.size() - SG.ch0;
            const int nreal = (nrows + kWave - 1) / kWave;
            // paired: an even number of chunks; a last odd chunk's partner
            // holds INF rows only (merged and packed, never unpacked)
            M.nchunks = SG.paired ? (nreal + 1) / 2 * 2 : nreal;
            M.lpair = SG.paired ? 1 : 0;
            M.ntiles = (int32_t)(G.np / kWave);
            M.nsub = 0;  // set once the step's rows per wave is known
            M.nrows = nrows;
            M.x0 = (int32_t)P.xs.size();
            M.k = I.k;
            M.np = (int32_t)G.np;
            M.tbase = I.abase;
            M.dbase0 = dcur;
            M.S = 0;
            M.tbits = I.tbits;
            M.mbase = 0;
            M.h2 = h2 ? 1 : 0;
            M.mtiles = M.h2 ? (int32_t)((G.np + 2 * kWave - 1) / (2 * kWave)) : M.ntiles;
            if (splits) {
                // parts so that a tile has about split_items items (a step
                // holds many instances' tiles, so few parts pay: fewer
                // atomics), each part >= 4 outer steps
                const int64_t steps = (I.stage - 1 <= 3) ? 1 : (1ll << (I.stage - 1 - 3));
                int64_t units = nrows;
                if (stars) {
                    // 16-bit cells: stars of rows sharing a base set
                    build_stars(rows.data(), nrows, I.k, tmp_stars);
                    M.star0 = (int32_t)P.stars.size();
                    M.nstars = (int32_t)tmp_stars.size();
                    P.stars.insert(P.stars.end(), tmp_stars.begin(), tmp_stars.end());
                    units = M.nstars;
                }
                M.S = (int32_t)std::max<int64_t>(
                    1, std::min<int64_t>((split_items(e) + units - 1) / units, steps / 4));
                M.mbase = P.macc_cells;
                P.macc_cells += (int64_t)nrows * G.np;
            }
            const int x0 = (int)P.xs.size();
            P.xs.insert(P.xs.end(), rows.begin(), rows.end());
            for (int q = 0; q < M.nchunks; q++) {
                c.dbase = dcur;
                c.fbase = fcur;
                c.x0 = x0 + q * kWave;
                c.nx = std::max(0, std::min(kWave, nrows - q * kWave));
                c.kind = 0;
                if (SG.paired) {
                    // labels of the pair in the first chunk's region, 128 per vertex
                    const int64_t pbase = dcur - (q & 1) * G.np * kWave;
                    c.lbase = pbase + (q & 1) * (kWave / 2);
                    c.lstride = 2 * kWave;
                    if ((q & 1) == 0) {
                        GChunk pc = c;
                        pc.lbase = pbase;
                        pc.nx = std::min(2 * kWave, nrows - q * kWave);
                        pairs[gsel].push_back(pc);
                    }
                } else {
                    c.lbase = dcur;
                    c.lstride = kWave;
                }
                if (q < nreal) P.ch.push_back(c);
                dcur += G.np * kWave;
                fcur += nflags;
            }
            // a layer step starts near-dense (a singleton-only step does not)
            P.dense_rounds = e->prm.dense_rounds > 0 ? e->prm.dense_rounds : layer_dense_rounds();
            if (splits && M.nstars > 0)
                starsplit.push_back(M);
            else if (splits)
                split.push_back(M);
            else
                P.mg.push_back(M);
        }
        SG.nch = (int)P.ch.size() - SG.ch0;
        SG.nmg = (int)P.mg.size() - SG.mg0;
        SG.ms0 = (int)P.mg.size();
        SG.nms = (int)(split.size() + starsplit.size());
        SG.nss = (int)starsplit.size();
        P.mg.insert(P.mg.end(), split.begin(), split.end());
        P.mg.insert(P.mg.end(), starsplit.begin(), starsplit.end());
    }
    // rows per merge wave: many rows while rows have few pairs (a wave's
    // fixed costs are spread over more work), one row per wave on the layers
    // whose rows have >= 128 pairs, so that a tile's items fill an XCD and its
    // lower rows are read from L2
    if (e->prm.merge_rows_per_wave > 0)
        P.xg = e->prm.merge_rows_per_wave;
    else if (maxstage >= 9)
        P.xg = 1;
    else if (maxstage <= 5)
        P.xg = 8;
    else
        P.xg = maxrows >= 1024 ? 8 : maxrows >= 256 ? 4 : maxrows >= 64 ? 2 : 1;
    // relaxation chunks: full staging chunks relax in place; the partial ones
    // of instances on one graph are packed into bins (first fit by rows,
    // largest first, a chunk's rows kept together); a bin of one chunk
    // relaxes that chunk in place
    fcur = 0;
    for (int gsel = 0; gsel < 4; gsel++) {
        StepGroup &SG = P.grp[gsel];
        SG.rch0 = (int)P.rch.size();
        SG.bin0 = (int)P.bins.size();
        auto add_relax = [&](GChunk c) {
            c.fbase = fcur;
            fcur += c.nblk > 0 ? c.nblk : c.np;
            P.rch.push_back(c);
        };
        if (SG.paired) {
            // one item per (pair, vertex); pairs relax in place
            for (const GChunk &pc : pairs[gsel]) add_relax(pc);
            SG.nrch = (int)P.rch.size() - SG.rch0;
            SG.nbin = 0;
            continue;
        }
        std::map<int32_t, std::vector<int>> partial;  // graph (voff) -> chunks
        for (int i = 0; i < SG.nch; i++) {
            const GChunk &c = P.ch[SG.ch0 + i];
            if (c.nx >= kWave || e->no_pack)
                add_relax(c);
            else
                partial[c.voff].push_back(SG.ch0 + i);
        }
        for (auto &kv : partial) {
            std::vector<int> &lst = kv.second;
            std::stable_sort(lst.begin(), lst.end(),
                             [&](int a, int b) { return P.ch[a].nx > P.ch[b].nx; });
            std::vector<std::pair<int, std::vector<int>>> bins;  // rows used, chunks
            for (int ci : lst) {
                size_t j = 0;
                while (j < bins.size() && bins[j].first + P.ch[ci].nx > kWave) j++;
                if (j == bins.size()) bins.push_back({0, {}});
                bins[j].first += P.ch[ci].nx;
                bins[j].second.push_back(ci);
            }
            for (auto &bn : bins) {
                if (bn.second.size() == 1) {
                    add_relax(P.ch[bn.second[0]]);
                    continue;
                }
                GChunk c = P.ch[bn.second[0]];
                GBin B;
                memset(&B, 0, sizeof(B));
                B.dbase = dcur;
                B.np = c.np;
                B.map0 = (int32_t)P.lanes.size();
                c.lbase = dcur;
                c.lstride = kWave;
                for (int ci : bn.second)
                    for (int j = 0; j < P.ch[ci].nx; j++)
                        P.lanes.push_back(GLane{P.ch[ci].lbase, j, 0});
                while ((int)P.lanes.size() < B.map0 + kWave) P.lanes.push_back(GLane{-1, 0, 0});
                P.bins.push_back(B);
                c.dbase = dcur;
                c.nx = bn.first;
                dcur += (int64_t)c.np * kWave;
                add_relax(c);
            }
        }
        SG.nrch = (int)P.rch.size() - SG.rch0;
        SG.nbin = (int)P.bins.size() - SG.bin0;
    }
    for (int gsel = 0; gsel < 4; gsel++) {
        StepGroup &SG = P.grp[gsel];
        SG.relax_pre.assign(1, 0);
        SG.unpack_pre.assign(1, 0);
        for (int i = 0; i < SG.nrch; i++) {
            const GChunk &c = P.rch[SG.rch0 + i];
            SG.relax_pre.push_back(SG.relax_pre.back() + (group_blocks(gsel) ? c.nblk : c.n));
        }
        for (int i = 0; i < SG.nch; i++) {
            const GChunk &c = P.ch[SG.ch0 + i];
            SG.unpack_pre.push_back(SG.unpack_pre.back() + (c.n + kWave - 1) / kWave);
        }
        SG.bin_pre.assign(1, 0);
        for (int i = 0; i < SG.nbin; i++)
            SG.bin_pre.push_back(SG.bin_pre.back() + P.bins[SG.bin0 + i].np / kWave);
        SG.merge_pre.assign(1, 0);
        for (int i = 0; i < SG.nmg; i++) {
            GMerge &M = P.mg[SG.mg0 + i];
            M.nsub = M.nchunks * (kWave / P.xg);
            SG.merge_pre.push_back(SG.merge_pre.back() + (int64_t)M.mtiles * M.nsub);
        }
        SG.split_pre.assign(1, 0);
        SG.star_pre.assign(1, 0);
        SG.pack_pre.assign(1, 0);
        for (int i = 0; i < SG.nms; i++) {
            GMerge &M = P.mg[SG.ms0 + i];
            M.nsub = M.nchunks * kWave;
            if (i < SG.nms - SG.nss)
                SG.split_pre.push_back(SG.split_pre.back() + (int64_t)M.mtiles * M.nrows * M.S);
            else
                SG.star_pre.push_back(SG.star_pre.back() + (int64_t)M.mtiles * M.nstars * M.S);
            SG.pack_pre.push_back(SG.pack_pre.back() + (int64_t)M.ntiles * M.nchunks);
        }
        const int64_t nitems = group_blocks(gsel) ? 0 : SG.relax_pre.back();
        SG.relax_blk.assign((size_t)((nitems + kWave - 1) / kWave), 0);
        int g = 0;
        for (size_t b = 0; b < SG.relax_blk.size(); b++) {
            const int64_t it = (int64_t)b * kWave;
            while (SG.relax_pre[g + 1] <= it) g++;
            SG.relax_blk[b] = g;
        }
    }
    P.seed_pre.push_back(0);
    for (int g : P.seed_chunks) {
        P.seed_ch.push_back(P.ch[g]);
        P.seed_pre.push_back(P.seed_pre.back() + (int64_t)P.ch[g].np * kWave);
    }
    P.place_pre.assign(1, 0);
    for (const GPlace &pl : P.places) P.place_pre.push_back(P.place_pre.back() + pl.k + 1);
    P.cells = dcur;
    P.flagbytes = fcur;
    // pack
    auto place = [&](size_t bytes) {
        size_t o = (P.bytes + 255) / 256 * 256;
        P.bytes = o + std::max<size_t>(bytes, 8);
        return o;
    };
    P.off_ch = place(sizeof(GChunk) * P.ch.size());
    P.off_mg = place(sizeof(GMerge) * P.mg.size());
    P.off_sp = place(sizeof(int64_t) * P.seed_pre.size());
    P.off_xs = place(sizeof(uint32_t) * P.xs.size());
    P.off_seedch = place(sizeof(GChunk) * P.seed_ch.size());
    P.off_rch = place(sizeof(GChunk) * P.rch.size());
    P.off_bins = place(sizeof(GBin) * P.bins.size());
    P.off_lanes = place(sizeof(GLane) * P.lanes.size());
    P.off_pl = place(sizeof(GPlace) * P.places.size());
    P.off_plp = place(sizeof(int64_t) * P.place_pre.size());
    P.off_mc = place(sizeof(int64_t) * P.member_cells.size());
    P.off_stars = place(sizeof(MergeStar) * P.stars.size());
    for (StepGroup &SG : P.grp) {
        SG.off_rp = place(sizeof(int64_t) * SG.relax_pre.size());
        SG.off_mp = place(sizeof(int64_t) * SG.merge_pre.size());
        SG.off_up = place(sizeof(int64_t) * SG.unpack_pre.size());
        SG.off_rb = place(sizeof(int32_t) * SG.relax_blk.size());
        SG.off_sp = place(sizeof(int64_t) * SG.split_pre.size());
        SG.off_pp = place(sizeof(int64_t) * SG.pack_pre.size());
        SG.off_bp = place(sizeof(int64_t) * SG.bin_pre.size());
        SG.off_stp = place(sizeof(int64_t) * SG.star_pre.size());
    }
    P.blob.assign(P.bytes, 0);
    auto put = [&](size_t off, const void *src, size_t n) {
        if (n) memcpy(P.blob.data() + off, src, n);
    };
    put(P.off_ch, P.ch.data(), sizeof(GChunk) * P.ch.size());
    put(P.off_mg, P.mg.data(), sizeof(GMerge) * P.mg.size());
    put(P.off_sp, P.seed_pre.data(), sizeof(int64_t) * P.seed_pre.size());
    put(P.off_xs, P.xs.data(), sizeof(uint32_t) * P.xs.size());
    put(P.off_seedch, P.seed_ch.data(), sizeof(GChunk) * P.seed_ch.size());
    put(P.off_rch, P.rch.data(), sizeof(GChunk) * P.rch.size());
    put(P.off_bins, P.bins.data(), sizeof(GBin) * P.bins.size());
    put(P.off_lanes, P.lanes.data(), sizeof(GLane) * P.lanes.size());
    put(P.off_pl, P.places.data(), sizeof(GPlace) * P.places.size());
    put(P.off_plp, P.place_pre.data(), sizeof(int64_t) * P.place_pre.size());
    put(P.off_mc, P.member_cells.data(), sizeof(int64_t) * P.member_cells.size());
    put(P.off_stars, P.stars.data(), sizeof(MergeStar) * P.stars.size());
    for (StepGroup &SG : P.grp) {
        put(SG.off_rp, SG.relax_pre.data(), sizeof(int64_t) * SG.relax_pre.size());
        put(SG.off_mp, SG.merge_pre.data(), sizeof(int64_t) * SG.merge_pre.size());
        put(SG.off_up, SG.unpack_pre.data(), sizeof(int64_t) * SG.unpack_pre.size());
        put(SG.off_rb, SG.relax_blk.data(), sizeof(int32_t) * SG.relax_blk.size());
        put(SG.off_sp, SG.split_pre.data(), sizeof(int64_t) * SG.split_pre.size());
        put(SG.off_pp, SG.pack_pre.data(), sizeof(int64_t) * SG.pack_pre.size());
        put(SG.off_bp, SG.bin_pre.data(), sizeof(int64_t) * SG.bin_pre.size());
        put(SG.off_stp, SG.star_pre.data(), sizeof(int64_t) * SG.star_pre.size());
    }
}

// Host-side bounds of a step plan against the buffers its launches will
// touch: every staging chunk, bin and lane inside the staging, every flag
// slice inside the flag arrays, every table inside the arena, accumulators
// inside theirs, the descriptor blob the size its offsets assume.  Checked
// before anything is enqueued, so a plan that disagrees with its buffers is
// an error status, never an out-of-bounds access on the device.
int check_plan(emv_engine *e, const StepPlan &P) {
    auto fail = [&](const char *what, long long a, long long b) {
        eset(e, "step plan out of bounds: %s (%lld > %lld)", what, a, b);
        return EMV_ERR_STATE;
    };
    if (P.blob.size() != P.bytes) return fail("descriptor blob", (long long)P.blob.size(), (long long)P.bytes);
    if (P.cells > e->stage_cells) return fail("staging cells", P.cells, e->stage_cells);
    if (P.flagbytes > e->flag_bytes) return fail("flag words", P.flagbytes, e->flag_bytes);
    if (P.macc_cells * (P.pbits / 8) > e->macc_bytes)
        return fail("merge accumulators", P.macc_cells * (P.pbits / 8), e->macc_bytes);
    // end (cells) of a chunk's labels: lstride 64 labels of up to 4 bytes,
    // np * 64 cells; a paired chunk's 64 16-bit rows at 128 per vertex end
    // 32 cells (64 rows) before that, a pair's relaxation item (all 128 rows
    // from the pair's base) at np * 64
    auto label_end = [](const GChunk &c, bool item) {
        return c.lbase + (int64_t)c.np * kWave - ((c.lstride == 2 * kWave && !item) ? kWave / 2 : 0);
    };
    for (const GChunk &c : P.ch) {
        if (c.dbase < 0 || c.dbase + (int64_t)c.np * kWave > e->stage_cells)
            return fail("chunk staging", c.dbase + (int64_t)c.np * kWave, e->stage_cells);
        if (c.lbase < 0 || label_end(c, false) > e->stage_cells ||
            (c.lstride != kWave && c.lstride != 2 * kWave))
            return fail("chunk labels", label_end(c, false), e->stage_cells);
        const int64_t tcells = 2 * (1ll << c.k) * (int64_t)c.np;
        if (c.tbase < 0 || c.tbase + tcells > e->arena.cap)
            return fail("instance table", c.tbase + tcells, e->arena.cap);
        if (c.nx < 1 || c.nx > kWave || c.n > c.np) return fail("chunk rows", c.nx, kWave);
    }
    for (const GChunk &c : P.rch) {
        const int64_t nf = c.nblk > 0 ? c.nblk : c.np;
        if (c.fbase < 0 || c.fbase + nf > e->flag_bytes)
            return fail("relaxation flags", c.fbase + nf, e->flag_bytes);
        if (c.dbase < 0 || c.dbase + (int64_t)c.np * kWave > e->stage_cells)
            return fail("relaxation staging", c.dbase + (int64_t)c.np * kWave, e->stage_cells);
        if (c.lbase < 0 || label_end(c, true) > e->stage_cells)
            return fail("relaxation labels", label_end(c, true), e->stage_cells);
    }
    for (const GBin &B : P.bins) {
        if (B.dbase < 0 || B.dbase + (int64_t)B.np * kWave > e->stage_cells)
            return fail("bin staging", B.dbase + (int64_t)B.np * kWave, e->stage_cells);
        if (B.map0 < 0 || (size_t)B.map0 + kWave > P.lanes.size())
            return fail("bin lanes", (long long)B.map0 + kWave, (long long)P.lanes.size());
    }
    for (const GLane &L : P.lanes)
        if (L.dbase >= 0 && (L.lane < 0 || L.lane >= kWave || L.dbase >= e->stage_cells))
            return fail("lane source", L.dbase, e->stage_cells);
    for (const GMerge &M : P.mg) {
        if (M.S > 0 && M.mbase + (int64_t)M.nrows * M.np > P.macc_cells)
            return fail("split rows", M.mbase + (int64_t)M.nrows * M.np, P.macc_cells);
        if (M.nstars > 0) {
            if (M.star0 < 0 || (size_t)M.star0 + M.nstars > P.stars.size() || M.S < 1 || !M.h2)
                return fail("merge stars", (long long)M.star0 + M.nstars, (long long)P.stars.size());
            for (int i = 0; i < M.nstars; i++) {
                const MergeStar &st = P.stars[(size_t)M.star0 + i];
                if (st.r < 1 || st.r > (uint32_t)kStarMax) return fail("star rows", st.r, kStarMax);
                for (uint32_t j = 0; j < st.r; j++)
                    if (st.xi[j] < 0 || st.xi[j] >= M.nrows)
                        return fail("star row index", st.xi[j], M.nrows);
            }
        }
        if (M.x0 < 0 || (size_t)M.x0 + M.nrows > P.xs.size())
            return fail("merge rows", (long long)M.x0 + M.nrows, (long long)P.xs.size());
    }
    for (int64_t c : P.member_cells)
        if (c < 0 || c >= e->stage_cells) return fail("member seed", c, e->stage_cells);
    return EMV_OK;
}

int desc_buffer(emv_engine *e, size_t idx, size_t bytes, char **out) {
    while (e->desc_bufs.size() <= idx) {
        e->desc_bufs.push_back(nullptr);
        e->desc_caps.push_back(0);
    }
    if ((size_t)e->desc_caps[idx] < bytes) {
        if (e->desc_bufs[idx]) (void)hipFree(e->desc_bufs[idx]);
        e->desc_bufs[idx] = nullptr;
        e->desc_caps[idx] = 0;
        EHIP(e, step_malloc(e, &e->desc_bufs[idx], bytes * 2));
        e->desc_caps[idx] = (int64_t)bytes * 2;
    }
    *out = (char *)e->desc_bufs[idx];
    return EMV_OK;
}

template <typename DT, int DMAX, bool PK = false>
void launch_grelax(emv_engine *e, char *db, const StepPlan &P, const StepGroup &SG, int dense,
                   int r) {
    // no more workgroups than 64-item blocks of waves
    const int64_t need = (SG.relax_pre.back() + (int64_t)kBlock - 1) / kBlock;
    const int grid = (int)std::max<int64_t>(
        1, std::min<int64_t>((int64_t)e->num_cu * e->prm.relax_blocks_per_cu, need));
    // the worklist numbers items within one group: only for single-group steps
    int groups = 0;
    for (const StepGroup &G : P.grp) groups += G.nrch > 0 ? 1 : 0;
    const uint32_t qcap = groups > 1 ? 0u : e->qcap;
    hipLaunchKernelGGL((k_grelax<DT, DMAX, PK>), dim3(grid), dim3(kBlock), 0, e->stream,
                       (void *)e->d_D, (const GChunk *)(db + P.off_rch) + SG.rch0,
                       (const int64_t *)(db + SG.off_rp), (const int32_t *)(db + SG.off_rb),
                       SG.nrch, e->d_rp, e->d_adj, e->d_ell, dense, e->d_flag[r & 1],
                       e->d_flag[(r + 1) & 1], (const uint32_t *)e->d_q[r & 1],
                       e->d_q[(r + 1) & 1], qcap, e->d_any, e->d_cnt, r, P.dense_rounds);
}

template <typename DT>
void grelax_group(emv_engine *e, char *db, const StepPlan &P, const StepGroup &SG, int dense,
                  int r) {
    if constexpr (std::is_same<DT, uint16_t>::value) {
        if (SG.paired) {
            // pairs of 16-bit chunks: two labels per 32-bit word
            switch (e->ell_deg) {
            case 4: launch_grelax<uint32_t, 4, true>(e, db, P, SG, dense, r); break;
            case 8: launch_grelax<uint32_t, 8, true>(e, db, P, SG, dense, r); break;
            case 16: launch_grelax<uint32_t, 16, true>(e, db, P, SG, dense, r); break;
            default: launch_grelax<uint32_t, 0, true>(e, db, P, SG, dense, r); break;
            }
            return;
        }
    }
    switch (e->ell_deg) {
    case 4: launch_grelax<DT, 4>(e, db, P, SG, dense, r); break;
    case 8: launch_grelax<DT, 8>(e, db, P, SG, dense, r); break;
    case 16: launch_grelax<DT, 16>(e, db, P, SG, dense, r); break;
    default: launch_grelax<DT, 0>(e, db, P, SG, dense, r); break;
    }
}

// Block-local round over a block group: one workgroup per (chunk, block)
// visit, the block's labels in LDS; as many workgroups per CU as its LDS
// holds (at most two of 1024 threads).
template <typename DT, typename DW>
void launch_gblk_w(emv_engine *e, char *db, const StepPlan &P, const StepGroup &SG, int dense,
                   int r) {
    const size_t bytes = blk_lds_bytes<DT>();
    const int per_cu = (int)std::max<size_t>(1, std::min<size_t>(2, e->lds_cu / bytes));
    const int grid = (int)std::max<int64_t>(
        1, std::min<int64_t>((int64_t)e->num_cu * per_cu, SG.relax_pre.back()));
    hipLaunchKernelGGL((k_gblk<DT, DW>), dim3(grid), dim3(kBT), bytes, e->stream,
                       (void *)e->d_D, (const GChunk *)(db + P.off_rch) + SG.rch0,
                       (const int64_t *)(db + SG.off_rp), SG.nrch, e->d_boff, e->d_bvert,
                       e->d_brp, e->d_bext, e->d_badj, e->d_vblk, (const DW *)e->d_bdist,
                       e->d_bdoff, P.rch[SG.rch0].fbase, dense, e->d_flag[r & 1], e->d_flag[(r + 1) & 1],
                       e->d_any, r);
}

template <typename DT>
void launch_gblk(emv_engine *e, char *db, const StepPlan &P, const StepGroup &SG, int dense,
                 int r) {
    if (e->bdw == 32)
        launch_gblk_w<DT, uint32_t>(e, db, P, SG, dense, r);
    else
        launch_gblk_w<DT, uint16_t>(e, db, P, SG, dense, r);
}

// One round over the step's groups (one launch each, sharing the round's
// flags and "round has work" words).
void grelax_round(emv_engine *e, char *db, const StepPlan &P, int dense, int r) {
    if (P.grp[0].nrch > 0) grelax_group<uint32_t>(e, db, P, P.grp[0], dense, r);
    if (P.grp[1].nrch > 0) grelax_group<uint16_t>(e, db, P, P.grp[1], dense, r);
    if (P.grp[2].nrch > 0) launch_gblk<uint32_t>(e, db, P, P.grp[2], dense, r);
    if (P.grp[3].nrch > 0) launch_gblk<uint16_t>(e, db, P, P.grp[3], dense, r);
}

template <typename P_, int XG, typename DT>
void launch_gmerge_xg(emv_engine *e, char *db, const StepPlan &P, const StepGroup &SG) {
    const int64_t items = SG.merge_pre.back();
    if (items == 0) return;
    const int64_t blocks = (items + (kBlock / kWave) - 1) / (kBlock / kWave);
    const int nxcd = (e->num_cu % 32 == 0 && e->num_cu >= 64) ? e->num_cu / 32 : 1;
    const int64_t per = (blocks + nxcd - 1) / nxcd;
    hipLaunchKernelGGL((k_gmerge<P_, XG, DT>), dim3((unsigned)(per * nxcd)), dim3(kBlock), 0,
                       e->stream, e->d_arena, (const GMerge *)(db + P.off_mg) + SG.mg0,
                       (const int64_t *)(db + SG.off_mp),
                       SG.nmg, (const uint32_t *)(db + P.off_xs), (void *)e->d_D, (P_ *)e->d_code, nxcd,
                       per);
}

template <typename P_, typename DT>
void launch_gmerge_p(emv_engine *e, char *db, const StepPlan &P, const StepGroup &SG) {
    switch (P.xg) {
    case 8: launch_gmerge_xg<P_, 8, DT>(e, db, P, SG); break;
    case 4: launch_gmerge_xg<P_, 4, DT>(e, db, P, SG); break;
    case 2: launch_gmerge_xg<P_, 2, DT>(e, db, P, SG); break;
    default: launch_gmerge_xg<P_, 1, DT>(e, db, P, SG); break;
    }
}

template <typename P_, typename DT>
void launch_gsplit_p(emv_engine *e, char *db, const StepPlan &P, const StepGroup &SG) {
    const int64_t items = SG.split_pre.back();
    if (items == 0 && SG.star_pre.back() == 0) return;
    const int64_t blocks = (items + (kBlock / kWave) - 1) / (kBlock / kWave);
    const int nxcd = (e->num_cu % 32 == 0 && e->num_cu >= 64) ? e->num_cu / 32 : 1;
    const int64_t per = (blocks + nxcd - 1) / nxcd;
    const GMerge *mg = (const GMerge *)(db + P.off_mg) + SG.ms0;
    const int nrow = SG.nms - SG.nss;  // rows cut into parts, then the star merges
    if (items > 0)
        hipLaunchKernelGGL((k_gmerge_split<P_>), dim3((unsigned)(per * nxcd)), dim3(kBlock), 0,
                           e->stream, e->d_arena, mg, (const int64_t *)(db + SG.off_sp), nrow,
                           (const uint32_t *)(db + P.off_xs), (P_ *)e->d_macc, nxcd, per);
    if (SG.star_pre.back() > 0) {
        const int64_t sblocks = (SG.star_pre.back() + (kBlock / kWave) - 1) / (kBlock / kWave);
        const int64_t sper = (sblocks + nxcd - 1) / nxcd;
        hipLaunchKernelGGL((k_gmerge_star<P_>), dim3((unsigned)(sper * nxcd)), dim3(kBlock), 0,
                           e->stream, e->d_arena, mg + nrow, (const int64_t *)(db + SG.off_stp),
                           SG.nss, (const MergeStar *)(db + P.off_stars), (P_ *)e->d_macc, nxcd,
                           sper);
    }
    hipLaunchKernelGGL((k_gpack<P_, DT>), dim3((unsigned)SG.pack_pre.back()), dim3(kBlock), 0,
                       e->stream, (const P_ *)e->d_macc, mg, (const int64_t *)(db + SG.off_pp),
                       SG.nms, (void *)e->d_D, (P_ *)e->d_code);
}

template <typename DT>
void launch_gmerge_group(emv_engine *e, char *db, const StepPlan &P, const StepGroup &SG) {
    if (SG.nmg > 0) {
        if (P.pbits == 64)
            launch_gmerge_p<uint64_t, DT>(e, db, P, SG);
        else
            launch_gmerge_p<uint32_t, DT>(e, db, P, SG);
    }
    if (SG.nms > 0) {
        if (P.pbits == 64)
            launch_gsplit_p<uint64_t, DT>(e, db, P, SG);
        else
            launch_gsplit_p<uint32_t, DT>(e, db, P, SG);
    }
}

// Rows of partial chunks into their bins (in) or back (out).
template <typename DT, bool IN>
void launch_glanes(emv_engine *e, char *db, const StepPlan &P, const StepGroup &SG) {
    if (SG.nbin == 0) return;
    hipLaunchKernelGGL((k_glanes<DT, IN>), dim3((unsigned)SG.bin_pre.back()), dim3(kBlock), 0,
                       e->stream, (void *)e->d_D, (const GBin *)(db + P.off_bins) + SG.bin0,
                       (const int64_t *)(db + SG.off_bp), SG.nbin,
                       (const GLane *)(db + P.off_lanes));
}

template <bool IN>
void glanes_all(emv_engine *e, char *db, const StepPlan &P) {
    for (int gsel = 0; gsel < 4; gsel++) {
        if (gsel & 1)
            launch_glanes<uint16_t, IN>(e, db, P, P.grp[gsel]);
        else
            launch_glanes<uint32_t, IN>(e, db, P, P.grp[gsel]);
    }
}

// Launches of one step up to (not including) the relaxation rounds.
int enqueue_front(emv_engine *e, char *db, const StepPlan &P) {
    if (!P.seed_ch.empty()) {
        const int64_t cells = P.seed_pre.back();
        const int blocks = (int)std::min<int64_t>((cells + kBlock - 1) / kBlock, 65536);
        hipLaunchKernelGGL(k_gseed, dim3(blocks), dim3(kBlock), 0, e->stream,
                           (const GChunk *)(db + P.off_seedch), (const int64_t *)(db + P.off_sp),
                           (int)P.seed_ch.size(), e->d_kk, e->d_kkoff, e->d_D);
        if (!P.member_cells.empty()) {
            const int64_t cnt = (int64_t)P.member_cells.size();
            hipLaunchKernelGGL(k_gseed_cells,
                               dim3((unsigned)std::min<int64_t>((cnt + kBlock - 1) / kBlock, 65536)),
                               dim3(kBlock), 0, e->stream, (const int64_t *)(db + P.off_mc), cnt,
                               e->d_D);
        }
    }
    if (!P.places.empty())
        hipLaunchKernelGGL(k_gplace, dim3((unsigned)P.place_pre.back()), dim3(kBlock), 0,
                           e->stream, e->d_arena, (const GPlace *)(db + P.off_pl),
                           (const int64_t *)(db + P.off_plp), (int)P.places.size());
    if (P.macc_cells > 0) {
        // the split accumulators start at all ones (a fill kernel: graph-safe)
        const int64_t words = P.macc_cells * (P.pbits / 32);
        hipLaunchKernelGGL(k_fill<uint32_t>,
                           dim3((unsigned)std::min<int64_t>((words + kBlock - 1) / kBlock, 1 << 20)),
                           dim3(kBlock), 0, e->stream, (uint32_t *)e->d_macc, words, ~0u);
    }
    for (int gsel = 0; gsel < 4; gsel++) {
        if (gsel & 1)
            launch_gmerge_group<uint16_t>(e, db, P, P.grp[gsel]);
        else
            launch_gmerge_group<uint32_t>(e, db, P, P.grp[gsel]);
    }
    glanes_all<true>(e, db, P);
    hipLaunchKernelGGL(k_fill<uint32_t>, dim3(1), dim3(64), 0, e->stream, e->d_any, (int64_t)4,
                       0u);
    hipLaunchKernelGGL(k_fill<uint32_t>, dim3(1), dim3(64), 0, e->stream, e->d_cnt, (int64_t)4,
                       0u);
    EHIP(e, hipGetLastError());
    return EMV_OK;
}

template <typename DT>
void launch_gunpack(emv_engine *e, char *db, const StepPlan &P, const StepGroup &SG) {
    const int64_t blocks = SG.unpack_pre.back();
    if (blocks == 0) return;
    if (P.pbits == 64)
        hipLaunchKernelGGL((k_gunpack<DT, uint64_t>), dim3((unsigned)blocks), dim3(kBlock), 0,
                           e->stream, (const void *)e->d_D, (const uint64_t *)e->d_code,
                           (const GChunk *)(db + P.off_ch) + SG.ch0,
                           (const int64_t *)(db + SG.off_up), SG.nch,
                           (const uint32_t *)(db + P.off_xs), e->d_kk, e->d_kkoff, e->d_rp,
                           e->d_adj, e->d_arena, e->d_arena, e->d_rowmax);
    else
        hipLaunchKernelGGL((k_gunpack<DT, uint32_t>), dim3((unsigned)blocks), dim3(kBlock), 0,
                           e->stream, (const void *)e->d_D, (const uint32_t *)e->d_code,
                           (const GChunk *)(db + P.off_ch) + SG.ch0,
                           (const int64_t *)(db + SG.off_up), SG.nch,
                           (const uint32_t *)(db + P.off_xs), e->d_kk, e->d_kkoff, e->d_rp,
                           e->d_adj, e->d_arena, e->d_arena, e->d_rowmax);
}

int enqueue_back(emv_engine *e, char *db, const StepPlan &P) {
    glanes_all<false>(e, db, P);
    for (int gsel = 0; gsel < 4; gsel++) {
        if (gsel & 1)
            launch_gunpack<uint16_t>(e, db, P, P.grp[gsel]);
        else
            launch_gunpack<uint32_t>(e, db, P, P.grp[gsel]);
    }
    EHIP(e, hipGetLastError());
    return EMV_OK;
}

int run_relax_eager(emv_engine *e, char *db, const StepPlan &P, int *rounds_used) {
    int r = 0;
    grelax_round(e, db, P, 1, r);
    EHIP(e, hipGetLastError());
    r = 1;
    int rounds = 1, group = 2;
    for (;;) {
        for (int g = 0; g < group; g++, r++, rounds++) grelax_round(e, db, P, 0, r);
        EHIP(e, hipGetLastError());
        EHIP(e, hipMemcpyAsync(e->h_pin, e->d_any + (r & 3), sizeof(uint32_t),
                               hipMemcpyDeviceToHost, e->stream));
        EHIP(e, hipStreamSynchronize(e->stream));
        if (*e->h_pin == 0) break;
        group = std::min(group * 2, 32);
        if (rounds > 4000000) {
            eset(e, "relaxation did not converge");
            return EMV_ERR_RANGE;
        }
    }
    // the rounds that did work: the last group may have run empty rounds
    *rounds_used = rounds;
    e->relax_rounds += rounds;
    return EMV_OK;
}

int read_costs_and_finish(emv_engine *e, const std::vector<int64_t> &done);

// After a step: singleton instances get their bound (and P width), layer
// instances advance; instances past their last layer finish.
int advance(emv_engine *e, const StepPlan &P) {
    std::vector<int> single_slots;
    for (int64_t t : P.insts)
        if (e->inst[t].stage == 1) single_slots.push_back(e->inst[t].slot);
    std::vector<unsigned long long> rm;
    if (!single_slots.empty()) {
        rm.resize((size_t)e->max_slots * kWave);
        EHIP(e, hipMemcpyAsync(rm.data(), e->d_rowmax,
                               sizeof(unsigned long long) * e->max_slots * kWave,
                               hipMemcpyDeviceToHost, e->stream));
        EHIP(e, hipStreamSynchronize(e->stream));
    }
    std::vector<int64_t> done;
    for (int64_t t : P.insts) {
        Inst &I = e->inst[t];
        if (I.stage == 1) {
            unsigned long long B = 0;
            bool disc = false;
            for (int j = 0; j < I.k; j++) {
                const unsigned long long d = rm[(size_t)I.slot * kWave + j];
                if (d >= 0xFFFFFFFFull) disc = true;
                B += d;
            }
            if (disc) {
                I.status = EMV_ERR_DISCONNECTED;
                done.push_back(t);
                continue;
            }
            I.pbits = (bits_for(B) + I.k <= 32) ? 32 : 64;
            // 16-bit cells hold every finite cell below 0xFFFF (row 0's INF)
            I.tbits = (I.k >= 3 && B < 0xFFFFull && !(I.flags & EMV_TABLE32)) ? 16 : 32;
            // labels are at most B, candidates at most B + max w
            I.dbits = (B + (unsigned long long)e->graphs[I.graph].maxw < 0xFFFFull &&
                       !(I.flags & EMV_WIDE_LABELS))
                          ? 16
                          : 32;
            if (B + (unsigned long long)e->graphs[I.graph].maxw >= 0xFFFFFFFFull ||
                B >= (unsigned long long)INT32_MAX) {
                I.status = EMV_ERR_RANGE;
                done.push_back(t);
                continue;
            }
            I.stage = 2;
        } else {
            I.stage++;
        }
        if (I.stage > I.k) done.push_back(t);
    }
    return read_costs_and_finish(e, done);
}

int traceback_inst(emv_engine *e, Inst &I);

// Group query (k groups): the reduced instance joins a virtual terminal to
// every member of its group by an edge of weight M = 1 + sum of weights.
// M exceeds every tree of the graph, so an optimal reduced tree uses each
// virtual terminal as a leaf: its cells at graph vertices are
// |X| * M + f[X][v], f being this table (singleton rows = distance to the
// nearest member), and its answer cell f[C][virtual terminal k-1] =
// k * M + min over the last group's members s of f[C][s].  The traceback
// starts at the smallest such s (the reduced instance's record at the
// virtual terminal is its smallest neighbour on a shortest path).
int group_answer(emv_engine *e, Inst &I) {
    const GraphRec &G = e->graphs[I.graph];
    const int64_t C = ((int64_t)1 << (I.k - 1)) - 1;
    std::vector<int32_t> row((size_t)G.np);
    if (I.tbits == 16) {
        std::vector<uint16_t> r16((size_t)G.np);
        EHIP(e, hipMemcpyAsync(r16.data(), (const uint16_t *)(e->d_arena + I.abase) + C * G.np,
                               sizeof(uint16_t) * G.n, hipMemcpyDeviceToHost, e->stream));
        EHIP(e, hipStreamSynchronize(e->stream));
        for (int64_t v = 0; v < G.n; v++)
            row[(size_t)v] = r16[(size_t)v] == 0xFFFF ? INT32_MAX : (int32_t)r16[(size_t)v];
    } else {
        EHIP(e, hipMemcpyAsync(row.data(), e->d_arena + I.abase + C * G.np, sizeof(int32_t) * G.n,
                               hipMemcpyDeviceToHost, e->stream));
        EHIP(e, hipStreamSynchronize(e->stream));
    }
    int64_t best = INT64_MAX, root = -1;
    for (int32_t s : I.groups.back()) {
        const int64_t x = row[(size_t)s] == INT32_MAX ? INT64_MAX : (int64_t)row[(size_t)s];
        if (x < best || (x == best && x != INT64_MAX && s < root)) {
            best = x;
            root = s;
        }
    }
    I.cost = best;
    I.root = root;
    const int64_t M = 1 + G.wsum;
    I.reduced_cost = (best == INT64_MAX) ? INT64_MAX : best + (int64_t)I.k * M;
    return EMV_OK;
}

int read_costs_and_finish(emv_engine *e, const std::vector<int64_t> &done) {
    for (int64_t t : done) {
        Inst &I = e->inst[t];
        const GraphRec &G = e->graphs[I.graph];
        if (I.status == EMV_OK && !I.groups.empty()) {
            ECHK(group_answer(e, I));
            if (I.cost != INT64_MAX && (I.flags & EMV_LIST)) ECHK(traceback_inst(e, I));
        } else if (I.status == EMV_OK) {
            const int64_t C = (I.k == 1) ? 0 : (((int64_t)1 << (I.k - 1)) - 1);
            I.root = I.kk[I.k - 1];
            if (C == 0) {
                I.cost = INT64_MAX;
            } else {
                const int64_t cell = C * G.np + I.kk[I.k - 1];
                if (I.tbits == 16) {
                    uint16_t v = 0;
                    EHIP(e, hipMemcpyAsync(&v, (uint16_t *)(e->d_arena + I.abase) + cell,
                                           sizeof(uint16_t), hipMemcpyDeviceToHost, e->stream));
                    EHIP(e, hipStreamSynchronize(e->stream));
                    I.cost = (v == 0xFFFF) ? INT64_MAX : (int64_t)v;
                } else {
                    int32_t v = 0;
                    EHIP(e, hipMemcpyAsync(&v, e->d_arena + I.abase + cell, sizeof(int32_t),
                                           hipMemcpyDeviceToHost, e->stream));
                    EHIP(e, hipStreamSynchronize(e->stream));
                    I.cost = (v == INT32_MAX) ? INT64_MAX : (int64_t)v;
                }
                if (I.flags & EMV_LIST) ECHK(traceback_inst(e, I));
            }
            I.reduced_cost = I.cost;
        }
        I.state = ST_DONE;
        I.t_done = now_ms();
        if (!(I.flags & EMV_KEEP_TABLE) && I.abase >= 0) {
            e->arena.release(I.abase, I.acells);
            I.abase = -1;
        }
        if (I.slot >= 0) {
            e->free_slots.push_back(I.slot);
            I.slot = -1;
        }
    }
    return EMV_OK;
}

int traceback_inst(emv_engine *e, Inst &I) {
    const GraphRec &G = e->graphs[I.graph];
    const int64_t S2 = 1ll << I.k;
    const int64_t C = (((int64_t)1 << (I.k - 1)) - 1);
    const int64_t cap = 2 * (G.n + I.k) + 64, stcap = 4 * I.k + 64;
    int64_t *d_pairs = nullptr, *d_stack = nullptr, *d_res = nullptr;
    EHIP(e, hipMalloc((void **)&d_pairs, sizeof(int64_t) * 2 * cap));
    if (hipMalloc((void **)&d_stack, sizeof(int64_t) * 2 * stcap) != hipSuccess ||
        hipMalloc((void **)&d_res, sizeof(int64_t)) != hipSuccess) {
        dfree(d_pairs);
        dfree(d_stack);
        eset(e, "traceback allocation failed");
        return EMV_ERR_NOMEM;
    }
    hipLaunchKernelGGL(k_traceback, dim3(1), dim3(64), 0, e->stream,
                       e->d_arena + I.abase + S2 * G.np, (int64_t)G.np, I.k, I.root, C,
                       d_pairs, cap, d_stack, stcap, d_res);
    int64_t cnt = 0;
    hipError_t x1 = hipMemcpyAsync(&cnt, d_res, sizeof(int64_t), hipMemcpyDeviceToHost, e->stream);
    hipError_t x2 = hipStreamSynchronize(e->stream);
    if (x1 == hipSuccess && x2 == hipSuccess && cnt >= 0) {
        I.pairs.resize((size_t)(2 * cnt));
        if (cnt > 0)
            x1 = hipMemcpy(I.pairs.data(), d_pairs, sizeof(int64_t) * 2 * cnt,
                           hipMemcpyDeviceToHost);
    }
    dfree(d_pairs);
    dfree(d_stack);
    dfree(d_res);
    if (x1 != hipSuccess || x2 != hipSuccess || cnt < 0) {
        eset(e, "traceback failed");
        return cnt < 0 ? EMV_ERR_CAPACITY : EMV_ERR_HIP;
    }
    return EMV_OK;
}

// Admission: FIFO; the head waits until its table fits the arena.
int admit(emv_engine *e) {
    while (!e->queue.empty() && !e->free_slots.empty()) {
        const int64_t t = e->queue.front();
        Inst &I = e->inst[t];
        const GraphRec &G = e->graphs[I.graph];
        const int64_t cells = 2 * ((int64_t)1 << I.k) * G.np;
        const int64_t off = e->arena.alloc(cells);
        if (off < 0) break;
        e->queue.pop_front();
        I.abase = off;
        I.acells = cells;
        I.slot = e->free_slots.back();
        e->free_slots.pop_back();
        I.state = ST_RUNNING;
        I.stage = 1;
        I.t_admit = now_ms();
        std::vector<int32_t> k64(kWave, -1);
        // group queries seed their members separately (k_gseed_cells)
        for (int j = 0; j < I.k; j++) k64[j] = I.groups.empty() ? I.kk[j] : -1;
        EHIP(e, hipMemcpyAsync(e->d_kk + (int64_t)I.slot * kWave, k64.data(),
                               sizeof(int32_t) * kWave, hipMemcpyHostToDevice, e->stream));
        // table: row 0 INF with record -1 (every other real cell is written by
        // the solve; padding columns only feed dropped lanes); rowmax = 0
        const int64_t S2 = 1ll << I.k;
        // the cell width is decided by the bound after the singleton stage;
        // with k >= 3 the first layer step writes row 0 (k_gplace)
        I.tbits = 32;
        if (I.k < 3)
            hipLaunchKernelGGL(k_fill<int32_t>, dim3(4), dim3(kBlock), 0, e->stream,
                               e->d_arena + off, G.np, INT32_MAX);
        hipLaunchKernelGGL(k_fill<int32_t>, dim3(4), dim3(kBlock), 0, e->stream,
                           e->d_arena + off + S2 * G.np, G.np, -1);
        hipLaunchKernelGGL(k_fill<unsigned long long>, dim3(1), dim3(64), 0, e->stream,
                           e->d_rowmax + (int64_t)I.slot * kWave, (int64_t)kWave, 0ull);
        EHIP(e, hipGetLastError());
    }
    return EMV_OK;
}

// One eager step over every running instance.
int step_eager(emv_engine *e, int *rounds_used) {
    ECHK(admit(e));
    std::vector<int64_t> run;
    for (size_t t = 0; t < e->inst.size(); t++)
        if (e->inst[t].state == ST_RUNNING) run.push_back((int64_t)t);
    *rounds_used = 0;
    if (run.empty()) return EMV_OK;
    StepPlan P;
    build_plan(e, run, P);
    ECHK(ensure_stage(e, P.cells, P.flagbytes));
    ECHK(ensure_macc(e, P.macc_cells * (P.pbits / 8)));
    ECHK(check_plan(e, P));
    char *db = nullptr;
    ECHK(desc_buffer(e, 0, P.bytes, &db));
    EHIP(e, hipMemcpyAsync(db, P.blob.data(), P.bytes, hipMemcpyHostToDevice, e->stream));
    const bool trace = getenv("EMV_TRACE") != nullptr;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    if (trace)
        for (auto &x : ev) EHIP(e, hipEventCreate(&x));
    if (trace) EHIP(e, hipEventRecord(ev[0], e->stream));
    ECHK(enqueue_front(e, db, P));
    if (trace) EHIP(e, hipEventRecord(ev[1], e->stream));
    ECHK(run_relax_eager(e, db, P, rounds_used));
    if (trace) EHIP(e, hipEventRecord(ev[2], e->stream));
    ECHK(enqueue_back(e, db, P));
    e->steps++;
    if (trace) {
        EHIP(e, hipEventRecord(ev[3], e->stream));
        EHIP(e, hipEventSynchronize(ev[3]));
        float ms[3] = {0, 0, 0};
        for (int i = 0; i < 3; i++) EHIP(e, hipEventElapsedTime(&ms[i], ev[i], ev[i + 1]));
        for (auto &x : ev) (void)hipEventDestroy(x);
        int64_t items[4];
        for (int g = 0; g < 4; g++) items[g] = P.grp[g].relax_pre.back();
        fprintf(stderr,
                "emv-trace: step %lld instances=%zu chunks=%zu relax_chunks=%zu bins=%zu "
                "ms merge %.2f relax %.2f unpack %.2f "
                "rounds=%d items u32/u16 vertex %lld/%lld block %lld/%lld\n",
                (long long)e->steps, P.insts.size(), P.ch.size(), P.rch.size(), P.bins.size(),
                ms[0], ms[1], ms[2], *rounds_used,
                (long long)items[0], (long long)items[1], (long long)items[2],
                (long long)items[3]);
    }
    return advance(e, P);
}

int engine_open(int device, emv_engine **out) {
    emv_engine *e = new (std::nothrow) emv_engine();
    if (!e) return EMV_ERR_NOMEM;
    e->device = device;
    e->prm.relax_blocks_per_cu = 64;
    e->prm.merge_rows_per_wave = 0;
    e->prm.use_graph = 1;
    e->prm.graph_round_margin = 2;
    e->prm.dense_rounds = 0;
    e->prm.split_items = 0;
    *out = e;
    if (hipSetDevice(device) != hipSuccess) return EMV_ERR_HIP;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) return EMV_ERR_HIP;
    e->num_cu = prop.multiProcessorCount;
    if (prop.maxSharedMemoryPerMultiProcessor > 0) e->lds_cu = prop.maxSharedMemoryPerMultiProcessor;
    EHIP(e, hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking));
    EHIP(e, hipMalloc((void **)&e->d_any, sizeof(uint32_t) * 4));
    EHIP(e, hipMalloc((void **)&e->d_cnt, sizeof(uint32_t) * 4));
    if (const char *qc = getenv("EMV_QCAP")) e->qcap = (uint32_t)std::max(0, atoi(qc));
    e->no_split = getenv("EMV_NO_SPLIT") != nullptr;
    e->no_pack = getenv("EMV_NO_PACK") != nullptr;
    if (const char *ms = getenv("EMV_MERGE_STAR")) e->star_merge = atoi(ms) != 0;
    if (const char *sp = getenv("EMV_STAR_PAIRS")) e->star_pairs = (uint32_t)std::max(1, atoi(sp));
    if (const char *si = getenv("EMV_SPLIT_ITEMS")) e->split_items = std::max(1, atoi(si));
    EHIP(e, hipMalloc((void **)&e->d_q[0], sizeof(uint32_t) * std::max<uint32_t>(e->qcap, 1)));
    EHIP(e, hipMalloc((void **)&e->d_q[1], sizeof(uint32_t) * std::max<uint32_t>(e->qcap, 1)));
    EHIP(e, hipHostMalloc((void **)&e->h_pin, sizeof(uint32_t) * 16, hipHostMallocDefault));
    return EMV_OK;
}

}  // namespace

// ================================================================ C ABI

extern "C" int emv_engine_default_params(emv_engine_params *p) {
    if (!p) return EMV_ERR_ARG;
    p->relax_blocks_per_cu = 64;
    p->merge_rows_per_wave = 0;
    p->use_graph = 1;
    p->graph_round_margin = 2;
    p->dense_rounds = 0;
    p->split_items = 0;
    return EMV_OK;
}

extern "C" int emv_engine_set_params(emv_engine *e, const emv_engine_params *p) {
    if (!e || !p) return EMV_ERR_ARG;
    if (p->relax_blocks_per_cu < 1 || p->relax_blocks_per_cu > 256 ||
        (p->merge_rows_per_wave != 0 && p->merge_rows_per_wave != 1 &&
         p->merge_rows_per_wave != 2 && p->merge_rows_per_wave != 4 &&
         p->merge_rows_per_wave != 8) ||
        p->graph_round_margin < -1000 || p->graph_round_margin > 1000 ||
        p->dense_rounds < 0 || p->dense_rounds > 64 || p->split_items < 0 ||
        p->split_items > 4096) {
        eset(e, "engine parameters out of range");
        return EMV_ERR_ARG;
    }
    e->prm = *p;
    return EMV_OK;
}

extern "C" const char *emv_engine_last_error(const emv_engine *e) {
    return e ? e->err : "no engine";
}

extern "C" void emv_engine_destroy(emv_engine *e) {
    if (!e) return;
    (void)hipSetDevice(e->device);
    if (e->stream) (void)hipStreamSynchronize(e->stream);
    dfree(e->d_rp);
    dfree(e->d_adj);
    dfree(e->d_ell);
    dfree(e->d_boff);
    dfree(e->d_bvert);
    dfree(e->d_brp);
    dfree(e->d_bext);
    dfree(e->d_vblk);
    dfree(e->d_badj);
    dfree(e->d_bdist);
    dfree(e->d_bdoff);
    dfree(e->d_arena);
    dfree(e->d_kk);
    dfree(e->d_kkoff);
    dfree(e->d_rowmax);
    dfree(e->d_D);
    dfree(e->d_code);
    dfree(e->d_flag[0]);
    dfree(e->d_flag[1]);
    dfree(e->d_macc);
    dfree(e->d_any);
    dfree(e->d_cnt);
    dfree(e->d_q[0]);
    dfree(e->d_q[1]);
    if (e->gexec) (void)hipGraphExecDestroy(e->gexec);
    if (e->ggraph) (void)hipGraphDestroy(e->ggraph);
    dfree(e->d_conv);
    for (void *p : e->desc_bufs)
        if (p) (void)hipFree(p);
    if (e->h_pin) (void)hipHostFree(e->h_pin);
    if (e->stream) (void)hipStreamDestroy(e->stream);
    delete e;
}

static int check_instance(emv_engine *e, const GraphRec &G, int64_t k, const int64_t *kk) {
    if (k < 2 || k > 30 || !kk) {
        eset(e, "k=%lld outside [2, 30]", (long long)k);
        return EMV_ERR_ARG;
    }
    for (int64_t j = 0; j < k; j++)
        if (kk[j] < 0 || kk[j] >= G.n) {
            eset(e, "terminal %lld out of range", (long long)kk[j]);
            return EMV_ERR_ARG;
        }
    // The singleton stage relaxes on 32-bit labels: a distance (at most
    // (n-1) * max w) plus an edge must stay below the all-ones INF.  The
    // int32 table is decided by the bound B = sum of the singleton rows'
    // maxima once they are known (advance: B >= 2^31 - 1 ends the request
    // with EMV_ERR_RANGE); group queries keep B small, since their rows hold
    // distances to the nearest member, not the reduction's M.
    if ((double)G.n * (double)G.maxw >= 4294967295.0 || (1ll << k) + G.n >= INT32_MAX) {
        eset(e, "graph distances exceed the engine's 32-bit labels (use emv_solve)");
        return EMV_ERR_RANGE;
    }
    return EMV_OK;
}

static int submit(emv_engine *e, int graph, int64_t k, const int64_t *kk, uint32_t flags,
                  int64_t *ticket, std::vector<std::vector<int32_t>> *groups = nullptr) {
    const GraphRec &G = e->graphs[graph];
    ECHK(check_instance(e, G, k, kk));
    const int64_t cells = 2 * ((int64_t)1 << k) * G.np;
    if ((cells + 255) / 256 * 256 > e->arena.cap) {
        eset(e, "request needs %lld table bytes, more than the arena's %lld", (long long)(cells * 4),
             (long long)(e->arena.cap * 4));
        return EMV_ERR_ADMISSION;
    }
    Inst I;
    I.ticket = e->next_ticket++;
    I.graph = graph;
    I.k = (int)k;
    I.kk.assign(kk, kk + k);
    if (groups) I.groups.swap(*groups);
    I.flags = flags;
    I.t_submit = now_ms();
    e->inst.push_back(I);
    e->queue.push_back(I.ticket);
    *ticket = I.ticket;
    return EMV_OK;
}

// ---------------------------------------------------------------- batch

extern "C" int emv_batch_create(int device, int64_t ninst, const emv_instance *in,
                                emv_engine **out) {
    if (!out || ninst < 1 || !in) return EMV_ERR_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return EMV_ERR_NODEV;
    if (device < 0 || device >= ndev) return EMV_ERR_ARG;
    emv_engine *e = nullptr;
    int rc = engine_open(device, &e);
    if (rc != EMV_OK) {
        emv_engine_destroy(e);
        return rc;
    }
    // every instance has its own graph, so no lane packing: 16-bit
    // vertex-wise chunks relax in pairs (EMV_PAIR=0: one by one)
    {
        const char *pe = getenv("EMV_PAIR");
        e->pair_chunks = !(pe && atoi(pe) == 0);
    }
    std::vector<std::vector<int64_t>> edges((size_t)ninst);
    std::vector<int64_t> ns((size_t)ninst);
    for (int64_t i = 0; i < ninst; i++) {
        if (in[i].n < 1 || in[i].m < 0 || (in[i].m > 0 && !in[i].edges)) {
            eset(e, "instance %lld: bad graph", (long long)i);
            *out = e;
            return EMV_ERR_ARG;
        }
        ns[i] = in[i].n;
        edges[i].assign(in[i].edges, in[i].edges + 3 * in[i].m);
    }
    *out = e;
    ECHK(upload_graphs(e, edges, ns));
    int64_t cells = 0;
    for (int64_t i = 0; i < ninst; i++) {
        ECHK(check_instance(e, e->graphs[i], in[i].k, in[i].kk));
        cells += (2 * ((int64_t)1 << in[i].k) * e->graphs[i].np + 255) / 256 * 256;
    }
    EHIP(e, hipMalloc((void **)&e->d_arena, sizeof(int32_t) * cells));
    e->arena.reset(cells);
    ECHK(ensure_slots(e, (int)ninst));
    for (int64_t i = 0; i < ninst; i++) {
        int64_t t;
        ECHK(submit(e, (int)i, in[i].k, in[i].kk, 0, &t));
    }
    return EMV_OK;
}

// Re-queues every instance of a batch (tables are rewritten by the solve).
static int batch_reset(emv_engine *e, uint32_t flags) {
    for (Inst &I : e->inst) {
        if (I.abase >= 0) e->arena.release(I.abase, I.acells);
        if (I.slot >= 0) e->free_slots.push_back(I.slot);
        I.abase = -1;
        I.slot = -1;
        I.state = ST_QUEUED;
        I.stage = 1;
        I.status = EMV_OK;
        I.flags = flags;
        I.pairs.clear();
        I.t_submit = now_ms();
    }
    e->queue.clear();
    for (Inst &I : e->inst) e->queue.push_back(I.ticket);
    return EMV_OK;
}

static int batch_solve_eager(emv_engine *e, uint32_t flags, emv_batch_stats *st) {
    ECHK(batch_reset(e, flags));
    e->rounds_per_step.clear();
    for (;;) {
        bool any_running = false;
        for (const Inst &I : e->inst)
            if (I.state != ST_DONE) any_running = true;
        if (!any_running) break;
        int used = 0;
        ECHK(step_eager(e, &used));
        e->rounds_per_step.push_back(used);
    }
    (void)st;
    return EMV_OK;
}

// Lockstep replay of steps 2..kmax captured in one graph.
static int batch_solve_graph(emv_engine *e, uint32_t flags, bool *ok) {
    *ok = false;
    ECHK(batch_reset(e, flags));
    int used = 0;
    ECHK(step_eager(e, &used));  // singletons + bounds, eager
    for (const Inst &I : e->inst)
        if (I.state == ST_DONE && I.status != EMV_OK) return EMV_OK;  // let eager report it
    // plans of the remaining lockstep steps
    std::vector<std::vector<int64_t>> runs;
    int kmax = 0;
    for (const Inst &I : e->inst) kmax = std::max(kmax, I.k);
    for (int m = 2; m <= kmax; m++) {
        std::vector<int64_t> run;
        for (const Inst &I : e->inst)
            if (I.state == ST_RUNNING && I.k >= m) run.push_back(I.ticket);
        runs.push_back(run);
    }
    // rounds per step from the last eager solve (+ margin)
    if ((int)e->rounds_per_step.size() < kmax) return EMV_OK;
    // the plans are a function of the instances' states after the singleton
    // step and of the engine's settings: a repeated solve reuses them (and
    // the descriptors already on the device)
    std::vector<int64_t> sig = {kmax, e->prm.relax_blocks_per_cu, e->prm.merge_rows_per_wave,
                                split_items(e), e->prm.dense_rounds, e->no_split ? 1 : 0,
                                e->no_pack ? 1 : 0};
    for (const Inst &I : e->inst)
        sig.insert(sig.end(), {(int64_t)I.state, (int64_t)I.stage, (int64_t)I.k,
                               (int64_t)I.graph, (int64_t)I.dbits, (int64_t)I.pbits,
                               (int64_t)I.tbits, I.abase, (int64_t)I.slot,
                               (int64_t)(I.flags & EMV_MERGE_1V)});
    std::vector<StepPlan> &plans = e->gplans;
    std::vector<char *> dbs(runs.size());
    if (!(sig == e->gplan_sig && plans.size() == runs.size())) {
        e->gplan_sig.clear();
        plans.assign(runs.size(), StepPlan());
        int64_t max_cells = 0, max_flags = 0, max_macc = 0;
        // build with the stage each instance will have at that step
        std::vector<int> saved_stage(e->inst.size());
        for (size_t t = 0; t < e->inst.size(); t++) saved_stage[t] = e->inst[t].stage;
        for (size_t s = 0; s < runs.size(); s++) {
            for (int64_t t : runs[s]) e->inst[t].stage = (int)s + 2;
            build_plan(e, runs[s], plans[s]);
            max_cells = std::max(max_cells, plans[s].cells);
            max_flags = std::max(max_flags, plans[s].flagbytes);
            max_macc = std::max(max_macc, plans[s].macc_cells * (plans[s].pbits / 8));
        }
        for (size_t t = 0; t < e->inst.size(); t++) e->inst[t].stage = saved_stage[t];
        ECHK(ensure_stage(e, max_cells, max_flags));
        ECHK(ensure_macc(e, max_macc));
        for (const StepPlan &P : plans) ECHK(check_plan(e, P));
        for (size_t s = 0; s < plans.size(); s++) {
            ECHK(desc_buffer(e, 1 + s, plans[s].bytes, &dbs[s]));
            EHIP(e, hipMemcpyAsync(dbs[s], plans[s].blob.data(), plans[s].bytes,
                                   hipMemcpyHostToDevice, e->stream));
        }
        e->gplan_sig = sig;
    } else {
        for (size_t s = 0; s < plans.size(); s++) dbs[s] = (char *)e->desc_bufs[1 + s];
    }
    if (e->conv_cap < plans.size()) {
        dfree(e->d_conv);
        e->conv_cap = 0;
        EHIP(e, hipMalloc((void **)&e->d_conv, sizeof(uint32_t) * std::max<size_t>(plans.size(), 1)));
        e->conv_cap = plans.size();
    }
    uint32_t *d_conv = e->d_conv;
    EHIP(e, hipStreamSynchronize(e->stream));
    // the replay's key: everything its nodes depend on
    std::vector<char> key;
    auto add_key = [&](const void *p, size_t n) {
        key.insert(key.end(), (const char *)p, (const char *)p + n);
    };
    const void *bufs[] = {e->d_D,    e->d_code, e->d_flag[0], e->d_flag[1], e->d_macc,
                          e->d_arena, e->d_any, e->d_cnt,     d_conv};
    add_key(bufs, sizeof(bufs));
    add_key(&e->prm, sizeof(e->prm));
    add_key(sig.data(), sizeof(int64_t) * sig.size());
    for (size_t s = 0; s < plans.size(); s++) {
        const int R = std::max(1, e->rounds_per_step[s + 1] + e->prm.graph_round_margin);
        add_key(&R, sizeof(R));
        add_key(&dbs[s], sizeof(dbs[s]));
    }
    if (!(e->gexec && key == e->gkey)) {
        if (e->gexec) (void)hipGraphExecDestroy(e->gexec);
        if (e->ggraph) (void)hipGraphDestroy(e->ggraph);
        e->gexec = nullptr;
        e->ggraph = nullptr;
        e->gkey.clear();
        if (hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal) != hipSuccess) {
            eset(e, "hipStreamBeginCapture failed");
            return EMV_ERR_HIP;
        }
        int rc = EMV_OK;
        for (size_t s = 0; s < plans.size() && rc == EMV_OK; s++) {
            rc = enqueue_front(e, dbs[s], plans[s]);
            const int R = std::max(1, e->rounds_per_step[s + 1] + e->prm.graph_round_margin);
            grelax_round(e, dbs[s], plans[s], 1, 0);
            for (int r = 1; r < R; r++) grelax_round(e, dbs[s], plans[s], 0, r);
            hipLaunchKernelGGL(k_copy_word, dim3(1), dim3(64), 0, e->stream, e->d_any, R, d_conv,
                               (int)s);
            if (rc == EMV_OK) rc = enqueue_back(e, dbs[s], plans[s]);
        }
        hipGraph_t graph = nullptr;
        hipError_t ce = hipStreamEndCapture(e->stream, &graph);
        if (rc != EMV_OK || ce != hipSuccess) {
            if (graph) (void)hipGraphDestroy(graph);
            if (rc == EMV_OK) {
                eset(e, "graph capture failed: %s", hipGetErrorString(ce));
                rc = EMV_ERR_HIP;
            }
            return rc;
        }
        hipGraphExec_t exec = nullptr;
        if (hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) != hipSuccess) {
            (void)hipGraphDestroy(graph);
            eset(e, "hipGraphInstantiate failed");
            return EMV_ERR_HIP;
        }
        e->ggraph = graph;
        e->gexec = exec;
        e->gkey.swap(key);
    }
    hipError_t ie = hipGraphLaunch(e->gexec, e->stream);
    if (ie == hipSuccess) ie = hipStreamSynchronize(e->stream);
    std::vector<uint32_t> conv(plans.size(), 1);
    if (ie == hipSuccess)
        ie = hipMemcpy(conv.data(), d_conv, sizeof(uint32_t) * plans.size(), hipMemcpyDeviceToHost);
    if (ie != hipSuccess) {
        eset(e, "graph replay failed: %s", hipGetErrorString(ie));
        return EMV_ERR_HIP;
    }
    for (uint32_t w : conv)
        if (w != 0) return EMV_OK;  // a step ran out of rounds: caller re-solves eagerly
    // bookkeeping the eager steps would have done
    for (size_t s = 0; s < plans.size(); s++) {
        e->steps++;
        e->relax_rounds += std::max(1, e->rounds_per_step[s + 1] + e->prm.graph_round_margin);
        std::vector<int64_t> done;
        for (int64_t t : runs[s]) {
            Inst &I = e->inst[t];
            I.stage++;
            if (I.stage > I.k) done.push_back(t);
        }
        ECHK(read_costs_and_finish(e, done));
    }
    *ok = true;
    return EMV_OK;
}

extern "C" int emv_batch_solve(emv_engine *e, uint32_t flags, int64_t *costs,
                               emv_batch_stats *st) {
    if (!e) return EMV_ERR_ARG;
    EHIP(e, hipSetDevice(e->device));
    const double t0 = now_ms();
    const int64_t steps0 = e->steps, rounds0 = e->relax_rounds;
    bool done = false;
    int graph_used = 0, graph_fallback = 0;
    if (e->prm.use_graph && !e->rounds_per_step.empty()) {
        ECHK(batch_solve_graph(e, flags, &done));
        if (done) graph_used = 1;
        else graph_fallback = 1;
    }
    if (!done) {
        // flags may be dirty after a replay that ran out of rounds
        if (graph_fallback && e->flag_bytes > 0) clear_flags(e);
        ECHK(batch_solve_eager(e, flags, st));
    }
    int rc = EMV_OK;
    for (size_t i = 0; i < e->inst.size(); i++) {
        if (costs) costs[i] = e->inst[i].cost;
        if (e->inst[i].status != EMV_OK && rc == EMV_OK) {
            rc = e->inst[i].status;
            eset(e, "instance %zu failed (%d)", i, rc);
        }
    }
    if (st) {
        memset(st, 0, sizeof(*st));
        st->total_ms = now_ms() - t0;
        st->steps = e->steps - steps0;
        st->relax_rounds = e->relax_rounds - rounds0;
        st->graph_used = graph_used;
        st->graph_fallback = graph_fallback;
        int64_t cells = 0;
        for (const Inst &I : e->inst) cells += ((int64_t)1 << I.k) * e->graphs[I.graph].n;
        st->dp_cells = cells;
    }
    return rc;
}

extern "C" int emv_engine_read_rows(emv_engine *e, int64_t ticket, int64_t X0, int64_t nrows,
                                    int64_t *f_out, int64_t *code_out) {
    if (!e || ticket < 0 || ticket >= (int64_t)e->inst.size()) return EMV_ERR_ARG;
    const Inst &I = e->inst[ticket];
    if (I.abase < 0 || I.state != ST_DONE) {
        eset(e, "no table kept for ticket %lld", (long long)ticket);
        return EMV_ERR_STATE;
    }
    const GraphRec &G = e->graphs[I.graph];
    const int64_t S2 = 1ll << I.k;
    if (X0 < 0 || nrows < 0 || X0 + nrows > S2) return EMV_ERR_ARG;
    EHIP(e, hipSetDevice(e->device));
    std::vector<int32_t> buf((size_t)G.np);
    for (int64_t r = 0; r < nrows; r++) {
        const int64_t X = X0 + r;
        if (f_out && I.tbits == 16) {
            const uint16_t *b16 = reinterpret_cast<const uint16_t *>(buf.data());
            EHIP(e, hipMemcpy(buf.data(), (const uint16_t *)(e->d_arena + I.abase) + X * G.np,
                              sizeof(uint16_t) * G.n, hipMemcpyDeviceToHost));
            for (int64_t v = 0; v < G.n; v++)
                f_out[r * G.n + v] = b16[v] == 0xFFFF ? INT64_MAX : (int64_t)b16[v];
        } else if (f_out) {
            EHIP(e, hipMemcpy(buf.data(), e->d_arena + I.abase + X * G.np, sizeof(int32_t) * G.n,
                              hipMemcpyDeviceToHost));
            for (int64_t v = 0; v < G.n; v++)
                f_out[r * G.n + v] = buf[v] == INT32_MAX ? INT64_MAX : (int64_t)buf[v];
        }
        if (code_out) {
            EHIP(e, hipMemcpy(buf.data(), e->d_arena + I.abase + S2 * G.np + X * G.np,
                              sizeof(int32_t) * G.n, hipMemcpyDeviceToHost));
            for (int64_t v = 0; v < G.n; v++) code_out[r * G.n + v] = buf[v];
        }
    }
    return EMV_OK;
}

extern "C" int emv_engine_pairs(emv_engine *e, int64_t ticket, int64_t *pairs, int64_t cap,
                                int64_t *npairs) {
    if (!e || !npairs || ticket < 0 || ticket >= (int64_t)e->inst.size()) return EMV_ERR_ARG;
    const Inst &I = e->inst[ticket];
    *npairs = (int64_t)I.pairs.size() / 2;
    if (pairs) {
        if (cap < *npairs) return EMV_ERR_CAPACITY;
        memcpy(pairs, I.pairs.data(), sizeof(int64_t) * I.pairs.size());
    }
    return EMV_OK;
}

extern "C" int emv_engine_release(emv_engine *e, int64_t ticket) {
    if (!e || ticket < 0 || ticket >= (int64_t)e->inst.size()) return EMV_ERR_ARG;
    Inst &I = e->inst[ticket];
    if (I.state == ST_DONE && I.abase >= 0) {
        e->arena.release(I.abase, I.acells);
        I.abase = -1;
    }
    return EMV_OK;
}

// --------------------------------------------------------------- server

extern "C" int emv_server_create(int device, int64_t n, int64_t m, const int64_t *edges,
                                 int64_t arena_bytes, int max_inflight, emv_engine **out) {
    if (!out || n < 1 || m < 0 || (m > 0 && !edges) || max_inflight < 1 || max_inflight > 4096)
        return EMV_ERR_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return EMV_ERR_NODEV;
    if (device < 0 || device >= ndev) return EMV_ERR_ARG;
    emv_engine *e = nullptr;
    int rc = engine_open(device, &e);
    *out = e;
    if (rc != EMV_OK) return rc;
    e->server = true;
    std::vector<std::vector<int64_t>> E(1);
    E[0].assign(edges, edges + 3 * m);
    std::vector<int64_t> ns(1, n);
    ECHK(upload_graphs(e, E, ns));
    if (arena_bytes <= 0) {
        size_t fr = 0, tot = 0;
        EHIP(e, hipMemGetInfo(&fr, &tot));
        // keep room for staging and the runtime
        arena_bytes = (int64_t)(fr * 0.6);
    }
    const int64_t cells = arena_bytes / 4 / 256 * 256;
    EHIP(e, hipMalloc((void **)&e->d_arena, sizeof(int32_t) * std::max<int64_t>(cells, 256)));
    e->arena.reset(cells);
    ECHK(ensure_slots(e, max_inflight));
    return EMV_OK;
}

extern "C" int emv_server_submit(emv_engine *e, int64_t k, const int64_t *kk, uint32_t flags,
                                 int64_t *ticket) {
    if (!e || !ticket) return EMV_ERR_ARG;
    return submit(e, 0, k, kk, flags, ticket);
}

extern "C" int emv_server_submit_groups(emv_engine *e, int64_t g, const int64_t *offsets,
                                        const int64_t *members, uint32_t flags, int64_t *ticket) {
    if (!e || !ticket || !offsets || !members) return EMV_ERR_ARG;
    if (g < 2 || g > 30) {
        eset(e, "groups=%lld outside [2, 30]", (long long)g);
        return EMV_ERR_ARG;
    }
    const GraphRec &G = e->graphs[0];
    std::vector<std::vector<int32_t>> groups((size_t)g);
    std::vector<int64_t> first((size_t)g);
    for (int64_t j = 0; j < g; j++) {
        if (offsets[j + 1] <= offsets[j] || offsets[j] < 0) {
            eset(e, "group %lld is empty", (long long)j);
            return EMV_ERR_ARG;
        }
        for (int64_t i = offsets[j]; i < offsets[j + 1]; i++) {
            if (members[i] < 0 || members[i] >= G.n) {
                eset(e, "member %lld of group %lld out of range", (long long)members[i], (long long)j);
                return EMV_ERR_ARG;
            }
            groups[(size_t)j].push_back((int32_t)members[i]);
        }
        first[(size_t)j] = members[offsets[j]];
    }
    return submit(e, 0, g, first.data(), flags, ticket, &groups);
}

extern "C" int emv_server_step(emv_engine *e, int64_t *running, int64_t *queued) {
    if (!e) return EMV_ERR_ARG;
    EHIP(e, hipSetDevice(e->device));
    int used = 0;
    ECHK(step_eager(e, &used));
    int64_t r = 0, q = 0;
    for (const Inst &I : e->inst) {
        if (I.state == ST_RUNNING) r++;
        if (I.state == ST_QUEUED) q++;
    }
    if (running) *running = r;
    if (queued) *queued = q;
    return EMV_OK;
}

extern "C" int emv_server_poll(emv_engine *e, int64_t ticket, emv_request_info *info) {
    if (!e || !info || ticket < 0 || ticket >= (int64_t)e->inst.size()) return EMV_ERR_ARG;
    const Inst &I = e->inst[ticket];
    info->state = I.state;
    info->status = I.status;
    info->cost = I.cost;
    info->k = I.k;
    info->stage = I.stage;
    info->queue_ms = (I.state == ST_QUEUED) ? 0.0 : I.t_admit - I.t_submit;
    info->latency_ms = (I.state == ST_DONE) ? I.t_done - I.t_submit : 0.0;
    info->npairs = (int64_t)I.pairs.size() / 2;
    info->reduced_cost = I.reduced_cost;
    info->groups = I.groups.empty() ? 0 : I.k;
    info->root = I.root;
    return EMV_OK;
}

extern "C" int emv_server_cancel(emv_engine *e, int64_t ticket) {
    if (!e || ticket < 0 || ticket >= (int64_t)e->inst.size()) return EMV_ERR_ARG;
    if (!e->server) {
        eset(e, "cancel: requests of a batch cannot be cancelled");
        return EMV_ERR_STATE;
    }
    Inst &I = e->inst[ticket];
    if (I.state == ST_DONE) {
        eset(e, "cancel: request %lld has finished", (long long)ticket);
        return EMV_ERR_STATE;
    }
    if (I.state == ST_QUEUED) {
        auto it = std::find(e->queue.begin(), e->queue.end(), ticket);
        if (it != e->queue.end()) e->queue.erase(it);
    } else {
        // Every launch of this request is on the engine's stream, so the
        // next request admitted to its table region or slot is ordered
        // after them.
        if (I.abase >= 0) e->arena.release(I.abase, I.acells);
        if (I.slot >= 0) e->free_slots.push_back(I.slot);
        I.abase = -1;
        I.slot = -1;
    }
    I.state = ST_DONE;
    I.status = EMV_ERR_CANCELLED;
    I.cost = INT64_MAX;
    I.reduced_cost = INT64_MAX;
    I.pairs.clear();
    I.t_done = now_ms();
    return EMV_OK;
}

extern "C" int emv_engine_fail_alloc(emv_engine *e, int64_t after) {
    if (!e || after < -1) return EMV_ERR_ARG;
    e->fail_alloc = after;
    return EMV_OK;
}

extern "C" int64_t emv_server_group_weight(const emv_engine *e) {
    return (e && !e->graphs.empty()) ? 1 + e->graphs[0].wsum : -1;
}

extern "C" int emv_engine_certify(emv_engine *e, int64_t ticket, emv_certificate *out) {
    if (!e || !out || ticket < 0 || ticket >= (int64_t)e->inst.size()) return EMV_ERR_ARG;
    memset(out, 0, sizeof(*out));
    out->first_X = out->first_v = -1;
    const Inst &I = e->inst[ticket];
    if (I.abase < 0 || I.state != ST_DONE || I.status != EMV_OK) {
        eset(e, "no table kept for ticket %lld", (long long)ticket);
        return EMV_ERR_STATE;
    }
    EHIP(e, hipSetDevice(e->device));
    const GraphRec &G = e->graphs[I.graph];
    const int64_t S2 = 1ll << I.k;
    std::vector<std::vector<int32_t>> srcs = I.groups;
    if (srcs.empty())
        for (int j = 0; j < I.k; j++) srcs.push_back(std::vector<int32_t>(1, I.kk[(size_t)j]));
    const auto t0 = std::chrono::steady_clock::now();
    unsigned long long cnt[4];
    EHIP(e, certify_table(e->stream, e->d_arena + I.abase, rows_layout(G.np), I.tbits / 8,
                          e->d_arena + I.abase + S2 * G.np, G.np, (int)G.n, I.k,
                          e->d_rp + G.voff, e->d_adj, srcs, I.kk, I.groups.empty() ? 0 : 1, cnt));
    out->ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    out->cells = S2 * G.n;
    out->bad_value = (int64_t)cnt[0];
    out->not_finite = (int64_t)cnt[1];
    out->bad_record = (int64_t)cnt[2];
    if (cnt[3] != ~0ull) {
        out->first_X = (int64_t)(cnt[3] >> 32);
        out->first_v = (int64_t)(cnt[3] & 0xFFFFFFFFull);
    }
    return EMV_OK;
}

extern "C" int64_t emv_server_free_bytes(const emv_engine *e) {
    return e ? e->arena.largest() * 4 : -1;
}

extern "C" int64_t emv_request_footprint(const emv_engine *e, int64_t k) {
    if (!e || e->graphs.empty() || k < 1 || k > 30) return -1;
    return 2 * ((int64_t)1 << k) * e->graphs[0].np * 4;
}
