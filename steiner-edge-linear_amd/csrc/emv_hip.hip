// emv_hip.hip -- host orchestration and C ABI of the gfx950 EMV Steiner DP.
//
// Solve sequence (reader-el.c:2141-2343 is the reference driver):
//   1. k single-source shortest paths in one staging chunk (lane t = source
//      kk[t], reader-el.c:1988-2031): one dense pull round, then flagged
//      rounds until no label drops.
//   2. Bound B = sum_t max_v dist(kk[t], v) picks the table type (int32 when
//      B < 2^31-1), the merge accumulator (32-bit when (2B << k) fits) and
//      the staging label width (32-bit when 2B + max w fits).
//   3. For each layer m = 2..k (reader-el.c:2033-2134), in batches of rows:
//        merge  -> vertex-major staging (label, packed (sum << k) | Y)
//        relax  -> dense pull round, then flagged rounds until no label drops
//        unpack -> records from the converged labels; table rows f[X], b[X]
//   4. Answer f[C][q]; optional device traceback.
#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <new>
#include <type_traits>
#include <vector>

#include "emv_hip.h"
#include "emv_blocks.hpp"
#include "emv_certify.hpp"
#include "emv_kernels.hpp"

using namespace emv;

struct emv_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    char err[512] = {0};
    int nxcd = 8;
    int num_cu = 256;
    int relax_nxcd = 1;  // relaxation sweep: 1 = grid-wide order, nxcd = per-XCD ranges
    int relax_wg = 64;   // vertex-wise rounds: workgroups per CU (EMV_RELAX_WG)
    std::vector<char> batch_split;  // per layer batch of a solve: merged by the split kernels
    // graph (CSR on the device)
    int64_t n = 0, m = 0;
    int32_t *d_rp = nullptr;
    Edge *d_adj = nullptr;
    Edge *d_ell = nullptr;  // padded fixed-degree copy (ell_deg > 0)
    int ell_deg = 0;
    // vertex blocks of the block-local relaxation (nblk > 0 when the graph
    // has the locality for it)
    int nblk = 0;
    int32_t *d_boff = nullptr, *d_bvert = nullptr, *d_brp = nullptr, *d_bext = nullptr,
            *d_vblk = nullptr;
    Edge *d_badj = nullptr;
    // in-block distances from border slots (bdw 16 / 32 bits; 0: none, sweeps only)
    void *d_bdist = nullptr;
    int64_t *d_bdoff = nullptr;
    int bdw = 0;
    size_t lds_block = 65536, lds_cu = 65536;  // LDS limits of the device
    bool blk_attr[3] = {false, false, false};  // dynamic-LDS attribute set, per label width
    int64_t maxw = 0;
    int sssp_bits = 64;  // label width of the last shortest-path chunk
    // EMV_TRACE: per-relaxation counters of the block kernel on stderr
    // (visits, sweeps, swept slots)
    unsigned long long *d_stats = nullptr;
    // table; the allocation is reused across solves
    void *d_f = nullptr;
    int32_t *d_b = nullptr;
    size_t f_cap = 0, b_cap = 0;
    int64_t np = 0;
    FLayout flay = rows_layout(0);  // cell layout of the table (tiles for 16/32-bit cells)
    FLayout fclay = rows_layout(0); // cell layout of the int32 copy of an int64 table
    int tbytes = 0;  // nonzero while a kept table is readable
    int64_t kept_cost = 0;  // answer cell of the last solve
    std::vector<int32_t> kept_kk;  // terminals of the last solve
    int64_t tk = 0;
    // relaxation workspace, sized in chunks of 64 rows
    void *d_D = nullptr;          // labels, 8 bytes per cell reserved
    void *d_code = nullptr;       // merge's packed words, 8 bytes per cell reserved
    bool merge_h2 = false;  // this solve merges 16-bit cells two vertices per lane
    int dense_rounds = 0;   // vertex-wise rounds of a relaxation that take every item
    bool merge_star = false;  // ... in stars of rows sharing a base set
    bool merge_cube = false;  // ... the small-subset layers from rows staged in LDS (cubes)
    bool merge_star32 = true; // 32-bit cells (int32 table, compact copy) merge in stars too
    bool chunk128 = true;     // 16-bit labels staged in chunks of 128 rows (packed pairs) ...
    int cw = 64;              // ... the current layer batch's chunk width in rows
    bool cube_attr = false;   // dynamic-LDS attribute of the cube kernel set
    std::vector<MergeCube> h_cubes;  // cubes of the last batch (kept for the async copy)
    std::vector<CubeRow> h_cube_rows;
    std::vector<uint32_t> h_cube_tab;  // staged-row offsets per cube (k_merge_cube_h2 rowtab)
    uint32_t *d_cube_tab = nullptr;
    size_t cube_tab_cap = 0;
    MergeCube *d_cubes = nullptr;
    CubeRow *d_cube_rows = nullptr;
    size_t cubes_cap = 0, cube_rows_cap = 0;
    double star_min_cells = 34359738368.0;  // from layer batches of this many pair cells on
    const uint32_t *h_xs = nullptr;         // host copy of the batch's rows (star building)
    std::vector<MergeStar> h_stars;         // stars of the last batch (kept for the async copy)
    MergeStar *d_stars = nullptr;
    size_t stars_cap = 0;
    std::vector<MergeGrid> h_grids;         // grids of the last batch (16-bit cells)
    MergeGrid *d_grids = nullptr;
    size_t grids_cap = 0;
    void *d_rspan = nullptr;  // per staged row: min and max merged value
    size_t rspan_cap = 0;
    // int64 tables: compact int32 copy of the rows (cell - row base) that the
    // merges read while every row fits
    int32_t *d_fcopy = nullptr;
    int64_t *d_fbase = nullptr;
    size_t fcopy_cap = 0, fbase_cap = 0;
    bool fcopy_on = false;
    uint32_t *d_flag[2] = {nullptr, nullptr};  // per (chunk, vertex) activity, alternating rounds
    uint32_t *d_q[2] = {nullptr, nullptr};    // worklists of flagged items, alternating rounds
    uint32_t qcap = 0;
    uint32_t *d_cnt = nullptr;                // 4 rotating worklist counters
    void *d_M = nullptr;                      // packed accumulators of a split merge
    size_t m_cap = 0;
    int64_t ws_chunks = 0;
    uint32_t *d_any = nullptr;               // 4 rotating "round has work" words
    unsigned long long *d_rowmax = nullptr;  // per singleton row
    int64_t *d_misc = nullptr;               // traceback result
    uint32_t *d_xs = nullptr;                // subset ids of a batch
    int64_t xs_cap = 0;
    int32_t *d_src = nullptr;                // shortest-path sources
    uint32_t *h_pin = nullptr;               // pinned readback of a work word
    unsigned long long *d_visits = nullptr;  // items visited by the layer rounds (EMV_COUNT_VISITS)
    bool count_visits = false;
    unsigned char *d_scratch = nullptr;      // this device's copy of `scratch` (emv_sssp_noalloc)
};

// ------------------------------------------------ emv_sssp_noalloc scratch
// Device memory of emv_sssp_noalloc, which calls no allocator: a file-scope
// array that every GPU the library runs on holds until the process ends.
// Layout for a graph of n vertices: labels uint64[n] at byte 0, the two
// round-parity flag arrays uint32[n] at 8n and 12n, four round words uint32
// at 16n.  Host code reaches it only through the address hipGetSymbolAddress
// gives for the current device (emv_ctx::d_scratch); every call on a device
// shares that device's array.
constexpr size_t kScratchBytes = (size_t)32 << 20;
constexpr int64_t kScratchMaxN = (int64_t)((kScratchBytes - 256) / 16);
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

// Labels INF except 0 at s; both flag arrays and the round words cleared
// (a fill kernel, so that no memset node is needed under capture).
__global__ void k_ss_init(uint64_t *__restrict__ lab, uint32_t *__restrict__ flg, int64_t n,
                          int64_t s, uint32_t *__restrict__ words) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        lab[i] = (i == s) ? 0ull : ~0ull;
        flg[i] = 0u;
        flg[n + i] = 0u;
    }
    if (blockIdx.x == 0 && threadIdx.x < 4) words[threadIdx.x] = 0u;
}

// One pull round of the single-source labels: a flagged vertex (every vertex
// when `dense`) takes the least neighbour label + w; a vertex whose label
// dropped flags its neighbours for the next round and raises
// words[(r + 1) & 3].  Round r returns at once when round r - 1 changed
// nothing, and clears words[(r + 2) & 3] for the round after next.  Labels
// only drop, so the rounds end at the shortest-path distances in any order.
__global__ __launch_bounds__(256) void k_ss_round(uint64_t *__restrict__ lab,
                                                  uint32_t *__restrict__ Fcur,
                                                  uint32_t *__restrict__ Fnext,
                                                  const int32_t *__restrict__ rp,
                                                  const Edge *__restrict__ adj, int64_t n, int r,
                                                  int dense, uint32_t *__restrict__ words) {
    if (blockIdx.x == 0 && threadIdx.x == 0) words[(r + 2) & 3] = 0u;
    if (!dense && words[r & 3] == 0u) return;
    bool raised = false;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += stride) {
        if (!dense) {
            if (Fcur[v] == 0u) continue;
            Fcur[v] = 0u;
        }
        const uint64_t cur = lab[v];
        uint64_t best = cur;
        for (int32_t e = rp[v]; e < rp[v + 1]; e++) {
            const uint64_t lu = lab[adj[e].v];
            if (lu != ~0ull && lu + (uint64_t)adj[e].w < best) best = lu + (uint64_t)adj[e].w;
        }
        if (best < cur) {
            lab[v] = best;
            raised = true;
            for (int32_t e = rp[v]; e < rp[v + 1]; e++) Fnext[adj[e].v] = 1u;
        }
    }
    if (raised) words[(r + 1) & 3] = 1u;
}

// Distances (INT64_MAX if unreached) and predecessors (the smallest
// neighbour u with d[u] + w(u,v) == d[v]; -1 at the source and when
// unreached), the rule of emv_sssp.
__global__ void k_ss_out(const uint64_t *__restrict__ lab, const int32_t *__restrict__ rp,
                         const Edge *__restrict__ adj, int64_t n, int64_t *__restrict__ d,
                         int64_t *__restrict__ pred) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += stride) {
        const uint64_t dv = lab[v];
        int64_t p = -1;
        if (dv != 0ull && dv != ~0ull) {
            for (int32_t e = rp[v]; e < rp[v + 1]; e++) {
                const int64_t u = adj[e].v;
                const uint64_t lu = lab[u];
                if (lu != ~0ull && lu + (uint64_t)adj[e].w == dv && (p < 0 || u < p)) p = u;
            }
        }
        if (d) d[v] = (dv == ~0ull) ? INT64_MAX : (int64_t)dv;
        if (pred) pred[v] = p;
    }
}

}  // namespace

namespace {

// A switch in the environment: set and not "0".
bool env_flag(const char *name) {
    const char *e = getenv(name);
    return e && atoi(e) != 0;
}

// A default-on switch turned off: set to "0".
bool env_off(const char *name) {
    const char *e = getenv(name);
    return e && atoi(e) == 0;
}

void set_err(emv_ctx *c, const char *fmt, ...) {
    if (!c) return;
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(c->err, sizeof(c->err), fmt, ap);
    va_end(ap);
}

#define HIPCHK(ctx, expr)                                                                   \
    do {                                                                                    \
        hipError_t _e = (expr);                                                             \
        if (_e != hipSuccess) {                                                             \
            set_err((ctx), "%s failed: %s (%s:%d)", #expr, hipGetErrorString(_e), __FILE__, \
                    __LINE__);                                                              \
            return (_e == hipErrorOutOfMemory || _e == hipErrorMemoryAllocation)            \
                       ? EMV_ERR_NOMEM                                                      \
                       : EMV_ERR_HIP;                                                       \
        }                                                                                   \
    } while (0)

#define CHK(expr)                    \
    do {                             \
        int _r = (expr);             \
        if (_r != EMV_OK) return _r; \
    } while (0)

int64_t round_up(int64_t x, int64_t a) { return (x + a - 1) / a * a; }

// Row stride of the table: n rounded up to 64-vertex tiles, plus one tile so
// that the stride is an odd multiple of 256 B.  The rows of one tile then
// rotate over the L2 and HBM channels instead of all landing on the channel
// that a power-of-two stride would pick.
int64_t row_stride(int64_t n) {
    int64_t np = round_up(n, kWave);
    if ((np / kWave) % 2 == 0) np += kWave;
    return np;
}

// The bits of x deposited into the set bits of mask, lowest first.
uint32_t deposit_host(uint32_t x, uint32_t mask) {
    uint32_t r = 0;
    for (uint32_t bb = 1; mask != 0; bb <<= 1) {
        const uint32_t low = mask & (~mask + 1u);
        if (x & bb) r |= low;
        mask ^= low;
    }
    return r;
}

int bits_for(uint64_t x) {  // number of bits to hold values 0..x
    int b = 0;
    while (b < 64 && (x >> b) != 0) b++;
    return b;
}

int64_t choose(int64_t n, int64_t r) {
    if (r < 0 || r > n) return 0;
    int64_t a = 1;
    for (int64_t i = 1; i <= r; i++) a = a * (n - r + i) / i;
    return a;
}

template <typename T>
void dfree(T *&p) {
    if (p) (void)hipFree((void *)p);
    p = nullptr;
}

// Grid of `items` workgroups rounded up to a multiple of the XCD count.
int grid_for(int64_t items, int nxcd, int64_t *per) {
    *per = (items + nxcd - 1) / nxcd;
    return (int)(*per * nxcd);
}

int free_table(emv_ctx *c) {
    dfree(c->d_f);
    dfree(c->d_b);
    c->f_cap = c->b_cap = 0;
    dfree(c->d_fcopy);
    dfree(c->d_fbase);
    c->fcopy_cap = c->fbase_cap = 0;
    c->fcopy_on = false;
    c->tbytes = 0;
    c->tk = 0;
    return EMV_OK;
}

void invalidate_table(emv_ctx *c) {
    c->tbytes = 0;
    c->tk = 0;
}

int ensure_table(emv_ctx *c, size_t fbytes, size_t bbytes) {
    if (c->f_cap < fbytes) {
        dfree(c->d_f);
        c->f_cap = 0;
        HIPCHK(c, hipMalloc(&c->d_f, fbytes));
        c->f_cap = fbytes;
    }
    if (c->b_cap < bbytes) {
        dfree(c->d_b);
        c->b_cap = 0;
        HIPCHK(c, hipMalloc((void **)&c->d_b, bbytes));
        c->b_cap = bbytes;
    }
    return EMV_OK;
}

int64_t chunk_cells(const emv_ctx *c) { return c->np * kWave; }

int ensure_fcopy(emv_ctx *c, int64_t cells, int64_t rows) {
    if (c->fcopy_cap < (size_t)cells * 4) {
        dfree(c->d_fcopy);
        c->fcopy_cap = 0;
        HIPCHK(c, hipMalloc((void **)&c->d_fcopy, (size_t)cells * 4));
        c->fcopy_cap = (size_t)cells * 4;
    }
    if (c->fbase_cap < (size_t)rows * 8) {
        dfree(c->d_fbase);
        c->fbase_cap = 0;
        HIPCHK(c, hipMalloc((void **)&c->d_fbase, (size_t)rows * 8));
        c->fbase_cap = (size_t)rows * 8;
    }
    return EMV_OK;
}

void free_ws(emv_ctx *c) {
    dfree(c->d_M);
    c->m_cap = 0;
    dfree(c->d_rspan);
    c->rspan_cap = 0;
    dfree(c->d_stars);
    c->stars_cap = 0;
    dfree(c->d_grids);
    c->grids_cap = 0;
    dfree(c->d_cubes);
    dfree(c->d_cube_rows);
    dfree(c->d_cube_tab);
    c->cubes_cap = c->cube_rows_cap = c->cube_tab_cap = 0;
    dfree(c->d_q[0]);
    dfree(c->d_q[1]);
    c->qcap = 0;
    dfree(c->d_D);
    dfree(c->d_code);
    dfree(c->d_flag[0]);
    dfree(c->d_flag[1]);
    c->ws_chunks = 0;
}

// Clears both activity-flag arrays (a completed relaxation leaves them clear;
// this also covers one that stopped on an error).
int clear_flags(emv_ctx *c) {
    const int64_t items = c->ws_chunks * c->np;
    int blocks = (int)std::min<int64_t>((items + kBlock - 1) / kBlock, 65536);
    for (int i = 0; i < 2; i++)
        hipLaunchKernelGGL(k_fill<uint32_t>, dim3(blocks), dim3(kBlock), 0, c->stream,
                           c->d_flag[i], items, 0u);
    HIPCHK(c, hipGetLastError());
    return EMV_OK;
}

int ensure_ws(emv_ctx *c, int64_t chunks) {
    if (chunks <= c->ws_chunks) return EMV_OK;
    free_ws(c);
    const int64_t cells = chunks * chunk_cells(c);
    const int64_t items = chunks * c->np;
    HIPCHK(c, hipMalloc(&c->d_D, sizeof(uint64_t) * cells));
    HIPCHK(c, hipMalloc(&c->d_code, sizeof(uint64_t) * cells));
    HIPCHK(c, hipMalloc((void **)&c->d_flag[0], sizeof(uint32_t) * items));
    HIPCHK(c, hipMalloc((void **)&c->d_flag[1], sizeof(uint32_t) * items));
    // Worklist rounds are off unless EMV_QCAP sets a capacity: every append
    // is an atomic on one counter, which costs more per round than scanning
    // the flags in order (measured on the 8-regular and grid workloads).
    int64_t qcap = 0;
    if (const char *qc = getenv("EMV_QCAP")) qcap = std::min<int64_t>(items, std::max(0ll, atoll(qc)));
    HIPCHK(c, hipMalloc((void **)&c->d_q[0], sizeof(uint32_t) * std::max<int64_t>(qcap, 1)));
    HIPCHK(c, hipMalloc((void **)&c->d_q[1], sizeof(uint32_t) * std::max<int64_t>(qcap, 1)));
    c->qcap = (uint32_t)qcap;
    c->ws_chunks = chunks;
    return clear_flags(c);
}

int ensure_xs(emv_ctx *c, int64_t count) {
    if (count <= c->xs_cap) return EMV_OK;
    dfree(c->d_xs);
    c->xs_cap = 0;
    HIPCHK(c, hipMalloc((void **)&c->d_xs, sizeof(uint32_t) * count));
    c->xs_cap = count;
    return EMV_OK;
}

// Relaxation of `nchunks` staged chunks: one dense pull round over every
// (chunk, vertex), then flagged rounds, launched in groups between polls of
// the "round has work" word, until a round flags nothing.  Rounds launched
// after convergence return at once.

template <typename DT, int DMAX, bool PK = false>
void launch_relax_round(emv_ctx *c, int64_t nitems, int dense, int r) {
    // no more workgroups than 64-item blocks of waves (small instances)
    const int64_t need = (nitems + (int64_t)kBlock - 1) / kBlock;
    int grid = (int)std::max<int64_t>(1, std::min<int64_t>((int64_t)c->num_cu * c->relax_wg, need));
    if (c->relax_nxcd > 1) grid = (grid + c->relax_nxcd - 1) / c->relax_nxcd * c->relax_nxcd;
    hipLaunchKernelGGL((k_relax_v<DT, DMAX, PK>), dim3(grid), dim3(kBlock), 0, c->stream,
                       (DT *)c->d_D, c->np, (int)c->n, c->d_rp, c->d_adj, c->d_ell,
                       nitems, dense, c->d_flag[r & 1], c->d_flag[(r + 1) & 1],
                       (const uint32_t *)c->d_q[r & 1], c->d_q[(r + 1) & 1], c->qcap, c->d_any,
                       c->d_cnt, r, c->relax_nxcd, c->dense_rounds,
                       c->count_visits ? c->d_visits : nullptr);
}

template <typename DT>
void relax_round(emv_ctx *c, int64_t nitems, int dense, int r) {
    if constexpr (std::is_same<DT, uint16_t>::value) {
        if (c->cw == 2 * kWave) {
            // chunks of 128 16-bit rows: packed pairs in 32-bit words
            switch (c->ell_deg) {
            case 4: launch_relax_round<uint32_t, 4, true>(c, nitems, dense, r); break;
            case 8: launch_relax_round<uint32_t, 8, true>(c, nitems, dense, r); break;
            case 16: launch_relax_round<uint32_t, 16, true>(c, nitems, dense, r); break;
            default: launch_relax_round<uint32_t, 0, true>(c, nitems, dense, r); break;
            }
            return;
        }
    }
    switch (c->ell_deg) {
    case 4: launch_relax_round<DT, 4>(c, nitems, dense, r); break;
    case 8: launch_relax_round<DT, 8>(c, nitems, dense, r); break;
    case 16: launch_relax_round<DT, 16>(c, nitems, dense, r); break;
    default: launch_relax_round<DT, 0>(c, nitems, dense, r); break;
    }
}

// The block-local kernel for labels of type DT, if the device gives a
// workgroup enough LDS for one block of labels.
template <typename DT>
int blk_slot() {
    return sizeof(DT) == 2 ? 0 : sizeof(DT) == 4 ? 1 : 2;
}

template <typename DT>
bool use_blocks(emv_ctx *c) {
    const size_t bytes = blk_lds_bytes<DT>();
    if (c->nblk == 0 || bytes > c->lds_block) return false;
    if (bytes > 65536 && !c->blk_attr[blk_slot<DT>()]) {
        if (hipFuncSetAttribute((const void *)k_relax_blk<DT, uint16_t>,
                                hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)bytes) != hipSuccess ||
            hipFuncSetAttribute((const void *)k_relax_blk<DT, uint32_t>,
                                hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)bytes) != hipSuccess) {
            (void)hipGetLastError();
            return false;
        }
        c->blk_attr[blk_slot<DT>()] = true;
    }
    return true;
}

bool use_blocks_any(emv_ctx *c, int dbits) {
    if (dbits == 16) return use_blocks<uint16_t>(c);
    if (dbits == 32) return use_blocks<uint32_t>(c);
    return use_blocks<uint64_t>(c);
}

// Full sweeps of a block's first visit (EMV_BLOCK_FULL_SWEEPS overrides the
// default; large values keep full sweeps to the end).
int blk_full_sweeps() {
    const char *e = getenv("EMV_BLOCK_FULL_SWEEPS");
    const int v = e ? atoi(e) : 0;
    return v > 0 ? v : kBlkFullSweeps;
}

template <typename DT, typename DW>
void blk_round_w(emv_ctx *c, int64_t nitems, int dense, int r) {
    const size_t bytes = blk_lds_bytes<DT>();
    const int per_cu = (int)std::max<size_t>(1, std::min<size_t>(2, c->lds_cu / bytes));
    const int grid = (int)std::max<int64_t>(1, std::min<int64_t>((int64_t)c->num_cu * per_cu, nitems));
    hipLaunchKernelGGL((k_relax_blk<DT, DW>), dim3(grid), dim3(kBT), bytes,
                       c->stream, (DT *)c->d_D, c->np, c->nblk, c->d_boff, c->d_bvert, c->d_brp,
                       c->d_bext, c->d_badj, c->d_vblk, (const DW *)c->d_bdist, c->d_bdoff,
                       nitems, dense, c->d_flag[r & 1], c->d_flag[(r + 1) & 1], c->d_any, r,
                       c->d_stats, blk_full_sweeps());
}

template <typename DT>
void blk_round(emv_ctx *c, int64_t nitems, int dense, int r) {
    if (c->bdw == 32)
        blk_round_w<DT, uint32_t>(c, nitems, dense, r);
    else
        blk_round_w<DT, uint16_t>(c, nitems, dense, r);
}

template <typename DT>
int relax_chunks(emv_ctx *c, int64_t nchunks, int64_t *rounds_out) {
    const bool blocks = use_blocks<DT>(c);
    const int64_t nitems = nchunks * (blocks ? (int64_t)c->nblk : c->n);
    hipLaunchKernelGGL(k_fill<uint32_t>, dim3(1), dim3(64), 0, c->stream, c->d_any, (int64_t)4,
                       0u);
    hipLaunchKernelGGL(k_fill<uint32_t>, dim3(1), dim3(64), 0, c->stream, c->d_cnt, (int64_t)4,
                       0u);
    int r = 0;
    if (blocks) {
        blk_round<DT>(c, nitems, 1, r);
        if (hipGetLastError() != hipSuccess) {
            // the device refused the block kernel's LDS: vertex-wise rounds
            c->lds_block = std::min(c->lds_block, blk_lds_bytes<DT>() - 1);
            return relax_chunks<DT>(c, nchunks, rounds_out);
        }
    } else {
        relax_round<DT>(c, nitems, 1, r);
    }
    HIPCHK(c, hipGetLastError());
    r = 1;
    int64_t rounds = 1;
    int group = 2;
    for (;;) {
        for (int g = 0; g < group; g++, r++) {
            if (blocks)
                blk_round<DT>(c, nitems, 0, r);
            else
                relax_round<DT>(c, nitems, 0, r);
            rounds++;
        }
        HIPCHK(c, hipGetLastError());
        HIPCHK(c, hipMemcpyAsync(c->h_pin, c->d_any + (r & 3), sizeof(uint32_t),
                                 hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        if (*c->h_pin == 0) break;
        group = std::min(group * 2, 32);
        if (rounds > (int64_t)2 * (c->n + 8)) {
            set_err(c, "relaxation did not converge");
            return EMV_ERR_RANGE;
        }
    }
    *rounds_out += rounds;
    if (c->d_stats && blocks) {
        unsigned long long st[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        HIPCHK(c, hipMemcpy(st, c->d_stats, sizeof(st), hipMemcpyDeviceToHost));
        HIPCHK(c, hipMemset(c->d_stats, 0, sizeof(st)));
        const double v = st[0] ? (double)st[0] : 1.0;
        fprintf(stderr,
                "emv-trace: relax chunks=%lld items=%lld rounds=%lld visits=%llu sweeps=%llu "
                "swept_slots=%llu idle_visits=%llu lowered_border=%llu cycles/visit load=%.0f "
                "sweep=%.0f write=%.0f lds=%zu/%zu grid=%d\n",
                (long long)nchunks, (long long)nitems, (long long)rounds, st[0], st[1], st[2],
                st[6], st[7], st[3] / v, st[4] / v, st[5] / v, c->lds_block, c->lds_cu,
                c->num_cu * (int)std::max<size_t>(
                                1, std::min<size_t>(2, c->lds_cu / blk_lds_bytes<DT>())));
    }
    return EMV_OK;
}

struct SolveState {
    int k;
    int64_t S2;  // 2^k
    int tbytes;
    int pbits;
    int dbits;
    bool rel32;  // 64-bit labels: try 32-bit labels relative to each row's base, per batch
};

// Rows per merge wave: few rows per wave when the batch has few rows, so
// that every tile still spreads over many waves.
// Rows with at least this many pairs take the split merge (EMV_SPLIT_PAIRS
// overrides kSplitPairs; measurements).
uint32_t split_pairs() {
    static const uint32_t v = [] {
        const char *e = getenv("EMV_SPLIT_PAIRS");
        const long x = e ? atol(e) : 0;
        return x > 0 ? (uint32_t)x : kSplitPairs;
    }();
    return v;
}

// EMV_MERGE_ROWS=<1|2|4|8> caps it (measurements).
int merge_rows_per_wave(int nx) {
    static const int cap = [] {
        const char *e = getenv("EMV_MERGE_ROWS");
        const int v = e ? atoi(e) : 8;
        return (v == 1 || v == 2 || v == 4) ? v : 8;
    }();
    int r = 1;
    if (nx >= 1024) r = 8;
    else if (nx >= 256) r = 4;
    else if (nx >= 64) r = 2;
    return std::min(r, cap);
}

// Row bases of the compact copy when the merges read it (int32 view).
template <typename T>
const int64_t *merge_fbase(const emv_ctx *c) {
    return (std::is_same<T, int32_t>::value && c->fcopy_on) ? c->d_fbase : nullptr;
}

// Cell layout of the cells a merge of T reads: the int32 copy's while an
// int64 table merges from it, else the table's.
template <typename T>
FLayout merge_layout(const emv_ctx *c) {
    return (std::is_same<T, int32_t>::value && c->fcopy_on) ? c->fclay : c->flay;
}

template <typename T>
bool use_merge_h2(const emv_ctx *c) {
    return std::is_same<T, uint16_t>::value && c->merge_h2;
}

template <typename T, typename P, typename DT, int XG, int CW>
int launch_merge_xg_cw(emv_ctx *c, const T *f, int nx, int64_t nchunks, int k) {
    const int nsub = (int)(nchunks * (CW / XG));
    if constexpr (std::is_same<T, uint16_t>::value) {
        if (use_merge_h2<T>(c)) {
            const int64_t ntiles = (c->np + 2 * kWave - 1) / (2 * kWave);
            int64_t per;
            const int grid = grid_for((ntiles * nsub + 3) / 4, c->nxcd, &per);
            hipLaunchKernelGGL((k_merge_h2<P, DT, XG, CW>), dim3(grid), dim3(kBlock), 0,
                               c->stream, f, c->flay, c->np, c->d_xs, nx, nsub, ntiles, k, (DT *)c->d_D,
                               (P *)c->d_code, c->nxcd, per);
            HIPCHK(c, hipGetLastError());
            return EMV_OK;
        }
    }
    const int64_t ntiles = c->np / kWave;
    int64_t per;
    const int grid = grid_for((ntiles * nsub + 3) / 4, c->nxcd, &per);
    hipLaunchKernelGGL((k_merge_v<T, P, DT, XG, CW>), dim3(grid), dim3(kBlock), 0, c->stream, f,
                       merge_layout<T>(c), c->np, c->d_xs, nx, nsub, ntiles, k, (DT *)c->d_D, (P *)c->d_code, c->nxcd,
                       per, merge_fbase<T>(c));
    HIPCHK(c, hipGetLastError());
    return EMV_OK;
}

// Chunks of 128 rows exist only for 16-bit labels with 32-bit packed words.
template <typename P, typename DT>
constexpr bool wide_chunks_ok() {
    return std::is_same<DT, uint16_t>::value && std::is_same<P, uint32_t>::value;
}

template <typename T, typename P, typename DT, int XG>
int launch_merge_xg(emv_ctx *c, const T *f, int nx, int64_t nchunks, int k) {
    if constexpr (wide_chunks_ok<P, DT>())
        if (c->cw == 2 * kWave) return launch_merge_xg_cw<T, P, DT, XG, 2 * kWave>(c, f, nx, nchunks, k);
    return launch_merge_xg_cw<T, P, DT, XG, kWave>(c, f, nx, nchunks, k);
}

// Accumulators M -> staging for the batch's nchunks chunks.
template <typename P, typename DT>
void launch_pack(emv_ctx *c, int nx, int64_t nchunks, int k) {
    const int64_t vtiles = c->np / kWave;
    if constexpr (wide_chunks_ok<P, DT>()) {
        if (c->cw == 2 * kWave) {
            hipLaunchKernelGGL((k_pack_v<P, DT, 2 * kWave>), dim3((unsigned)(2 * nchunks * vtiles)),
                               dim3(kBlock), 0, c->stream, (const P *)c->d_M, c->np, nx, k,
                               (DT *)c->d_D, (P *)c->d_code, vtiles);
            return;
        }
    }
    hipLaunchKernelGGL((k_pack_v<P, DT>), dim3((unsigned)(nchunks * vtiles)), dim3(kBlock), 0,
                       c->stream, (const P *)c->d_M, c->np, nx, k, (DT *)c->d_D, (P *)c->d_code,
                       vtiles);
}

// Split merge for layers whose rows have >= kSplitPairs pairs: fill the
// packed accumulators, fold the parts, pack into the staging.

template <typename T, typename P, typename DT>
int launch_merge_split(emv_ctx *c, const T *f, int nx, int64_t nchunks, int k, int m) {
    const size_t bytes = sizeof(P) * (size_t)nx * (size_t)c->np;
    if (c->m_cap < bytes) {
        dfree(c->d_M);
        c->m_cap = 0;
        HIPCHK(c, hipMalloc(&c->d_M, bytes));
        c->m_cap = bytes;
    }
    const int64_t cells = (int64_t)nx * c->np;
    const int fb = (int)std::min<int64_t>((cells + kBlock - 1) / kBlock, 1 << 20);
    hipLaunchKernelGGL(k_fill<P>, dim3(fb), dim3(kBlock), 0, c->stream, (P *)c->d_M, cells, ~(P)0);
    // enough items per tile to occupy an XCD's waves; parts of >= 4 outer steps
    const int64_t steps = (m - 1 <= 3) ? 1 : (1ll << (m - 1 - 3));
    const int64_t S = std::max<int64_t>(1, std::min<int64_t>((512 + nx - 1) / nx, steps / 4));
    bool h2 = false;
    if constexpr (std::is_same<T, uint16_t>::value) {
        if (use_merge_h2<T>(c)) {
            const int64_t ntiles = (c->np + 2 * kWave - 1) / (2 * kWave);
            int64_t per;
            const int grid = grid_for((ntiles * nx * S + 3) / 4, c->nxcd, &per);
            hipLaunchKernelGGL((k_merge_split_h2<P>), dim3(grid), dim3(kBlock), 0, c->stream, f,
                               c->flay, c->np, c->d_xs, nx, (int)S, ntiles, k, (P *)c->d_M, c->nxcd,
                               per);
            h2 = true;
        }
    }
    if (!h2) {
        const int64_t ntiles = c->np / kWave;
        int64_t per;
        const int grid = grid_for((ntiles * nx * S + 3) / 4, c->nxcd, &per);
        hipLaunchKernelGGL((k_merge_split<T, P>), dim3(grid), dim3(kBlock), 0, c->stream, f,
                           merge_layout<T>(c), c->np, c->d_xs, nx, (int)S, ntiles, k, (P *)c->d_M, c->nxcd, per,
                           merge_fbase<T>(c));
    }
    launch_pack<P, DT>(c, nx, nchunks, k);
    HIPCHK(c, hipGetLastError());
    return EMV_OK;
}

// EMV_MERGE_STAR=0: 16-bit merges one row per item (k_merge_h2 /
// k_merge_split_h2) instead of stars of rows sharing a base set.
bool merge_star_enabled() {
    static const bool on = [] {
        const char *e = getenv("EMV_MERGE_STAR");
        return !(e && atoi(e) == 0);
    }();
    return on;
}

// Waves a star launch gives each tile (its stars times their parts);
// EMV_STAR_WPT overrides 512.  At n = 2^18, k = 14 the merges take 114.6 /
// 114.5 ms at 256 / 512 and 119.5 / 121.1 ms at 128 / 1024 (one part per
// star: 129.4 ms).
int64_t star_waves_per_tile() {
    static const int64_t v = [] {
        const char *e = getenv("EMV_STAR_WPT");
        return e ? std::max<int64_t>(1, atoll(e)) : (int64_t)512;
    }();
    return v;
}

// Grids instead of stars for 16-bit cells (EMV_MERGE_GRID=0: stars; read
// per layer batch).
bool merge_grid_enabled() { return !env_off("EMV_MERGE_GRID"); }

// Layer batch of m >= 5 elements on 16-bit cells merged by grids
// (build_grids) into the accumulators, then packed into the staging.
template <typename P, typename DT>
int launch_merge_grid(emv_ctx *c, const uint16_t *f, int nx, int64_t nchunks, int k, int m) {
    build_grids(c->h_xs, nx, k, c->h_grids);
    const int ngrids = (int)c->h_grids.size();
    if (c->grids_cap < (size_t)ngrids) {
        dfree(c->d_grids);
        c->grids_cap = 0;
        HIPCHK(c, hipMalloc((void **)&c->d_grids, sizeof(MergeGrid) * (size_t)ngrids));
        c->grids_cap = (size_t)ngrids;
    }
    // h_grids is rebuilt only by the next batch's merge, after this batch's
    // relaxation has synchronised the stream
    HIPCHK(c, hipMemcpyAsync(c->d_grids, c->h_grids.data(), sizeof(MergeGrid) * (size_t)ngrids,
                             hipMemcpyHostToDevice, c->stream));
    const size_t bytes = sizeof(P) * (size_t)nx * (size_t)c->np;
    if (c->m_cap < bytes) {
        dfree(c->d_M);
        c->m_cap = 0;
        HIPCHK(c, hipMalloc(&c->d_M, bytes));
        c->m_cap = bytes;
    }
    const int64_t cells = (int64_t)nx * c->np;
    const int fb = (int)std::min<int64_t>((cells + kBlock - 1) / kBlock, 1 << 20);
    hipLaunchKernelGGL(k_fill<P>, dim3(fb), dim3(kBlock), 0, c->stream, (P *)c->d_M, cells, ~(P)0);
    // outer steps of a grid: subsets of S (m - 2 elements) without its two
    // lowest; parts as for the stars, each at least four outer steps
    const int64_t steps = 1ll << (m - 4);
    // waves per tile: EMV_STAR_WPT, else 256 (n = 2^18, k = 14: merges 85.7 /
    // 88.4 / 97.0 / 97.9 ms at 256 / 512 / 1,024 / 128; one part per grid
    // 104.7 ms)
    const int64_t wpt = getenv("EMV_STAR_WPT") ? star_waves_per_tile() : 256;
    const int64_t parts =
        std::max<int64_t>(1, std::min<int64_t>((wpt + ngrids - 1) / ngrids, steps / 4));
    const int64_t ntiles = (c->np + 2 * kWave - 1) / (2 * kWave);
    int64_t per;
    const int grid = grid_for((ntiles * ngrids * parts + 3) / 4, c->nxcd, &per);
    hipLaunchKernelGGL((k_merge_grid_h2<P>), dim3(grid), dim3(kBlock), 0, c->stream, f, c->flay,
                       c->np, (const MergeGrid *)c->d_grids, ngrids, (int)parts, ntiles, k,
                       (P *)c->d_M, c->nxcd, per);
    launch_pack<P, DT>(c, nx, nchunks, k);
    HIPCHK(c, hipGetLastError());
    return EMV_OK;
}

template <typename T, typename P, typename DT>
int launch_merge_star(emv_ctx *c, const T *f, int nx, int64_t nchunks, int k, int m) {
    if constexpr (std::is_same<T, uint16_t>::value)
        if (m >= 5 && merge_grid_enabled())
            return launch_merge_grid<P, DT>(c, f, nx, nchunks, k, m);
    // stars of up to kStarMax rows, but of at most four over the int32 copy
    // of an int64 table (row bases per pair: more rows cost it occupancy and
    // scalar registers; config 5's merges take 421 ms at four, 627 at eight)
    const bool rel = !std::is_same<T, uint16_t>::value && merge_fbase<T>(c) != nullptr;
    build_stars(c->h_xs, nx, k, c->h_stars, rel ? kStarRel : kStarMax);
    const int nstars = (int)c->h_stars.size();
    if (c->stars_cap < (size_t)nstars) {
        dfree(c->d_stars);
        c->stars_cap = 0;
        HIPCHK(c, hipMalloc((void **)&c->d_stars, sizeof(MergeStar) * (size_t)nstars));
        c->stars_cap = (size_t)nstars;
    }
    // h_stars is rebuilt only by the next batch's merge, after this batch's
    // relaxation has synchronised the stream
    HIPCHK(c, hipMemcpyAsync(c->d_stars, c->h_stars.data(), sizeof(MergeStar) * (size_t)nstars,
                             hipMemcpyHostToDevice, c->stream));
    const size_t bytes = sizeof(P) * (size_t)nx * (size_t)c->np;
    if (c->m_cap < bytes) {
        dfree(c->d_M);
        c->m_cap = 0;
        HIPCHK(c, hipMalloc(&c->d_M, bytes));
        c->m_cap = bytes;
    }
    const int64_t cells = (int64_t)nx * c->np;
    const int fb = (int)std::min<int64_t>((cells + kBlock - 1) / kBlock, 1 << 20);
    hipLaunchKernelGGL(k_fill<P>, dim3(fb), dim3(kBlock), 0, c->stream, (P *)c->d_M, cells, ~(P)0);
    // outer steps of a star: subsets of S's high part (S has m - 1 elements)
    const int64_t steps = (m - 1 <= 3) ? 1 : (1ll << (m - 1 - 3));
    // parts per star: about star_waves_per_tile() waves per tile, each taking
    // at least four outer steps
    const int64_t wpt = star_waves_per_tile();
    const int64_t parts =
        std::max<int64_t>(1, std::min<int64_t>((wpt + nstars - 1) / nstars, steps / 4));
    int64_t per;
    if constexpr (std::is_same<T, uint16_t>::value) {
        const int64_t ntiles = (c->np + 2 * kWave - 1) / (2 * kWave);
        const int grid = grid_for((ntiles * nstars * parts + 3) / 4, c->nxcd, &per);
        hipLaunchKernelGGL((k_merge_star_h2<P>), dim3(grid), dim3(kBlock), 0, c->stream, f,
                           c->flay, c->np,
                           (const MergeStar *)c->d_stars, nstars, (int)parts, ntiles, k,
                           (P *)c->d_M, c->nxcd, per);
    } else {
        // 32-bit cells (int32 table or compact copy): 64-vertex tiles
        const int64_t ntiles = c->np / kWave;
        const int grid = grid_for((ntiles * nstars * parts + 3) / 4, c->nxcd, &per);
        const int64_t *fb = merge_fbase<T>(c);
        if (fb)
            hipLaunchKernelGGL((k_merge_star_w<P, true, kStarRel>), dim3(grid), dim3(kBlock), 0, c->stream,
                               f, merge_layout<T>(c), fb, c->np, (const MergeStar *)c->d_stars,
                               nstars, (int)parts,
                               ntiles, k, (P *)c->d_M, c->nxcd, per);
        else
            hipLaunchKernelGGL((k_merge_star_w<P, false, kStarMax>), dim3(grid), dim3(kBlock), 0,
                               c->stream,
                               f, merge_layout<T>(c), fb, c->np, (const MergeStar *)c->d_stars,
                               nstars, (int)parts,
                               ntiles, k, (P *)c->d_M, c->nxcd, per);
    }
    launch_pack<P, DT>(c, nx, nchunks, k);
    HIPCHK(c, hipGetLastError());
    return EMV_OK;
}

// Cubes covering the batch's rows xs[0..nx) (all of m elements): each
// uncovered row, in order, goes with the u-element superset U (u = min(k,
// kCubeU)) that holds the most uncovered rows of the batch, and every
// uncovered row of U is assigned to it.  parts splits a row's pairs so that
// a cube has about two tasks per wave.
void build_cubes(const uint32_t *xs, int nx, int k, int m, std::vector<MergeCube> &cubes,
                 std::vector<CubeRow> &rows) {
    cubes.clear();
    rows.clear();
    const int u = std::min(k, kCubeU);
    std::vector<int32_t> idx((size_t)1 << k, -1);  // batch index of an uncovered row
    for (int i = 0; i < nx; i++) idx[xs[i]] = i;
    // the m-subsets of a u-set, as local masks (Gosper's order)
    std::vector<uint32_t> loc;
    for (uint32_t L = (1u << m) - 1u; L < (1u << u);) {
        loc.push_back(L);
        const uint32_t lo = L & (~L + 1u), r = L + lo;
        L = (((r ^ L) >> 2) / lo) | r;
    }
    for (int i = 0; i < nx; i++) {
        const uint32_t X = xs[i];
        if (idx[X] < 0) continue;
        // supersets U = X + (u - m) of the other elements
        std::vector<int> rest;
        for (int b = 0; b < k; b++)
            if (!(X >> b & 1u)) rest.push_back(b);
        const int need = u - m;
        uint32_t bestU = X;
        int best = -1;
        std::vector<int> pick(need);
        for (int q = 0; q < need; q++) pick[q] = q;
        for (;;) {
            uint32_t U = X;
            for (int q = 0; q < need; q++) U |= 1u << rest[pick[q]];
            int cnt = 0;
            for (uint32_t L : loc) {
                const uint32_t G = deposit_host(L, U);
                if (idx[G] >= 0) cnt++;
            }
            if (cnt > best) {
                best = cnt;
                bestU = U;
            }
            int q = need - 1;
            while (q >= 0 && pick[q] == (int)rest.size() - need + q) q--;
            if (q < 0) break;
            pick[q]++;
            for (int r2 = q + 1; r2 < need; r2++) pick[r2] = pick[r2 - 1] + 1;
        }
        MergeCube cb;
        memset(&cb, 0, sizeof(cb));
        cb.U = bestU;
        cb.e0 = (int32_t)rows.size();
        for (uint32_t L : loc) {
            const uint32_t G = deposit_host(L, bestU);
            if (idx[G] >= 0) {
                rows.push_back(CubeRow{L, idx[G]});
                idx[G] = -1;
            }
        }
        cb.nx = (int32_t)rows.size() - cb.e0;
        const int64_t npairs = ((int64_t)1 << (m - 1)) - 1;
        const int64_t want = (2 * kCubeWaves + cb.nx - 1) / cb.nx;
        cb.parts = (int32_t)std::max<int64_t>(1, std::min<int64_t>(want, npairs / 8));
        cubes.push_back(cb);
    }
}

// Whether the batch's layer takes the cube merge: 16-bit cells merged two
// vertices per lane, layers 3 .. u - 1 (a cube holds several rows of the
// layer), and a device that gives a workgroup the cube's LDS.
bool use_cube(emv_ctx *c, int k, int m) {
    static const int maxm = [] {  // EMV_CUBE_MAX_M: the largest layer merged in cubes
        const char *e = getenv("EMV_CUBE_MAX_M");
        return e ? atoi(e) : 64;
    }();
    const int u = std::min(k, kCubeU);
    if (!c->merge_cube || m < 3 || m > u - 1 || m > maxm) return false;
    (void)u;
    if ((size_t)kCubeLds > c->lds_block) return false;
    if (!c->cube_attr) {
        if (hipFuncSetAttribute((const void *)k_merge_cube_h2<uint32_t>,
                                hipFuncAttributeMaxDynamicSharedMemorySize, kCubeLds) !=
                hipSuccess ||
            hipFuncSetAttribute((const void *)k_merge_cube_h2<uint64_t>,
                                hipFuncAttributeMaxDynamicSharedMemorySize, kCubeLds) !=
                hipSuccess) {
            (void)hipGetLastError();
            c->merge_cube = false;
            return false;
        }
        c->cube_attr = true;
    }
    return true;
}

template <typename P, typename DT>
int launch_merge_cube(emv_ctx *c, const uint16_t *f, int nx, int64_t nchunks, int k, int m) {
    build_cubes(c->h_xs, nx, k, m, c->h_cubes, c->h_cube_rows);
    const int ncubes = (int)c->h_cubes.size();
    const size_t nrows = c->h_cube_rows.size();
    if (nrows != (size_t)nx) {
        set_err(c, "cube cover assigned %zu of %d rows", nrows, nx);
        return EMV_ERR_STATE;
    }
    // staged-row table: local row t = w + 16 q of cube g at [g][w][q], the
    // row's table offset in 128-byte units (rows are 2 np bytes, np a
    // multiple of 64), all ones for rows the merges never read
    constexpr int kPer = (1 << kCubeU) / kCubeWaves;
    const int64_t units = 2 * c->flay.rcells / 128;
    if ((((int64_t)1 << k) - 1) * units >= (int64_t)0xFFFFFFFF) {
        set_err(c, "table offsets exceed the cube table's 32 bits");
        return EMV_ERR_RANGE;
    }
    c->h_cube_tab.assign((size_t)ncubes << kCubeU, 0xFFFFFFFFu);
    for (int g = 0; g < ncubes; g++) {
        const uint32_t U = c->h_cubes[(size_t)g].U;
        const uint32_t nloc = 1u << __builtin_popcount(U);
        for (uint32_t t = 1; t < nloc; t++) {
            if (__builtin_popcount(t) >= m) continue;
            const uint32_t w = t % kCubeWaves, q = t / kCubeWaves;
            c->h_cube_tab[((size_t)g << kCubeU) + w * kPer + q] =
                (uint32_t)((int64_t)deposit_host(t, U) * units);
        }
    }
    if (c->cube_tab_cap < c->h_cube_tab.size()) {
        dfree(c->d_cube_tab);
        c->cube_tab_cap = 0;
        HIPCHK(c, hipMalloc((void **)&c->d_cube_tab, sizeof(uint32_t) * c->h_cube_tab.size()));
        c->cube_tab_cap = c->h_cube_tab.size();
    }
    HIPCHK(c, hipMemcpyAsync(c->d_cube_tab, c->h_cube_tab.data(),
                             sizeof(uint32_t) * c->h_cube_tab.size(), hipMemcpyHostToDevice,
                             c->stream));
    if (c->cubes_cap < (size_t)ncubes) {
        dfree(c->d_cubes);
        c->cubes_cap = 0;
        HIPCHK(c, hipMalloc((void **)&c->d_cubes, sizeof(MergeCube) * (size_t)ncubes));
        c->cubes_cap = (size_t)ncubes;
    }
    if (c->cube_rows_cap < nrows) {
        dfree(c->d_cube_rows);
        c->cube_rows_cap = 0;
        HIPCHK(c, hipMalloc((void **)&c->d_cube_rows, sizeof(CubeRow) * nrows));
        c->cube_rows_cap = nrows;
    }
    // h_cubes / h_cube_rows are rebuilt only by the next batch's merge, after
    // this batch's relaxation has synchronised the stream
    HIPCHK(c, hipMemcpyAsync(c->d_cubes, c->h_cubes.data(), sizeof(MergeCube) * (size_t)ncubes,
                             hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->d_cube_rows, c->h_cube_rows.data(), sizeof(CubeRow) * nrows,
                             hipMemcpyHostToDevice, c->stream));
    const size_t bytes = sizeof(P) * (size_t)nx * (size_t)c->np;
    if (c->m_cap < bytes) {
        dfree(c->d_M);
        c->m_cap = 0;
        HIPCHK(c, hipMalloc(&c->d_M, bytes));
        c->m_cap = bytes;
    }
    const int64_t cells = (int64_t)nx * c->np;
    const int fb = (int)std::min<int64_t>((cells + kBlock - 1) / kBlock, 1 << 20);
    hipLaunchKernelGGL(k_fill<P>, dim3(fb), dim3(kBlock), 0, c->stream, (P *)c->d_M, cells, ~(P)0);
    const int64_t ntiles = (c->np + 2 * kWave - 1) / (2 * kWave);
    // one persistent workgroup per CU (its LDS holds one), the items of each
    // XCD a contiguous range
    const int64_t items = ntiles * ncubes;
    const int64_t per = (items + c->nxcd - 1) / c->nxcd;
    const int grid = (int)std::max<int64_t>(
        c->nxcd, std::min<int64_t>((int64_t)c->num_cu, per * c->nxcd) / c->nxcd * c->nxcd);
    hipLaunchKernelGGL((k_merge_cube_h2<P>), dim3(grid), dim3(kCubeWaves * kWave),
                       (size_t)kCubeLds, c->stream, f, c->flay, c->np,
                       (const MergeCube *)c->d_cubes, (const CubeRow *)c->d_cube_rows,
                       (const uint32_t *)c->d_cube_tab, ncubes, ntiles, k, (P *)c->d_M, c->nxcd,
                       per);
    launch_pack<P, DT>(c, nx, nchunks, k);
    HIPCHK(c, hipGetLastError());
    return EMV_OK;
}

template <typename T, typename P, typename DT>
int launch_merge(emv_ctx *c, const T *f, int nx, int64_t nchunks, int k, int m) {
    if constexpr (std::is_same<T, uint16_t>::value) {
        if (use_merge_h2<T>(c) && c->h_xs && use_cube(c, k, m))
            return launch_merge_cube<P, DT>(c, f, nx, nchunks, k, m);
        // stars cost two more launches (fill, pack) and an accumulator pass
        // per row than whole rows: only layer batches of >= star_min_cells
        // pair cells take them (default 2^35; at n = 2^18, k = 14 the layers
        // from 7 elements on: layers 4-6 merged whole were 4-5 ms faster)
        const double pair_cells = (double)nx * (double)((1u << (m - 1)) - 1u) * (double)c->n;
        if (use_merge_h2<T>(c) && c->merge_star && c->h_xs && pair_cells >= c->star_min_cells)
            return launch_merge_star<T, P, DT>(c, f, nx, nchunks, k, m);
    }
    if constexpr (std::is_same<T, int32_t>::value) {
        // int32 table or the compact copy of an int64 table
        const double pair_cells = (double)nx * (double)((1u << (m - 1)) - 1u) * (double)c->n;
        if (c->merge_star32 && c->h_xs && pair_cells >= c->star_min_cells)
            return launch_merge_star<T, P, DT>(c, f, nx, nchunks, k, m);
    }
    if (((1u << (m - 1)) - 1u) >= split_pairs())
        return launch_merge_split<T, P, DT>(c, f, nx, nchunks, k, m);
    switch (merge_rows_per_wave(nx)) {
    case 8: return launch_merge_xg<T, P, DT, 8>(c, f, nx, nchunks, k);
    case 4: return launch_merge_xg<T, P, DT, 4>(c, f, nx, nchunks, k);
    case 2: return launch_merge_xg<T, P, DT, 2>(c, f, nx, nchunks, k);
    default: return launch_merge_xg<T, P, DT, 1>(c, f, nx, nchunks, k);
    }
}

// 64-bit labels: whether every row of the staged batch has merged values
// spanning less than 2^32 - 1 - max w.  Relaxation keeps each label between
// its row's minimum and maximum merged value, so the rows then relax exactly
// on 32-bit labels relative to the row minimum (INF = all ones stays out of
// reach of label + w).  If so, the relative labels replace the staging.
template <typename P>
int relative_labels_fit(emv_ctx *c, int nx, int64_t nchunks, int k, bool *fit, bool *copy_fit) {
    *fit = false;
    *copy_fit = false;
    const size_t need = 2 * sizeof(unsigned long long) * (size_t)nchunks * kWave;
    if (c->rspan_cap < need) {
        dfree(c->d_rspan);
        c->rspan_cap = 0;
        HIPCHK(c, hipMalloc(&c->d_rspan, need));
        c->rspan_cap = need;
    }
    unsigned long long *rs = (unsigned long long *)c->d_rspan;
    const int64_t rows = nchunks * kWave;
    hipLaunchKernelGGL(k_fill_span<unsigned long long>, dim3((unsigned)((rows + kBlock - 1) / kBlock)), dim3(kBlock), 0,
                       c->stream, rs, rows);
    const int64_t vper = 64;  // vertices per wave
    const int64_t vblocks = (c->n + vper * (kBlock / kWave) - 1) / (vper * (kBlock / kWave));
    hipLaunchKernelGGL((k_row_span<P>), dim3((unsigned)vblocks, (unsigned)nchunks), dim3(kBlock), 0,
                       c->stream, (const P *)c->d_code, c->np, (int)c->n, nx, k, rs, vper);
    HIPCHK(c, hipGetLastError());
    std::vector<unsigned long long> h((size_t)rows * 2);
    HIPCHK(c, hipMemcpyAsync(h.data(), rs, sizeof(unsigned long long) * h.size(),
                             hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    const unsigned long long lim = 0xFFFFFFFFull - 1ull - (unsigned long long)c->maxw;
    bool cfit = true;  // every final cell of the rows also fits the int32 compact copy
    for (int xi = 0; xi < nx; xi++) {
        if (h[2 * xi + 1] < h[2 * xi] || h[2 * xi + 1] - h[2 * xi] >= lim) return EMV_OK;
        if (h[2 * xi + 1] - h[2 * xi] >= (unsigned long long)INT32_MAX) cfit = false;
    }
    *copy_fit = cfit;
    const int64_t cells = nchunks * c->np * kWave;
    const int fb = (int)std::min<int64_t>((cells + kBlock - 1) / kBlock, 1 << 20);
    hipLaunchKernelGGL((k_rel_labels<P>), dim3(fb), dim3(kBlock), 0, c->stream,
                       (const P *)c->d_code, c->np, (int)c->n, nx, k,
                       (const unsigned long long *)rs, (uint32_t *)c->d_D, cells);
    HIPCHK(c, hipGetLastError());
    *fit = true;
    return EMV_OK;
}

template <typename T, typename P, typename DT>
int run_layers(emv_ctx *c, const SolveState &st, T *f, int32_t *b, emv_result *res,
               std::vector<hipEvent_t> &ev) {
    const int k = st.k;
    const int64_t np = c->np;
    const int64_t n = c->n;
    const int64_t rows_per_batch = std::min<int64_t>(c->ws_chunks * kWave, c->xs_cap);
    // 16-bit labels relaxed vertex-wise stage in chunks of 128 rows (one
    // 256-byte gather per neighbour); everything else in chunks of 64.  A
    // batch of rows_per_batch rows then needs at most as many bytes of staging
    // as 64-row chunks of 8-byte cells reserve (2 + 4 bytes per cell).
    int cw = kWave;
    if constexpr (wide_chunks_ok<P, DT>())
        if (c->chunk128 && !use_blocks<DT>(c)) cw = 2 * kWave;
    c->cw = cw;
    res->relax_chunk_rows = cw;
    std::vector<uint32_t> xs;
    for (int m = 2; m <= k; m++) {
        xs.clear();
        // the m-subsets in increasing numeric order (reader-el.c:2040-2046)
        for (uint64_t X = (1ull << m) - 1; X < (uint64_t)st.S2;) {
            xs.push_back((uint32_t)X);
            uint64_t lo = X & (~X + 1);
            uint64_t r = X + lo;
            X = (((r ^ X) >> 2) / lo) | r;
        }
        for (size_t off = 0; off < xs.size(); off += (size_t)rows_per_batch) {
            const int nx = (int)std::min<int64_t>(rows_per_batch, (int64_t)(xs.size() - off));
            const int64_t nchunks = (nx + cw - 1) / cw;
            HIPCHK(c, hipMemcpyAsync(c->d_xs, xs.data() + off, sizeof(uint32_t) * nx,
                                     hipMemcpyHostToDevice, c->stream));
            hipEvent_t e0, e1, e2, e3;
            for (hipEvent_t *pe : {&e0, &e1, &e2, &e3}) {
                HIPCHK(c, hipEventCreate(pe));
                ev.push_back(*pe);
            }
            c->batch_split.push_back(((1u << (m - 1)) - 1u) >= split_pairs() ? 1 : 0);
            HIPCHK(c, hipEventRecord(e0, c->stream));
            c->h_xs = xs.data() + off;
            // merge -> staging (int64 tables: from the compact copy while it holds)
            if (sizeof(T) == 8 && c->fcopy_on) {
                CHK((launch_merge<int32_t, P, DT>(c, c->d_fcopy, nx, nchunks, k, m)));
                res->compact_merge_layers++;
            } else {
                CHK((launch_merge<T, P, DT>(c, f, nx, nchunks, k, m)));
            }
            c->h_xs = nullptr;
            res->merge_launches++;
            HIPCHK(c, hipEventRecord(e1, c->stream));
            const int64_t vtiles = (n + kWave - 1) / kWave;
            int64_t per;
            const int grid = grid_for(nchunks * (cw / kWave) * vtiles, c->nxcd, &per);
            bool rel = false, copy_fit = false;
            if constexpr (sizeof(DT) == 8) {
                if (st.rel32) CHK(relative_labels_fit<P>(c, nx, nchunks, k, &rel, &copy_fit));
            }
            if (!rel || !copy_fit) c->fcopy_on = false;
            // a layer's relaxation starts near-dense: its first rounds take
            // every item (the single-source phase keeps flags from round 0)
            c->dense_rounds = layer_dense_rounds();
            if (rel) {
                // relax + records on 32-bit labels relative to each row's base
                CHK(relax_chunks<uint32_t>(c, nchunks, &res->relax_rounds));
                HIPCHK(c, hipEventRecord(e2, c->stream));
                hipLaunchKernelGGL((k_unpack_v<T, uint32_t, P>), dim3(grid), dim3(kBlock), 0,
                                   c->stream, (const uint32_t *)c->d_D, (const P *)c->d_code, np,
                                   c->d_xs, nx, (int)n, k, c->d_rp, c->d_adj, f, c->flay, b, c->nxcd,
                                   per, vtiles, (const unsigned long long *)c->d_rspan,
                                   c->fcopy_on ? c->d_fcopy : nullptr, c->fclay, c->d_fbase);
                res->relative_label_layers++;
            } else {
                CHK(relax_chunks<DT>(c, nchunks, &res->relax_rounds));
                HIPCHK(c, hipEventRecord(e2, c->stream));
                bool wide = false;
                if constexpr (wide_chunks_ok<P, DT>()) {
                    // (32-bit packed words imply a table of at most 4-byte cells)
                    if (cw == 2 * kWave && sizeof(T) <= 4) {
                        // one workgroup per (chunk, vertex tile)
                        int64_t wper;
                        const int wgrid = grid_for(nchunks * vtiles, c->nxcd, &wper);
                        hipLaunchKernelGGL((k_unpack_w<T, P>), dim3(wgrid),
                                           dim3(kWideWaves * kWave), 0, c->stream,
                                           (const uint32_t *)c->d_D, (const P *)c->d_code, np,
                                           c->d_xs, nx, (int)n, k, c->d_rp, c->d_adj, f, c->flay,
                                           b, c->nxcd, wper, vtiles);
                        wide = true;
                    } else if (cw == 2 * kWave) {
                        hipLaunchKernelGGL((k_unpack_v<T, DT, P, 2 * kWave>), dim3(grid),
                                           dim3(kBlock), 0, c->stream, (const DT *)c->d_D,
                                           (const P *)c->d_code, np, c->d_xs, nx, (int)n, k,
                                           c->d_rp, c->d_adj, f, c->flay, b, c->nxcd, per, vtiles,
                                           (const unsigned long long *)nullptr,
                                           (int32_t *)nullptr, c->fclay,
                                           (int64_t *)nullptr);
                        wide = true;
                    }
                }
                if (!wide)
                    hipLaunchKernelGGL((k_unpack_v<T, DT, P>), dim3(grid), dim3(kBlock), 0,
                                       c->stream, (const DT *)c->d_D, (const P *)c->d_code, np,
                                       c->d_xs, nx, (int)n, k, c->d_rp, c->d_adj, f, c->flay, b,
                                       c->nxcd, per, vtiles, (const unsigned long long *)nullptr,
                                       (int32_t *)nullptr, c->fclay,
                                           (int64_t *)nullptr);
            }
            c->dense_rounds = 0;
            HIPCHK(c, hipGetLastError());
            HIPCHK(c, hipEventRecord(e3, c->stream));
        }
    }
    c->cw = kWave;
    return EMV_OK;
}

template <typename T>
int run_layers_t(emv_ctx *c, const SolveState &st, T *f, int32_t *b, emv_result *res,
                 std::vector<hipEvent_t> &ev) {
    if (st.pbits == 32) {
        if (st.dbits == 16) return run_layers<T, uint32_t, uint16_t>(c, st, f, b, res, ev);
        if (st.dbits == 32) return run_layers<T, uint32_t, uint32_t>(c, st, f, b, res, ev);
        return run_layers<T, uint32_t, uint64_t>(c, st, f, b, res, ev);
    }
    if (st.dbits == 16) return run_layers<T, uint64_t, uint16_t>(c, st, f, b, res, ev);
    if (st.dbits == 32) return run_layers<T, uint64_t, uint32_t>(c, st, f, b, res, ev);
    return run_layers<T, uint64_t, uint64_t>(c, st, f, b, res, ev);
}

void free_blocks(emv_ctx *c) {
    dfree(c->d_boff);
    dfree(c->d_bvert);
    dfree(c->d_brp);
    dfree(c->d_bext);
    dfree(c->d_vblk);
    dfree(c->d_badj);
    dfree(c->d_bdist);
    dfree(c->d_bdoff);
    c->bdw = 0;
    c->nblk = 0;
}

template <typename T>
int upload(emv_ctx *c, T *&d, const std::vector<T> &h) {
    HIPCHK(c, hipMalloc((void **)&d, sizeof(T) * std::max<size_t>(1, h.size())));
    if (!h.empty()) HIPCHK(c, hipMemcpy(d, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice));
    return EMV_OK;
}

int load_csr(emv_ctx *c, int64_t n, int64_t m, const std::vector<int32_t> &rp,
             const std::vector<Edge> &adj, int64_t maxw) {
    HIPCHK(c, hipSetDevice(c->device));
    invalidate_table(c);
    // the table and the workspace are sized by np: drop them when it changes
    const int64_t np = row_stride(n);
    if (np != c->np) {
        free_table(c);
        free_ws(c);
    }
    dfree(c->d_rp);
    dfree(c->d_adj);
    dfree(c->d_ell);
    c->ell_deg = 0;
    c->n = 0;
    c->m = 0;
    // padded fixed-degree copy when the maximum degree is small and the
    // padding at most doubles the adjacency
    int64_t maxdeg = 0;
    for (int64_t u = 0; u < n; u++) maxdeg = std::max<int64_t>(maxdeg, rp[u + 1] - rp[u]);
    int ell = 0;
    for (int d : {4, 8, 16})
        if (maxdeg <= d) {
            ell = d;
            break;
        }
    if (ell && (int64_t)ell * n > 2 * std::max<int64_t>(1, 2 * m)) ell = 0;
    if (getenv("EMV_NO_ELL")) ell = 0;
    free_blocks(c);
    // Block-local relaxation when most edges stay inside their block (grids,
    // meshes).  EMV_BLOCKS=0 turns it off, EMV_BLOCKS=1 forces it.
    {
        const int mode = blocks_mode();
        if (mode != 0 && n >= 2 && !adj.empty()) {
            BlockPart P;
            const double inside = build_blocks(n, rp.data(), adj.data(), P);
            if (blocks_wanted(mode, n, inside)) {
                CHK(upload(c, c->d_boff, P.boff));
                CHK(upload(c, c->d_bvert, P.bvert));
                CHK(upload(c, c->d_brp, P.brp));
                CHK(upload(c, c->d_bext, P.bext));
                CHK(upload(c, c->d_vblk, P.vblk));
                CHK(upload(c, c->d_badj, P.badj));
                c->nblk = P.nblk;
                if (closure_mode() != 0) {
                    BlockDist B;
                    block_distances(P, B);
                    c->bdw = dist_width(B.dmax);
                    if (c->bdw == 16) {
                        uint16_t *p = nullptr;
                        CHK(upload(c, p, narrow<uint16_t>(B.d)));
                        c->d_bdist = p;
                    } else if (c->bdw == 32) {
                        uint32_t *p = nullptr;
                        CHK(upload(c, p, narrow<uint32_t>(B.d)));
                        c->d_bdist = p;
                    }
                    if (c->bdw) CHK(upload(c, c->d_bdoff, B.bdoff));
                }
            }
        }
    }
    if (ell) {
        std::vector<Edge> e((size_t)(n * ell), Edge{-1, 0});
        for (int64_t u = 0; u < n; u++)
            for (int32_t i = rp[u]; i < rp[u + 1]; i++) e[(size_t)(u * ell + (i - rp[u]))] = adj[i];
        HIPCHK(c, hipMalloc((void **)&c->d_ell, sizeof(Edge) * e.size()));
        HIPCHK(c, hipMemcpy(c->d_ell, e.data(), sizeof(Edge) * e.size(), hipMemcpyHostToDevice));
        c->ell_deg = ell;
    }
    HIPCHK(c, hipMalloc((void **)&c->d_rp, sizeof(int32_t) * (n + 1)));
    HIPCHK(c, hipMalloc((void **)&c->d_adj, sizeof(Edge) * std::max<size_t>(1, adj.size())));
    HIPCHK(c, hipMemcpy(c->d_rp, rp.data(), sizeof(int32_t) * (n + 1), hipMemcpyHostToDevice));
    if (!adj.empty())
        HIPCHK(c, hipMemcpy(c->d_adj, adj.data(), sizeof(Edge) * adj.size(),
                            hipMemcpyHostToDevice));
    c->n = n;
    c->m = m;
    c->maxw = maxw;
    c->np = np;
    return EMV_OK;
}

int check_weight(emv_ctx *c, int64_t w) {
    if (w < 1 || w > INT32_MAX) {
        set_err(c, "edge weight %lld outside [1, 2^31-1]", (long long)w);
        return EMV_ERR_RANGE;
    }
    return EMV_OK;
}

// Shortest paths from up to 64 sources in staging chunk 0: 32-bit labels
// when n * max w stays below the all-ones label, else 64-bit (sssp_bits).
int sssp_chunk(emv_ctx *c, const std::vector<int32_t> &src, int64_t *rounds) {
    const int ns = (int)src.size();
    if (ns > kWave) {
        set_err(c, "%d sources exceed %d", ns, kWave);
        return EMV_ERR_RANGE;
    }
    CHK(ensure_ws(c, 1));
    const int64_t cells = chunk_cells(c);
    const int blocks = (int)std::min<int64_t>((cells + kBlock - 1) / kBlock, 65536);
    dfree(c->d_src);
    HIPCHK(c, hipMalloc((void **)&c->d_src, sizeof(int32_t) * std::max(ns, 1)));
    HIPCHK(c, hipMemcpyAsync(c->d_src, src.data(), sizeof(int32_t) * ns, hipMemcpyHostToDevice,
                             c->stream));
    c->sssp_bits = ((uint64_t)c->n * (uint64_t)c->maxw < 0xFFFFFFFFull) ? 32 : 64;
    if (c->sssp_bits == 32) {
        hipLaunchKernelGGL(k_fill<uint32_t>, dim3(blocks), dim3(kBlock), 0, c->stream,
                           (uint32_t *)c->d_D, cells, ~0u);
        hipLaunchKernelGGL((k_seed_v<uint32_t>), dim3(1), dim3(kWave), 0, c->stream,
                           (uint32_t *)c->d_D, c->d_src, ns);
        HIPCHK(c, hipGetLastError());
        return relax_chunks<uint32_t>(c, 1, rounds);
    }
    hipLaunchKernelGGL(k_fill<uint64_t>, dim3(blocks), dim3(kBlock), 0, c->stream,
                       (uint64_t *)c->d_D, cells, ~0ull);
    hipLaunchKernelGGL((k_seed_v<uint64_t>), dim3(1), dim3(kWave), 0, c->stream,
                       (uint64_t *)c->d_D, c->d_src, ns);
    HIPCHK(c, hipGetLastError());
    return relax_chunks<uint64_t>(c, 1, rounds);
}

// Singleton rows (f == nullptr: only the per-row maxima) from chunk 0.
template <typename T>
void unpack_singletons(emv_ctx *c, int64_t k, T *f, int32_t *b) {
    const dim3 grid((unsigned)((c->n + kBlock - 1) / kBlock), (unsigned)k);
    if (c->sssp_bits == 32)
        hipLaunchKernelGGL((k_unpack_singletons_v<T, uint32_t>), grid, dim3(kBlock), 0, c->stream,
                           (const uint32_t *)c->d_D, c->np, c->d_src, (int)k, (int)c->n, f, c->flay,
                           b, c->d_rowmax);
    else
        hipLaunchKernelGGL((k_unpack_singletons_v<T, uint64_t>), grid, dim3(kBlock), 0, c->stream,
                           (const uint64_t *)c->d_D, c->np, c->d_src, (int)k, (int)c->n, f, c->flay,
                           b, c->d_rowmax);
}

double ev_ms(hipEvent_t a, hipEvent_t b) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, a, b) != hipSuccess) return 0.0;
    return (double)ms;
}

struct EventList {
    std::vector<hipEvent_t> v;
    ~EventList() {
        for (auto e : v) (void)hipEventDestroy(e);
    }
};

}  // namespace

// =================================================================== C ABI

extern "C" int emv_open(emv_ctx **out, int device) {
    if (!out) return EMV_ERR_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return EMV_ERR_NODEV;
    if (device < 0 || device >= ndev) return EMV_ERR_ARG;
    emv_ctx *c = new (std::nothrow) emv_ctx();
    if (!c) return EMV_ERR_NOMEM;
    c->device = device;
    int rc = EMV_OK;
    do {
        if (hipSetDevice(device) != hipSuccess) { rc = EMV_ERR_HIP; break; }
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, device) != hipSuccess) { rc = EMV_ERR_HIP; break; }
        c->num_cu = prop.multiProcessorCount;
        c->lds_block = prop.sharedMemPerBlock;
        c->lds_cu = std::max<size_t>(prop.maxSharedMemoryPerMultiProcessor, prop.sharedMemPerBlock);
        // MI355X: 8 XCDs of 32 CUs; workgroups are dealt round-robin over XCDs.
        c->nxcd = (c->num_cu % 32 == 0 && c->num_cu >= 64) ? c->num_cu / 32 : 1;
        const char *rx = getenv("EMV_RELAX_XCD");
        c->relax_nxcd = (rx && atoi(rx) > 0) ? c->nxcd : 1;
        if (const char *rw = getenv("EMV_RELAX_WG"))
            c->relax_wg = std::max(1, std::min(256, atoi(rw)));
        if (getenv("EMV_TRACE") && atoi(getenv("EMV_TRACE")) > 0) {
            if (hipMalloc((void **)&c->d_stats, 8 * sizeof(unsigned long long)) != hipSuccess ||
                hipMemset(c->d_stats, 0, 8 * sizeof(unsigned long long)) != hipSuccess) {
                rc = EMV_ERR_NOMEM;
                break;
            }
        }
        if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
            rc = EMV_ERR_HIP;
            break;
        }
        // this device's copy of the scratch array (emv_sssp_noalloc): the
        // array's name is a host-side stand-in, its device address comes
        // from the runtime, per device, with the device current
        void *sp = nullptr;
        if (hipGetSymbolAddress(&sp, HIP_SYMBOL(scratch)) != hipSuccess || !sp) {
            rc = EMV_ERR_HIP;
            break;
        }
        c->d_scratch = static_cast<unsigned char *>(sp);
        if (hipMalloc((void **)&c->d_any, sizeof(uint32_t) * 4) != hipSuccess ||
            hipMalloc((void **)&c->d_cnt, sizeof(uint32_t) * 4) != hipSuccess ||
            hipMalloc((void **)&c->d_rowmax, sizeof(unsigned long long) * 64) != hipSuccess ||
            hipMalloc((void **)&c->d_misc, sizeof(int64_t) * 8) != hipSuccess ||
            hipMalloc((void **)&c->d_visits, sizeof(unsigned long long)) != hipSuccess ||
            hipHostMalloc((void **)&c->h_pin, sizeof(uint32_t) * 16, hipHostMallocDefault) !=
                hipSuccess) {
            rc = EMV_ERR_NOMEM;
            break;
        }
    } while (0);
    if (rc != EMV_OK) {
        emv_close(c);
        return rc;
    }
    *out = c;
    return EMV_OK;
}

extern "C" void emv_close(emv_ctx *c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    free_table(c);
    free_ws(c);
    dfree(c->d_rp);
    dfree(c->d_adj);
    dfree(c->d_ell);
    free_blocks(c);
    dfree(c->d_stats);
    dfree(c->d_any);
    dfree(c->d_cnt);
    dfree(c->d_rowmax);
    dfree(c->d_misc);
    dfree(c->d_visits);
    dfree(c->d_xs);
    dfree(c->d_src);
    if (c->h_pin) (void)hipHostFree(c->h_pin);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

extern "C" const char *emv_last_error(const emv_ctx *c) { return c ? c->err : "no context"; }

extern "C" int emv_load_edges(emv_ctx *c, int64_t n, int64_t m, const int64_t *e) {
    if (!c) return EMV_ERR_ARG;
    if (n < 1 || n >= ((int64_t)1 << 30) || m < 0 || (m > 0 && !e) || 2 * m >= INT32_MAX) {
        set_err(c, "bad graph size n=%lld m=%lld", (long long)n, (long long)m);
        return EMV_ERR_ARG;
    }
    std::vector<int32_t> rp(n + 1, 0);
    int64_t maxw = 0;
    for (int64_t j = 0; j < m; j++) {
        int64_t u = e[3 * j], v = e[3 * j + 1], w = e[3 * j + 2];
        if (u < 0 || v < 0 || u >= n || v >= n) {
            set_err(c, "edge %lld endpoint out of range", (long long)j);
            return EMV_ERR_ARG;
        }
        CHK(check_weight(c, w));
        maxw = std::max(maxw, w);
        rp[u + 1]++;
        rp[v + 1]++;
    }
    for (int64_t u = 0; u < n; u++) rp[u + 1] += rp[u];
    std::vector<Edge> adj((size_t)(2 * m));
    std::vector<int32_t> fill(rp.begin(), rp.end() - 1);
    // both directions in edge-list order (reader-el.c:798-826)
    for (int64_t j = 0; j < m; j++) {
        int32_t u = (int32_t)e[3 * j], v = (int32_t)e[3 * j + 1], w = (int32_t)e[3 * j + 2];
        adj[fill[u]++] = Edge{v, w};
        adj[fill[v]++] = Edge{u, w};
    }
    return load_csr(c, n, m, rp, adj, maxw);
}

extern "C" int emv_load_graph(emv_ctx *c, int64_t n, int64_t m, const int64_t *pos,
                              const int64_t *adj) {
    if (!c) return EMV_ERR_ARG;
    if (n < 1 || n >= ((int64_t)1 << 30) || m < 0 || !pos || !adj || 2 * m >= INT32_MAX) {
        set_err(c, "bad graph n=%lld m=%lld", (long long)n, (long long)m);
        return EMV_ERR_ARG;
    }
    std::vector<int32_t> rp(n + 1, 0);
    int64_t total = 0, maxw = 0;
    for (int64_t u = 0; u < n; u++) {
        int64_t deg = adj[pos[u]];
        if (deg < 0) {
            set_err(c, "negative degree at vertex %lld", (long long)u);
            return EMV_ERR_ARG;
        }
        total += deg;
        rp[u + 1] = (int32_t)total;
    }
    if (total != 2 * m) {
        set_err(c, "adjacency holds %lld entries, expected 2m = %lld", (long long)total,
                (long long)(2 * m));
        return EMV_ERR_ARG;
    }
    std::vector<Edge> out((size_t)total);
    for (int64_t u = 0; u < n; u++) {
        const int64_t *a = adj + pos[u];
        for (int64_t i = 0; i < a[0]; i++) {
            int64_t v = a[1 + 2 * i], w = a[2 + 2 * i];
            if (v < 0 || v >= n) {
                set_err(c, "neighbour %lld of %lld out of range", (long long)v, (long long)u);
                return EMV_ERR_ARG;
            }
            CHK(check_weight(c, w));
            maxw = std::max(maxw, w);
            out[rp[u] + i] = Edge{(int32_t)v, (int32_t)w};
        }
    }
    return load_csr(c, n, m, rp, out, maxw);
}

extern "C" int64_t emv_solve_footprint(const emv_ctx *c, int64_t k, uint32_t flags) {
    if (!c || c->n <= 0 || k < 1 || k > 30) return -1;
    const int64_t np = row_stride(c->n);
    const int64_t S2 = (int64_t)1 << k;
    const int tb = (flags & EMV_FORCE_I64) ? 8 : 4;
    const int64_t chunks = std::max<int64_t>(1, (choose(k, k / 2) + kWave - 1) / kWave);
    const int64_t ws = chunks * np * (kWave * (8 + 8) + 8);
    return S2 * np * (tb + 4) + ws;
}

extern "C" int emv_release_table(emv_ctx *c) {
    if (!c) return EMV_ERR_ARG;
    (void)hipSetDevice(c->device);
    return free_table(c);
}

extern "C" int emv_solve(emv_ctx *c, int64_t k, const int64_t *kk, uint32_t flags,
                         emv_result *res) {
    if (!c || !res) return EMV_ERR_ARG;
    memset(res, 0, sizeof(*res));
    if (c->n <= 0 || !c->d_rp) {
        set_err(c, "no graph loaded");
        return EMV_ERR_STATE;
    }
    if (k < 1 || k > 30 || !kk) {
        set_err(c, "k=%lld outside [1,30]", (long long)k);
        return EMV_ERR_ARG;
    }
    const int64_t n = c->n;
    for (int64_t t = 0; t < k; t++)
        if (kk[t] < 0 || kk[t] >= n) {
            set_err(c, "terminal %lld out of range", (long long)kk[t]);
            return EMV_ERR_ARG;
        }
    HIPCHK(c, hipSetDevice(c->device));
    auto t_start = std::chrono::steady_clock::now();
    invalidate_table(c);
    c->count_visits = false;

    SolveState st;
    st.k = (int)k;
    st.S2 = (int64_t)1 << k;
    const int64_t np = c->np;
    if (st.S2 + n >= INT32_MAX) {
        set_err(c, "2^k + n = %lld does not fit the int32 record", (long long)(st.S2 + n));
        return EMV_ERR_RANGE;
    }
    // workspace for the largest layer, chunks of 64 rows
    const int64_t want_chunks = std::max<int64_t>(1, (choose(k, k / 2) + kWave - 1) / kWave);
    CHK(ensure_ws(c, want_chunks));
    CHK(clear_flags(c));
    CHK(ensure_xs(c, c->ws_chunks * kWave));

    EventList evl;
    std::vector<hipEvent_t> &ev = evl.v;
    hipEvent_t e_begin, e_single1, e_end, e_tb0, e_tb1, e_zero0, e_zero1;
    for (hipEvent_t *p : {&e_begin, &e_single1, &e_end, &e_tb0, &e_tb1, &e_zero0, &e_zero1}) {
        HIPCHK(c, hipEventCreate(p));
        ev.push_back(*p);
    }

    HIPCHK(c, hipEventRecord(e_begin, c->stream));
    // ---- 1. singleton shortest paths in chunk 0, lane t = kk[t]
    std::vector<int32_t> src((size_t)k);
    for (int64_t t = 0; t < k; t++) src[t] = (int32_t)kk[t];
    CHK(sssp_chunk(c, src, &res->singleton_rounds));
    // ---- 2. bound from the per-row maxima
    hipLaunchKernelGGL(k_fill<unsigned long long>, dim3(1), dim3(64), 0, c->stream, c->d_rowmax,
                       (int64_t)64, 0ull);
    unpack_singletons<int32_t>(c, k, nullptr, nullptr);
    HIPCHK(c, hipGetLastError());
    std::vector<unsigned long long> rowmax((size_t)k);
    HIPCHK(c, hipMemcpyAsync(rowmax.data(), c->d_rowmax, sizeof(unsigned long long) * k,
                             hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    unsigned long long B = 0;
    for (int64_t t = 0; t < k; t++) {
        if (rowmax[t] == ~0ull) {
            set_err(c, "terminal %lld does not reach every vertex (graph disconnected)",
                    (long long)kk[t] + 1);
            return EMV_ERR_DISCONNECTED;
        }
        if (rowmax[t] >= (1ull << 56)) {
            set_err(c, "distance bound overflows");
            return EMV_ERR_RANGE;
        }
        B += rowmax[t];
    }
    // Every cell obeys f[X][v] <= sum over t in X of dist(kk[t], v) <= B, so
    // a merge sum f[Y][v] + f[X\Y][v] and every stored label are at most B,
    // and a relaxation candidate label + w is at most B + max w.
    const unsigned long long top = B + (unsigned long long)c->maxw;
    if (bits_for(B) + k > 64 || top >= (1ull << 62)) {
        set_err(c, "distance bound %llu exceeds the 64-bit packing", B);
        return EMV_ERR_RANGE;
    }
    // table cells: 16-bit while every finite cell stays below 0xFFFF (row 0
    // holds 0xFFFF as INF), else int32, else int64
    st.tbytes = (B < (unsigned long long)INT32_MAX && !(flags & EMV_FORCE_I64)) ? 4 : 8;
    if (st.tbytes == 4 && B < 0xFFFFull && !(flags & EMV_TABLE32)) st.tbytes = 2;
    st.pbits = (st.tbytes <= 4 && bits_for(B) + k <= 32 && !(flags & EMV_FORCE_P64)) ? 32 : 64;
    // 16-bit tables in 128-vertex tiles of 2^k rows, int32 tables and the
    // int32 copy of an int64 table in 64-vertex tiles (256 bytes a row either
    // way); int64 tables in rows (EMV_TILE_TABLE=0: every table in rows)
    const bool tiles = !env_off("EMV_TILE_TABLE");
    c->flay = (tiles && st.tbytes == 2)   ? tiles_layout(k, 7)
              : (tiles && st.tbytes == 4) ? tiles_layout(k, 6)
                                          : rows_layout(c->np);
    c->fclay = tiles ? tiles_layout(k, 6) : rows_layout(c->np);
    c->merge_h2 = st.tbytes == 2 && !(flags & EMV_MERGE_1V) && merge_h2_ok(16, k, c->flay.rcells);
    c->merge_star = c->merge_h2 && !(flags & EMV_MERGE_ROWS) && merge_star_enabled();
    // cubes staged in LDS: opt-in (EMV_MERGE_CUBE or EMV_CUBE=1); the stars are
    // faster on every layer measured (n = 2^18, k = 14: layers 3-8 in 0.49 /
    // 1.44 / 3.2 / 6.7 / 15.6 / 39 ms against 0.35 / 1.38 / 2.6 / 4.7 / 9.0 /
    // 14.9 ms): a cube re-stages ~500 lower rows per tile from L2, while the
    // stars' repeated row reads hit L1
    c->merge_cube = c->merge_h2 && ((flags & EMV_MERGE_CUBE) || env_flag("EMV_CUBE"));
    {
        // EMV_STAR32=0: 32-bit cells merge row by row
        const char *s32 = getenv("EMV_STAR32");
        c->merge_star32 = !(flags & EMV_MERGE_ROWS) && !(s32 && atoi(s32) == 0);
        // EMV_CHUNK128=0: 16-bit labels in chunks of 64 rows
        const char *w = getenv("EMV_CHUNK128");
        c->chunk128 = !(w && atoi(w) == 0);
        c->cw = kWave;
    }
    {
        // EMV_STAR_MIN_CELLS: pair cells from which a layer batch merges in stars
        const char *e = getenv("EMV_STAR_MIN_CELLS");
        c->star_min_cells = e ? atof(e) : 34359738368.0;
    }
    st.rel32 = !(flags & EMV_FORCE_P64) && !env_flag("EMV_NO_REL32");
    if (flags & EMV_FORCE_P64)
        st.dbits = 64;
    else if (top < 0xFFFFull && !(flags & EMV_WIDE_LABELS))
        st.dbits = 16;
    else
        st.dbits = (top < 0xFFFFFFFFull) ? 32 : 64;

    // ---- table allocation + fill
    const int64_t cells = st.S2 * np;  // records (and table cells in rows)
    // tiles: ceil(np / 2^ts) blocks of 2^k rows of 2^ts cells
    const int64_t tw = c->flay.ts < 62 ? (int64_t)1 << c->flay.ts : 1;
    const int64_t vcells = (np + tw - 1) / tw * tw;
    HIPCHK(c, hipEventRecord(e_zero0, c->stream));
    CHK(ensure_table(c, (size_t)(st.S2 * vcells) * st.tbytes, (size_t)cells * 4));
    {
        // Every cell of rows 1 .. 2^k-1 at v < n is written by the solve, and
        // the padding columns v >= n only feed lanes whose results are
        // dropped, so only row 0 (never written; INF and record -1 as the
        // reference's initialisation, reader-el.c:2244-2259) is filled.
        if (st.tbytes == 2)
            hipLaunchKernelGGL(k_fill_row<uint16_t>, dim3(4), dim3(kBlock), 0, c->stream,
                               (uint16_t *)c->d_f, c->flay, (int64_t)0, vcells, (uint16_t)0xFFFF);
        else if (st.tbytes == 4)
            hipLaunchKernelGGL(k_fill_row<int32_t>, dim3(4), dim3(kBlock), 0, c->stream,
                               (int32_t *)c->d_f, c->flay, (int64_t)0, vcells, INT32_MAX);
        else
            hipLaunchKernelGGL(k_fill<int64_t>, dim3(4), dim3(kBlock), 0, c->stream,
                               (int64_t *)c->d_f, np, INT64_MAX);
        hipLaunchKernelGGL(k_fill<int32_t>, dim3(4), dim3(kBlock), 0, c->stream, c->d_b, np, -1);
        HIPCHK(c, hipGetLastError());
    }
    HIPCHK(c, hipEventRecord(e_zero1, c->stream));
    if (st.tbytes == 2)
        unpack_singletons<uint16_t>(c, k, (uint16_t *)c->d_f, c->d_b);
    else if (st.tbytes == 4)
        unpack_singletons<int32_t>(c, k, (int32_t *)c->d_f, c->d_b);
    else
        unpack_singletons<int64_t>(c, k, (int64_t *)c->d_f, c->d_b);
    HIPCHK(c, hipGetLastError());
    c->fcopy_on = false;
    if (st.tbytes == 8 && st.dbits == 64 && st.rel32 && !env_flag("EMV_NO_COPY32")) {
        // compact int32 copy for the merges: singleton rows as they are
        // (base 0) while every distance fits
        bool ok = true;
        for (int64_t t = 0; t < k; t++) ok = ok && rowmax[t] < (unsigned long long)INT32_MAX;
        if (ok) {
            CHK(ensure_fcopy(c, cells, st.S2));
            hipLaunchKernelGGL(k_fill<int64_t>, dim3((unsigned)std::min<int64_t>((st.S2 + kBlock - 1) / kBlock, 4096)),
                               dim3(kBlock), 0, c->stream, c->d_fbase, st.S2, (int64_t)0);
            hipLaunchKernelGGL(k_copy_singletons<int32_t>, dim3((unsigned)((n + kBlock - 1) / kBlock), (unsigned)k),
                               dim3(kBlock), 0, c->stream, (const int64_t *)c->d_f, np, (int)n, (int)k,
                               c->d_fcopy, c->fclay);
            HIPCHK(c, hipGetLastError());
            c->fcopy_on = true;
        }
    }
    HIPCHK(c, hipEventRecord(e_single1, c->stream));
    c->tbytes = st.tbytes;
    c->tk = k;

    // ---- 3. layers
    c->count_visits = (flags & EMV_COUNT_VISITS) != 0;
    if (c->count_visits)
        hipLaunchKernelGGL(k_fill<unsigned long long>, dim3(1), dim3(64), 0, c->stream, c->d_visits,
                           (int64_t)1, 0ull);
    const size_t ev_layers0 = ev.size();
    c->batch_split.clear();
    if (st.tbytes == 2)
        CHK(run_layers_t<uint16_t>(c, st, (uint16_t *)c->d_f, c->d_b, res, ev));
    else if (st.tbytes == 4)
        CHK(run_layers_t<int32_t>(c, st, (int32_t *)c->d_f, c->d_b, res, ev));
    else
        CHK(run_layers_t<int64_t>(c, st, (int64_t *)c->d_f, c->d_b, res, ev));
    HIPCHK(c, hipEventRecord(e_end, c->stream));
    res->relax_item_visits = -1;
    if (c->count_visits) {
        unsigned long long nv = 0;
        HIPCHK(c, hipMemcpyAsync(&nv, c->d_visits, sizeof(nv), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        res->relax_item_visits = (int64_t)nv;
        c->count_visits = false;
    }

    // ---- 4. answer
    const int64_t C = (k == 1) ? 0 : (((int64_t)1 << (k - 1)) - 1);
    const int64_t q = kk[k - 1];
    int64_t cost = 0;
    if (C == 0) {
        cost = INT64_MAX;  // row 0 is never written (reader-el.c:2137 with k == 1)
    } else if (st.tbytes == 2) {
        uint16_t v16;
        HIPCHK(c, hipMemcpyAsync(&v16, (uint16_t *)c->d_f + c->flay.cell(C, q), 2,
                                 hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        cost = v16;
    } else if (st.tbytes == 4) {
        int32_t v32;
        HIPCHK(c, hipMemcpyAsync(&v32, (int32_t *)c->d_f + c->flay.cell(C, q), 4,
                                 hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        cost = v32;
    } else {
        HIPCHK(c, hipMemcpyAsync(&cost, (int64_t *)c->d_f + C * np + q, 8, hipMemcpyDeviceToHost,
                                 c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
    }

    // ---- traceback
    HIPCHK(c, hipEventRecord(e_tb0, c->stream));
    if ((flags & EMV_LIST) && C != 0) {
        // an optimal tree emits at most n-1 edges plus k shortcut pairs; the
        // pending-frame stack holds at most one frame per terminal split
        const int64_t cap = 2 * (n + k) + 64;
        const int64_t stcap = 4 * k + 64;
        int64_t *d_pairs = nullptr, *d_stack = nullptr;
        HIPCHK(c, hipMalloc((void **)&d_pairs, sizeof(int64_t) * 2 * cap));
        if (hipMalloc((void **)&d_stack, sizeof(int64_t) * 2 * stcap) != hipSuccess) {
            (void)hipFree(d_pairs);
            set_err(c, "traceback stack allocation failed");
            return EMV_ERR_NOMEM;
        }
        hipLaunchKernelGGL(k_traceback, dim3(1), dim3(64), 0, c->stream, c->d_b, np, (int)k, q, C,
                           d_pairs, cap, d_stack, stcap, c->d_misc);
        int64_t cnt = 0;
        hipError_t e1 = hipMemcpyAsync(&cnt, c->d_misc, sizeof(int64_t), hipMemcpyDeviceToHost,
                                       c->stream);
        hipError_t e2 = hipStreamSynchronize(c->stream);
        if (e1 != hipSuccess || e2 != hipSuccess || cnt < 0) {
            (void)hipFree(d_pairs);
            (void)hipFree(d_stack);
            set_err(c, "traceback failed (%lld)", (long long)cnt);
            return cnt < 0 ? EMV_ERR_CAPACITY : EMV_ERR_HIP;
        }
        res->npairs = cnt;
        res->pairs = (int64_t *)malloc(sizeof(int64_t) * 2 * std::max<int64_t>(cnt, 1));
        hipError_t e3 = hipSuccess;
        if (cnt > 0)
            e3 = hipMemcpy(res->pairs, d_pairs, sizeof(int64_t) * 2 * cnt, hipMemcpyDeviceToHost);
        (void)hipFree(d_pairs);
        (void)hipFree(d_stack);
        HIPCHK(c, e3);
    } else if (flags & EMV_LIST) {
        res->npairs = 0;
        res->pairs = (int64_t *)malloc(sizeof(int64_t) * 2);
    }
    HIPCHK(c, hipEventRecord(e_tb1, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));

    // ---- timings and traffic models
    res->cost = cost;
    res->k = (int32_t)k;
    res->n = n;
    res->table_bytes = st.tbytes;
    res->merge_bits = st.pbits;
    res->label_bits = st.dbits;
    res->relax_blocks = use_blocks_any(c, st.dbits) ? c->nblk : 0;
    res->block_dist_bits = res->relax_blocks ? c->bdw : 0;
    res->dp_cells = st.S2 * n;
    res->zero_ms = ev_ms(e_zero0, e_zero1);
    res->singleton_ms = ev_ms(e_begin, e_single1) - res->zero_ms;
    for (size_t i = ev_layers0; i + 3 < ev.size(); i += 4) {
        res->merge_ms += ev_ms(ev[i], ev[i + 1]);
        const size_t bi = (i - ev_layers0) / 4;
        if (bi < c->batch_split.size() && c->batch_split[bi])
            res->merge_split_ms += ev_ms(ev[i], ev[i + 1]);
        res->relax_ms += ev_ms(ev[i + 1], ev[i + 2]);
        res->unpack_ms += ev_ms(ev[i + 2], ev[i + 3]);
    }
    res->kernel_ms = ev_ms(e_begin, e_end) - res->zero_ms;
    res->traceback_ms = ev_ms(e_tb0, e_tb1);
    const double cb = st.tbytes, ab = 4.0, dn = (double)n, dm = (double)c->m;
    double p3 = 1.0;
    for (int64_t i = 0; i < k; i++) p3 *= 3.0;
    res->merge_alg_bytes = cb * dn * (p3 - 2.0 * (double)st.S2 + 1.0) +
                           (cb + ab) * dn * ((double)st.S2 - (double)k - 1.0);
    res->relax_alg_bytes = ((double)st.S2 - 1.0) * (cb * dn + cb * 2 * dm + (cb + ab) * dn) +
                           (double)k * (4.0 * (dn + 1) + 16.0 * dm);
    res->device_bytes = (int64_t)(c->f_cap + c->b_cap + c->fcopy_cap + c->fbase_cap) +
                        c->ws_chunks * np * (int64_t)(kWave * 16 + 8);
    auto t_end = std::chrono::steady_clock::now();
    res->total_ms = std::chrono::duration<double, std::milli>(t_end - t_start).count();
    c->kept_cost = cost;
    c->kept_kk.assign(src.begin(), src.end());
    if (!(flags & EMV_KEEP_TABLE)) invalidate_table(c);
    return EMV_OK;
}

extern "C" int emv_write_table(emv_ctx *c, const char *path, uint32_t flags) {
    if (!c || !path) return EMV_ERR_ARG;
    if (!c->d_f || c->tk == 0) {
        set_err(c, "no table kept (solve with EMV_KEEP_TABLE)");
        return EMV_ERR_STATE;
    }
    FILE *out = fopen(path, "wb");
    if (!out) {
        set_err(c, "cannot open '%s' for writing", path);
        return EMV_ERR_ARG;
    }
    const int64_t k = c->tk, n = c->n, S2 = (int64_t)1 << k;
    const int records = (flags & EMV_DUMP_RECORDS) ? 1 : 0;
    unsigned char hdr[64];
    memset(hdr, 0, sizeof(hdr));
    memcpy(hdr, "EMVTABLE", 8);
    const uint32_t version = 1, cell_bytes = 8, rec = (uint32_t)records;
    const int64_t inf = INT64_MAX, cost = c->kept_cost;
    memcpy(hdr + 8, &version, 4);
    memcpy(hdr + 12, &cell_bytes, 4);
    memcpy(hdr + 16, &k, 8);
    memcpy(hdr + 24, &n, 8);
    memcpy(hdr + 32, &inf, 8);
    memcpy(hdr + 40, &rec, 4);
    memcpy(hdr + 48, &cost, 8);
    bool ok = fwrite(hdr, 1, sizeof(hdr), out) == sizeof(hdr);
    // rows in batches through emv_read_rows (device cells -> int64, INF = INT64_MAX)
    const int64_t batch = std::max<int64_t>(1, std::min<int64_t>(S2, (64ll << 20) / (8 * std::max<int64_t>(n, 1))));
    std::vector<int64_t> fb((size_t)(batch * n)), cb;
    std::vector<int32_t> c32;
    for (int64_t X = 0; ok && X < S2; X += batch) {
        const int64_t nr = std::min(batch, S2 - X);
        const int st = emv_read_rows(c, X, nr, fb.data(), nullptr);
        if (st != EMV_OK) {
            fclose(out);
            return st;
        }
        ok = fwrite(fb.data(), 8, (size_t)(nr * n), out) == (size_t)(nr * n);
    }
    if (ok && records) {
        cb.resize((size_t)(batch * n));
        c32.resize((size_t)(batch * n));
        for (int64_t X = 0; ok && X < S2; X += batch) {
            const int64_t nr = std::min(batch, S2 - X);
            const int st = emv_read_rows(c, X, nr, nullptr, cb.data());
            if (st != EMV_OK) {
                fclose(out);
                return st;
            }
            for (int64_t i = 0; i < nr * n; i++) c32[(size_t)i] = (int32_t)cb[(size_t)i];
            ok = fwrite(c32.data(), 4, (size_t)(nr * n), out) == (size_t)(nr * n);
        }
    }
    if (fclose(out) != 0) ok = false;
    if (!ok) {
        set_err(c, "short write to '%s'", path);
        return EMV_ERR_ARG;
    }
    return EMV_OK;
}

extern "C" int emv_read_rows(emv_ctx *c, int64_t X0, int64_t nrows, int64_t *f_out,
                             int64_t *code_out) {
    if (!c) return EMV_ERR_ARG;
    if (!c->d_f || c->tk == 0) {
        set_err(c, "no table kept (solve with EMV_KEEP_TABLE)");
        return EMV_ERR_STATE;
    }
    const int64_t S2 = (int64_t)1 << c->tk;
    if (X0 < 0 || nrows < 0 || X0 + nrows > S2) {
        set_err(c, "rows [%lld, %lld) outside [0, %lld)", (long long)X0, (long long)(X0 + nrows),
                (long long)S2);
        return EMV_ERR_ARG;
    }
    HIPCHK(c, hipSetDevice(c->device));
    const int64_t n = c->n, np = c->np;
    std::vector<char> buf((size_t)(np * std::max(c->tbytes, 4)));
    for (int64_t r = 0; r < nrows; r++) {
        const int64_t X = X0 + r;
        if (f_out) {
            if (c->flay.ts < 62) {
                // tiles: 2^ts cells of row X per tile, one tile block apart
                const int64_t tw = (int64_t)1 << c->flay.ts;
                const size_t w = (size_t)tw * c->tbytes;
                HIPCHK(c, hipMemcpy2D(buf.data(), w,
                                      (char *)c->d_f + (size_t)c->flay.cell(X, 0) * c->tbytes,
                                      (size_t)c->flay.tcells * c->tbytes, w,
                                      (size_t)((n + tw - 1) / tw), hipMemcpyDeviceToHost));
            } else {
                HIPCHK(c, hipMemcpy(buf.data(), (char *)c->d_f + (size_t)(X * np) * c->tbytes,
                                    (size_t)n * c->tbytes, hipMemcpyDeviceToHost));
            }
            for (int64_t v = 0; v < n; v++) {
                int64_t val;
                if (c->tbytes == 2) {
                    const uint16_t x = ((uint16_t *)buf.data())[v];
                    val = (x == 0xFFFF) ? INT64_MAX : (int64_t)x;
                } else if (c->tbytes == 4) {
                    int32_t x = ((int32_t *)buf.data())[v];
                    val = (x == INT32_MAX) ? INT64_MAX : (int64_t)x;
                } else {
                    val = ((int64_t *)buf.data())[v];
                }
                f_out[r * n + v] = val;
            }
        }
        if (code_out) {
            HIPCHK(c, hipMemcpy(buf.data(), c->d_b + X * np, (size_t)n * 4,
                                hipMemcpyDeviceToHost));
            for (int64_t v = 0; v < n; v++) code_out[r * n + v] = ((int32_t *)buf.data())[v];
        }
    }
    return EMV_OK;
}

extern "C" int emv_sssp(emv_ctx *c, int64_t s, int64_t *d_out, int64_t *pred_out, double *ms) {
    if (!c) return EMV_ERR_ARG;
    if (c->n <= 0) {
        set_err(c, "no graph loaded");
        return EMV_ERR_STATE;
    }
    if (s < 0 || s >= c->n) {
        set_err(c, "source %lld out of range", (long long)s);
        return EMV_ERR_ARG;
    }
    HIPCHK(c, hipSetDevice(c->device));
    EventList evl;
    hipEvent_t a, b;
    HIPCHK(c, hipEventCreate(&a));
    evl.v.push_back(a);
    HIPCHK(c, hipEventCreate(&b));
    evl.v.push_back(b);
    HIPCHK(c, hipEventRecord(a, c->stream));
    int64_t rounds = 0;
    std::vector<int32_t> src(1, (int32_t)s);
    CHK(sssp_chunk(c, src, &rounds));
    int64_t *dd = nullptr, *dp = nullptr;
    HIPCHK(c, hipMalloc((void **)&dd, sizeof(int64_t) * c->n));
    if (hipMalloc((void **)&dp, sizeof(int64_t) * c->n) != hipSuccess) {
        (void)hipFree(dd);
        set_err(c, "allocation failed");
        return EMV_ERR_NOMEM;
    }
    if (c->sssp_bits == 32)
        hipLaunchKernelGGL((k_unpack_sssp_v<uint32_t>), dim3((c->n + 255) / 256), dim3(256), 0,
                           c->stream, (const uint32_t *)c->d_D, (int)c->n, c->d_rp, c->d_adj, dd,
                           dp);
    else
        hipLaunchKernelGGL((k_unpack_sssp_v<uint64_t>), dim3((c->n + 255) / 256), dim3(256), 0,
                           c->stream, (const uint64_t *)c->d_D, (int)c->n, c->d_rp, c->d_adj, dd,
                           dp);
    hipError_t e0 = hipEventRecord(b, c->stream);
    hipError_t e1 = hipStreamSynchronize(c->stream);
    if (ms) *ms = ev_ms(a, b);
    hipError_t e2 = hipSuccess, e3 = hipSuccess;
    if (d_out) e2 = hipMemcpy(d_out, dd, sizeof(int64_t) * c->n, hipMemcpyDeviceToHost);
    if (pred_out) e3 = hipMemcpy(pred_out, dp, sizeof(int64_t) * c->n, hipMemcpyDeviceToHost);
    (void)hipFree(dd);
    (void)hipFree(dp);
    HIPCHK(c, e0);
    HIPCHK(c, e1);
    HIPCHK(c, e2);
    HIPCHK(c, e3);
    return EMV_OK;
}

namespace {

// Runs the certificate over a table on the device and fills *out.
int run_certificate(emv_ctx *c, const void *f, FLayout fl, int tb, const int32_t *b, int64_t np, int k,
                    const std::vector<int32_t> &kk, emv_certificate *out) {
    std::vector<std::vector<int32_t>> srcs((size_t)k);
    for (int t = 0; t < k; t++) srcs[(size_t)t].push_back(kk[(size_t)t]);
    hipEvent_t a, e;
    EventList evl;
    HIPCHK(c, hipEventCreate(&a));
    evl.v.push_back(a);
    HIPCHK(c, hipEventCreate(&e));
    evl.v.push_back(e);
    HIPCHK(c, hipEventRecord(a, c->stream));
    unsigned long long cnt[4];
    HIPCHK(c, certify_table(c->stream, f, fl, tb, b, np, (int)c->n, k, c->d_rp, c->d_adj, srcs, kk, 0,
                            cnt));
    HIPCHK(c, hipEventRecord(e, c->stream));
    HIPCHK(c, hipEventSynchronize(e));
    out->cells = ((int64_t)1 << k) * c->n;
    out->bad_value = (int64_t)cnt[0];
    out->not_finite = (int64_t)cnt[1];
    out->bad_record = (int64_t)cnt[2];
    if (cnt[3] != ~0ull) {
        out->first_X = (int64_t)(cnt[3] >> 32);
        out->first_v = (int64_t)(cnt[3] & 0xFFFFFFFFull);
    }
    out->ms = ev_ms(a, e);
    return EMV_OK;
}

}  // namespace

extern "C" int emv_certify(emv_ctx *c, emv_certificate *out) {
    if (!c || !out) return EMV_ERR_ARG;
    memset(out, 0, sizeof(*out));
    out->first_X = out->first_v = -1;
    if (!c->d_f || c->tk == 0 || (int64_t)c->kept_kk.size() != c->tk) {
        set_err(c, "no table kept (solve with EMV_KEEP_TABLE)");
        return EMV_ERR_STATE;
    }
    HIPCHK(c, hipSetDevice(c->device));
    return run_certificate(c, c->d_f, c->flay, c->tbytes, c->d_b, c->np, (int)c->tk, c->kept_kk,
                           out);
}

extern "C" int emv_certify_table(emv_ctx *c, int64_t k, const int64_t *kk, const int64_t *f,
                                 const int32_t *code, emv_certificate *out) {
    if (!c || !out || !kk || !f || !code) return EMV_ERR_ARG;
    memset(out, 0, sizeof(*out));
    out->first_X = out->first_v = -1;
    if (c->n <= 0 || !c->d_rp) {
        set_err(c, "no graph loaded");
        return EMV_ERR_STATE;
    }
    if (k < 1 || k > 30) {
        set_err(c, "k=%lld outside [1,30]", (long long)k);
        return EMV_ERR_ARG;
    }
    std::vector<int32_t> kk32((size_t)k);
    for (int64_t t = 0; t < k; t++) {
        if (kk[t] < 0 || kk[t] >= c->n) {
            set_err(c, "terminal %lld out of range", (long long)kk[t]);
            return EMV_ERR_ARG;
        }
        kk32[(size_t)t] = (int32_t)kk[t];
    }
    HIPCHK(c, hipSetDevice(c->device));
    const size_t cells = ((size_t)1 << k) * (size_t)c->n;
    int64_t *d_f = nullptr;
    int32_t *d_b = nullptr;
    HIPCHK(c, hipMalloc((void **)&d_f, sizeof(int64_t) * cells));
    int rc = EMV_OK;
    hipError_t err = hipMalloc((void **)&d_b, sizeof(int32_t) * cells);
    if (err == hipSuccess) err = hipMemcpy(d_f, f, sizeof(int64_t) * cells, hipMemcpyHostToDevice);
    if (err == hipSuccess) err = hipMemcpy(d_b, code, sizeof(int32_t) * cells, hipMemcpyHostToDevice);
    if (err != hipSuccess) {
        set_err(c, "table upload failed: %s", hipGetErrorString(err));
        rc = (err == hipErrorOutOfMemory) ? EMV_ERR_NOMEM : EMV_ERR_HIP;
    } else {
        rc = run_certificate(c, d_f, rows_layout(c->n), 8, d_b, c->n, (int)k, kk32, out);
    }
    (void)hipFree(d_f);
    if (d_b) (void)hipFree(d_b);
    return rc;
}

extern "C" int emv_sssp_noalloc(emv_ctx *c, int64_t s, int64_t *d_dist, int64_t *d_pred) {
    if (!c) return EMV_ERR_ARG;
    if (c->n <= 0 || !c->d_rp) {
        set_err(c, "no graph loaded");
        return EMV_ERR_STATE;
    }
    if (s < 0 || s >= c->n) {
        set_err(c, "source %lld out of range", (long long)s);
        return EMV_ERR_ARG;
    }
    if (c->n > kScratchMaxN) {
        set_err(c, "n = %lld exceeds the %lld vertices the %zu-byte scratch holds (use emv_sssp)",
                (long long)c->n, (long long)kScratchMaxN, kScratchBytes);
        return EMV_ERR_CAPACITY;
    }
    if (!c->d_scratch) {
        set_err(c, "scratch address unknown");
        return EMV_ERR_STATE;
    }
    HIPCHK(c, hipSetDevice(c->device));
    const int64_t n = c->n;
    uint64_t *lab = reinterpret_cast<uint64_t *>(c->d_scratch);
    uint32_t *flg = reinterpret_cast<uint32_t *>(c->d_scratch + 8 * n);
    uint32_t *words = reinterpret_cast<uint32_t *>(c->d_scratch + 16 * n);
    const int grid = (int)std::max<int64_t>(
        1, std::min<int64_t>((n + 255) / 256, (int64_t)c->num_cu * 8));
    hipLaunchKernelGGL(k_ss_init, dim3(grid), dim3(256), 0, c->stream, lab, flg, n, s, words);
    hipLaunchKernelGGL(k_ss_round, dim3(grid), dim3(256), 0, c->stream, lab, flg, flg + n, c->d_rp,
                       c->d_adj, n, 0, 1, words);
    HIPCHK(c, hipGetLastError());
    int r = 1, group = 2;
    for (;;) {
        for (int g = 0; g < group; g++, r++)
            hipLaunchKernelGGL(k_ss_round, dim3(grid), dim3(256), 0, c->stream, lab,
                               flg + (r & 1) * n, flg + ((r + 1) & 1) * n, c->d_rp, c->d_adj, n,
                               r, 0, words);
        HIPCHK(c, hipGetLastError());
        HIPCHK(c, hipMemcpyAsync(c->h_pin, words + (r & 3), sizeof(uint32_t),
                                 hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        if (*c->h_pin == 0) break;
        group = std::min(group * 2, 64);
        if ((int64_t)r > 2 * (n + 8)) {
            set_err(c, "shortest paths did not converge");
            return EMV_ERR_RANGE;
        }
    }
    hipLaunchKernelGGL(k_ss_out, dim3(grid), dim3(256), 0, c->stream, lab, c->d_rp, c->d_adj, n,
                       d_dist, d_pred);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return EMV_OK;
}

extern "C" void emv_result_free(emv_result *res) {
    if (!res) return;
    free(res->pairs);
    res->pairs = nullptr;
    res->npairs = 0;
}
