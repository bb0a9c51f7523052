// emv_blocks.hpp -- host side of block-local relaxation: cutting a graph
// into blocks of at most kBV vertices and the rule that picks the block path.
//
// rp / adj are one graph's CSR with vertex ids local to the graph (rp may
// start at an offset into a shared adjacency: entries rp[u] .. rp[u+1] of
// adj belong to vertex u).
#pragma once
#include <stdint.h>
#include <stdlib.h>

#include <algorithm>
#include <functional>
#include <utility>
#include <vector>

#include "emv_kernels.hpp"

namespace emv {

// Blocks of at most kBV vertices grown breadth-first over unassigned
// vertices from the lowest unassigned id.  Within a block the border slots
// (vertices with a neighbour outside the block) come first, then the inner
// ones, each in BFS order; bnb[b] counts the border slots of block b.
// Returns the share of adjacency entries that stay inside their block.
struct BlockPart {
    int nblk = 0;
    std::vector<int32_t> boff, bvert, brp, bext, vblk, bnb;
    std::vector<Edge> badj;
};

inline double build_blocks(int64_t n, const int32_t *rp, const Edge *adj, BlockPart &P) {
    P.nblk = 0;
    P.vblk.assign((size_t)n, -1);
    std::vector<int32_t> order, slot((size_t)n, -1);  // order: members block by block, BFS order
    order.reserve((size_t)n);
    P.boff.assign(1, 0);
    for (int64_t seed = 0; seed < n; seed++) {
        if (P.vblk[seed] >= 0) continue;
        const int b = P.nblk++;
        const size_t q0 = order.size();
        order.push_back((int32_t)seed);
        P.vblk[seed] = b;
        for (size_t head = q0; head < order.size() && order.size() - q0 < (size_t)kBV; head++) {
            const int32_t u = order[head];
            for (int32_t e = rp[u]; e < rp[u + 1] && order.size() - q0 < (size_t)kBV; e++) {
                const int32_t v = adj[e].v;
                if (P.vblk[v] < 0) {
                    P.vblk[v] = b;
                    order.push_back(v);
                }
            }
        }
        P.boff.push_back((int32_t)order.size());
    }
    // slots: the border vertices of each block first, then the inner ones
    P.bvert.assign((size_t)n, -1);
    P.bnb.assign((size_t)P.nblk, 0);
    for (int b = 0; b < P.nblk; b++) {
        int32_t s = P.boff[b];
        for (int pass = 0; pass < 2; pass++)
            for (int32_t i = P.boff[b]; i < P.boff[b + 1]; i++) {
                const int32_t u = order[i];
                bool border = false;
                for (int32_t e = rp[u]; e < rp[u + 1] && !border; e++)
                    border = P.vblk[adj[e].v] != b;
                if (border != (pass == 0)) continue;
                slot[u] = s;
                P.bvert[s++] = u;
                if (pass == 0) P.bnb[b]++;
            }
    }
    P.brp.assign((size_t)n + 1, 0);
    P.bext.assign((size_t)n, 0);
    P.badj.clear();
    P.badj.reserve((size_t)(rp[n] - rp[0]));
    int64_t inside = 0;
    for (int64_t s = 0; s < n; s++) {
        const int32_t u = P.bvert[s];
        const int32_t b = P.vblk[u];
        const int32_t s0 = P.boff[b];
        P.brp[s] = (int32_t)P.badj.size();
        for (int32_t e = rp[u]; e < rp[u + 1]; e++)
            if (P.vblk[adj[e].v] == b) P.badj.push_back(Edge{slot[adj[e].v] - s0, adj[e].w});
        P.bext[s] = (int32_t)P.badj.size();
        inside += P.bext[s] - P.brp[s];
        for (int32_t e = rp[u]; e < rp[u + 1]; e++)
            if (P.vblk[adj[e].v] != b) P.badj.push_back(adj[e]);
    }
    P.brp[n] = (int32_t)P.badj.size();
    return rp[n] == rp[0] ? 0.0 : (double)inside / (double)(rp[n] - rp[0]);
}

// Distances inside each block (over the block's own edges only) from every
// border slot to every slot of the block: block b's matrix starts at
// bdoff[b], row i (border slot s0 + i) has kBRow entries, slot s0 + j at
// kBQ * (j % kBW) + j / kBW -- the bands of the block workgroup's waves
// (wave w holds slots w, w + kBW, ...), each padded to kBQ entries (padding
// 0).  Blocks
// are connected (grown breadth-first), so every entry is finite.  dmax is
// the largest entry.
struct BlockDist {
    std::vector<int64_t> bdoff;  // nblk + 1 entries
    std::vector<uint64_t> d;
    uint64_t dmax = 0;
};

inline void block_distances(const BlockPart &P, BlockDist &B) {
    B.bdoff.assign(1, 0);
    for (int b = 0; b < P.nblk; b++)
        B.bdoff.push_back(B.bdoff.back() + (int64_t)P.bnb[b] * kBRow);
    B.d.assign((size_t)B.bdoff.back(), 0);
    B.dmax = 0;
    std::vector<std::pair<uint64_t, int32_t>> heap;
    std::vector<uint64_t> row;
    for (int b = 0; b < P.nblk; b++) {
        const int32_t s0 = P.boff[b], cnt = P.boff[b + 1] - s0;
        for (int i = 0; i < P.bnb[b]; i++) {
            row.assign((size_t)cnt, ~0ull);
            heap.clear();
            row[i] = 0;
            heap.push_back({0ull, i});
            while (!heap.empty()) {  // Dijkstra, min-heap on (distance, slot)
                std::pop_heap(heap.begin(), heap.end(), std::greater<>());
                const auto [du, j] = heap.back();
                heap.pop_back();
                if (du != row[j]) continue;
                for (int32_t e = P.brp[s0 + j]; e < P.bext[s0 + j]; e++) {
                    const int32_t t = P.badj[e].v;
                    const uint64_t nd = du + (uint64_t)P.badj[e].w;
                    if (nd < row[t]) {
                        row[t] = nd;
                        heap.push_back({nd, t});
                        std::push_heap(heap.begin(), heap.end(), std::greater<>());
                    }
                }
            }
            uint64_t *out = B.d.data() + B.bdoff[b] + (int64_t)i * kBRow;
            for (int32_t j = 0; j < cnt; j++) {
                out[kBQ * (j % kBW) + j / kBW] = row[j];
                B.dmax = std::max(B.dmax, row[j]);
            }
        }
    }
}

// Storage width of the in-block distances: 16 or 32 bits, 0 when an entry
// needs more (the visits then sweep).
inline int dist_width(uint64_t dmax) {
    return dmax <= 0xFFFFull ? 16 : dmax <= 0xFFFFFFFFull ? 32 : 0;
}

template <typename T>
std::vector<T> narrow(const std::vector<uint64_t> &v) {
    return std::vector<T>(v.begin(), v.end());
}

// EMV_BLOCK_CLOSURE=0 keeps the sweeps on every visit (no distance matrices).
inline int closure_mode() {
    const char *e = getenv("EMV_BLOCK_CLOSURE");
    return e ? atoi(e) : 1;
}

// EMV_BLOCKS=0 turns block relaxation off, EMV_BLOCKS=1 forces it; unset
// (-1) leaves it to the graph.
inline int blocks_mode() {
    const char *eb = getenv("EMV_BLOCKS");
    return eb ? atoi(eb) : -1;
}

// Block relaxation pays when most adjacency entries stay inside their block
// (grids and meshes keep about 0.9 of them; a random 3-regular graph about
// 0.66, which its spanning trees alone explain) and the graph spans a few
// blocks at least.
inline bool blocks_wanted(int mode, int64_t n, double inside) {
    return mode == 1 || (mode != 0 && inside >= 0.8 && n >= 4 * kBV);
}

}  // namespace emv
