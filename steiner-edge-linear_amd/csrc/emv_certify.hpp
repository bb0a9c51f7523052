// emv_certify.hpp -- an independent check of a finished EMV table.
//
// The table of a connected graph with positive weights is the unique
// solution of the EMV fixed point (reader-el.c:1988-2130):
//   f[{t}][v] = min(0 if v is a source of row {t}, min_u f[{t}][u] + w(u,v))
//   f[X][v]   = min(min over splits f[Y][v] + f[X\Y][v], min_u f[X][u] + w(u,v))
// (rows by increasing |X|: given correct lower rows, the row's values are the
// shortest distances from a virtual source carrying the merged values, and
// positive weights make that solution unique).  Checking that every cell
// equals the least of its candidates therefore pins every cell, whatever
// order and whatever kernels produced the table.  The records are checked
// against their rule at the same time:
//   * row 0: INF and record -1 (reader-el.c:2244-2259);
//   * a singleton row {t}: record {t} at its sources; elsewhere the
//     shortcut 2^k + kk[t] (reader-el.c:2015-2027), or, for the rows of a
//     group query (sources = the group's members), the edge 2^k + u from the
//     smallest neighbour u with f[u] + w == f[v];
//   * any other row: the merge split Y when the cell equals the least split
//     sum, Y the smallest subset attaining it among those without X's top
//     element (the first strict minimum of the reference's increasing-Y walk,
//     reader-el.c:2070-2085); else the edge from the smallest neighbour u
//     with f[u] + w == f[v] (reader-el.c:2117-2126 with the equal-cost tie
//     broken by neighbour id).
//
// The kernel is deliberately plain: one thread per cell, all arithmetic in
// int64, every candidate recomputed from the table itself.  It shares no code
// with the solver's merge, relaxation or unpack kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <vector>

#include "emv_kernels.hpp"

namespace emv {

constexpr int64_t kCertInf = INT64_MAX;

// Cell i of a table of `tb`-byte cells (2: 0xFFFF = INF, 4: INT32_MAX = INF,
// 8: INT64_MAX = INF) as int64 with INF = INT64_MAX.
__device__ __forceinline__ int64_t cert_cell(const void *__restrict__ f, int tb, int64_t i) {
    if (tb == 2) {
        const uint16_t x = static_cast<const uint16_t *>(f)[i];
        return x == 0xFFFFu ? kCertInf : (int64_t)x;
    }
    if (tb == 4) {
        const int32_t x = static_cast<const int32_t *>(f)[i];
        return x == INT32_MAX ? kCertInf : (int64_t)x;
    }
    return static_cast<const int64_t *>(f)[i];
}

__device__ __forceinline__ int64_t cert_add(int64_t a, int64_t b) {
    return (a == kCertInf || b == kCertInf) ? kCertInf : a + b;
}

// out[0] cells whose value is not the least candidate, out[1] cells of rows
// X != 0 holding INF, out[2] records other than the rule's, out[3] the
// smallest failing (X << 32 | v) (starts at all ones).
// mark[v] bit t: v is a source of singleton row {t}.  kk[t]: the terminal of
// row {t} (shortcut records; unused with edge_singletons).
// Rows X = row0 .. row0 + nrows - 1; one workgroup per (row, 256 vertices).
[[maybe_unused]] static __global__ __launch_bounds__(256) void k_certify(
    const void *__restrict__ f, FLayout fl, int tb, const int32_t *__restrict__ b, int64_t np,
    int n, int k,
    const int32_t *__restrict__ rp, const Edge *__restrict__ adj, const uint32_t *__restrict__ mark,
    const int32_t *__restrict__ kk, int edge_singletons, int64_t row0, int64_t nrows,
    unsigned long long *__restrict__ out) {
    const int64_t vblocks = ((int64_t)n + 255) / 256;
    const int64_t items = nrows * vblocks;
    const int64_t S2 = 1ll << k;
    for (int64_t item = blockIdx.x; item < items; item += gridDim.x) {
        const int64_t X = row0 + item / vblocks;
        const int64_t v = (item % vblocks) * 256 + threadIdx.x;
        if (v >= n) continue;
        const int64_t cell = X * np + v;  // records: rows of np; cells: layout fl
        const int64_t fv = cert_cell(f, tb, fl.cell(X, v));
        const int32_t code = b[cell];
        bool bad_value = false, bad_record = false, not_finite = false;
        if (X == 0) {
            bad_value = fv != kCertInf;
            bad_record = code != -1;
        } else {
            not_finite = fv == kCertInf;
            // edge candidates: the least f[X][u] + w and the smallest u attaining it
            int64_t best_e = kCertInf;
            int64_t bu = -1;
            for (int32_t e = rp[v]; e < rp[v + 1]; e++) {
                const int64_t u = adj[e].v;
                const int64_t c = cert_add(cert_cell(f, tb, fl.cell(X, u)), (int64_t)adj[e].w);
                if (c < best_e || (c == best_e && c != kCertInf && u < bu)) {
                    best_e = c;
                    bu = u;
                }
            }
            const bool single = (X & (X - 1)) == 0;
            if (single) {
                const int t = __builtin_ctzll((unsigned long long)X);
                const bool src = (mark[v] >> t) & 1u;
                const int64_t expect = src ? 0 : best_e;
                bad_value = fv != expect;
                int64_t want;
                if (src)
                    want = X;
                else if (edge_singletons)
                    want = S2 + bu;
                else
                    want = S2 + kk[t];
                bad_record = (int64_t)code != want;
            } else {
                // split candidates: Y over the nonempty subsets of X without
                // its top element, in increasing order (first strict minimum)
                const int64_t top = 1ll << (63 - __builtin_clzll((unsigned long long)X));
                const int64_t Xl = X ^ top;
                int64_t best_s = kCertInf, bY = -1;
                int64_t Y = 0;
                do {
                    Y = (Y - Xl) & Xl;
                    const int64_t s = cert_add(cert_cell(f, tb, fl.cell(Y, v)),
                                               cert_cell(f, tb, fl.cell(X ^ Y, v)));
                    if (s < best_s) {
                        best_s = s;
                        bY = Y;
                    }
                } while (Y != Xl);
                const int64_t expect = best_s < best_e ? best_s : best_e;
                bad_value = fv != expect;
                const int64_t want = (fv == best_s) ? bY : S2 + bu;
                bad_record = (int64_t)code != want;
            }
        }
        if (bad_value) atomicAdd(&out[0], 1ull);
        if (not_finite) atomicAdd(&out[1], 1ull);
        if (bad_record) atomicAdd(&out[2], 1ull);
        if (bad_value || not_finite || bad_record)
            atomicMin(&out[3], ((unsigned long long)X << 32) | (unsigned long long)v);
    }
}

// Host side: runs k_certify over every row of one table on `stream` and
// fills out[0..3] as above (out[3] = ~0 when no cell fails).  srcs[t] lists
// the sources of singleton row {t}; kk[t] its terminal.  Returns the HIP
// error of the first failing call.
static inline hipError_t certify_table(hipStream_t stream, const void *f, FLayout fl, int tb,
                                       const int32_t *b, int64_t np, int n, int k,
                                       const int32_t *rp, const Edge *adj,
                                       const std::vector<std::vector<int32_t>> &srcs,
                                       const std::vector<int32_t> &kk, int edge_singletons,
                                       unsigned long long out[4]) {
    std::vector<uint32_t> mark((size_t)n, 0u);
    for (int t = 0; t < k && t < (int)srcs.size(); t++)
        for (int32_t v : srcs[(size_t)t]) mark[(size_t)v] |= 1u << t;
    std::vector<int32_t> kk32(kk);
    kk32.resize((size_t)k, -1);
    uint32_t *d_mark = nullptr;
    int32_t *d_kk = nullptr;
    unsigned long long *d_out = nullptr;
    const unsigned long long init[4] = {0ull, 0ull, 0ull, ~0ull};
    hipError_t err = hipMalloc((void **)&d_mark, sizeof(uint32_t) * (size_t)n);
    if (err == hipSuccess) err = hipMalloc((void **)&d_kk, sizeof(int32_t) * (size_t)k);
    if (err == hipSuccess) err = hipMalloc((void **)&d_out, sizeof(init));
    if (err == hipSuccess)
        err = hipMemcpyAsync(d_mark, mark.data(), sizeof(uint32_t) * (size_t)n,
                             hipMemcpyHostToDevice, stream);
    if (err == hipSuccess)
        err = hipMemcpyAsync(d_kk, kk32.data(), sizeof(int32_t) * (size_t)k,
                             hipMemcpyHostToDevice, stream);
    if (err == hipSuccess)
        err = hipMemcpyAsync(d_out, init, sizeof(init), hipMemcpyHostToDevice, stream);
    // rows in slices, so that no single launch runs for long
    const int64_t S2 = 1ll << k, vblocks = ((int64_t)n + 255) / 256;
    const int64_t rows = std::max<int64_t>(1, (1ll << 22) / vblocks);
    for (int64_t X0 = 0; err == hipSuccess && X0 < S2; X0 += rows) {
        const int64_t nr = std::min<int64_t>(rows, S2 - X0);
        const int64_t items = nr * vblocks;
        hipLaunchKernelGGL(k_certify, dim3((unsigned)std::min<int64_t>(items, 1 << 20)), dim3(256),
                           0, stream, f, fl, tb, b, np, n, k, rp, adj, d_mark, d_kk, edge_singletons,
                           X0, nr, d_out);
        err = hipGetLastError();
    }
    if (err == hipSuccess)
        err = hipMemcpyAsync(out, d_out, sizeof(init), hipMemcpyDeviceToHost, stream);
    if (err == hipSuccess) err = hipStreamSynchronize(stream);
    if (d_mark) (void)hipFree(d_mark);
    if (d_kk) (void)hipFree(d_kk);
    if (d_out) (void)hipFree(d_out);
    return err;
}

}  // namespace emv
