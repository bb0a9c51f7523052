/*
 * emv_hip.h -- C ABI of the MI355X (gfx950) Erickson-Monma-Veinott Steiner
 * tree dynamic program.  Plain C types only; no HIP or torch types cross it.
 *
 * Drop-in boundary.  The reference computes the DP inside
 *     index_t erickson_monma_veinott(steinerq_t *root, index_t list_soln)
 * (reader/reader-el.c:2141-2343), which allocates the tables (2212-2259),
 * runs emv_kernel (1964-2139, with dijkstra 1762-1880) and the traceback
 * build_tree/backtrack (1907-1944).  That one call is replaced by
 *     emv_open + emv_load_graph(root->n, root->m, root->pos, root->adj)
 *     + emv_solve(root->k, root->kk, EMV_LIST, &res)
 * and the host prints the reference's lines from the result.
 *
 * Conventions.  Every function returns EMV_OK (0) or a negative EMV_ERR_*
 * status, never aborts and never throws; emv_last_error() gives the text.
 * Vertex ids are 0-based.  The caller owns every input array and may free it
 * when the call returns.  The library owns device memory and the pairs array
 * of an emv_result until emv_result_free().  A context is driven by one host
 * thread at a time.  Weights must be integers in [1, 2^31-1]; the graph must
 * be connected (EMV_ERR_DISCONNECTED otherwise).
 */
#ifndef EMV_HIP_H
#define EMV_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define EMV_OK                 0
#define EMV_ERR_ARG           -1   /* bad argument */
#define EMV_ERR_HIP           -2   /* HIP runtime error */
#define EMV_ERR_NOMEM         -3   /* device or host allocation failed */
#define EMV_ERR_STATE         -4   /* call out of order (no graph, no table) */
#define EMV_ERR_DISCONNECTED  -5   /* a terminal does not reach every vertex */
#define EMV_ERR_RANGE         -6   /* values exceed the supported integer range */
#define EMV_ERR_NODEV         -7   /* no HIP device */
#define EMV_ERR_CAPACITY      -8   /* request exceeds a configured capacity */
#define EMV_ERR_ADMISSION     -9   /* serving: request can never fit in device memory */
#define EMV_ERR_CANCELLED     -10  /* serving: request cancelled before it finished */

/* emv_solve flags */
#define EMV_LIST        0x1u  /* run the device traceback; fill res->pairs */
#define EMV_FORCE_I64   0x2u  /* int64 table even when int32 bounds hold */
#define EMV_KEEP_TABLE  0x4u  /* keep the table on the device for emv_read_rows */
#define EMV_FORCE_P64   0x8u  /* 64-bit merge packing and relaxation labels */
#define EMV_WIDE_LABELS 0x10u /* no 16-bit relaxation labels even when they fit */
#define EMV_TABLE32     0x20u /* int32 table even when the bound allows 16-bit cells */
#define EMV_MERGE_1V    0x40u /* merges of 16-bit cells one vertex per lane, as wider cells */
#define EMV_MERGE_ROWS  0x80u /* merges of 16-bit cells row by row, not in stars of rows sharing a base set */
#define EMV_COUNT_VISITS 0x200u /* count the (chunk, vertex) items the vertex-wise layer rounds visit */
#define EMV_MERGE_CUBE  0x100u /* layers 3..8 of 16-bit cells merged from rows staged in LDS (cubes);
                                  measured slower than the stars on MI355X, so not the default */

typedef struct emv_ctx emv_ctx;

typedef struct {
    int64_t cost;            /* f[C][q], C = 2^(k-1)-1, q = kk[k-1] (reader-el.c:2137) */
    int32_t k;
    int32_t table_bytes;     /* 2 (16-bit cells, when the distance bound is below 0xFFFF),
                                4 (int32 table) or 8 (int64 table) */
    int32_t merge_bits;      /* 32 or 64: width of the packed merge accumulator */
    int32_t label_bits;      /* width of the relaxation labels: 16, 32 or 64 */
    int64_t n;
    int64_t dp_cells;        /* 2^k * n */
    /* timings, milliseconds, from HIP events on the solve stream */
    double zero_ms;          /* table allocation + initial fill */
    double singleton_ms;     /* k single-source shortest paths */
    double merge_ms;         /* sum over layers of the subset-merge kernels */
    double relax_ms;         /* sum over layers of the relaxation rounds */
    double kernel_ms;        /* singleton + merge + relax + unpack, as the reference's [kernel:] */
    double traceback_ms;
    double total_ms;
    /* algorithmic traffic models (bytes) */
    double merge_alg_bytes;  /* cb*n*(3^k-2^(k+1)+1) + (cb+ab)*n*(2^k-k-1) */
    double relax_alg_bytes;  /* (2^k-1)*(cb*n + cb*2m + (cb+ab)*n) + k*(4(n+1)+16m) */
    int64_t merge_launches;
    int64_t relax_rounds;    /* relaxation rounds launched, layers 2..k */
    int64_t device_bytes;    /* device memory held during the solve */
    int64_t npairs;          /* traceback: number of (v,u) pairs */
    int64_t *pairs;          /* 2*npairs, 0-based, reference DFS order (reader-el.c:1907-1944) */
    int64_t relax_blocks;    /* vertex blocks of the block-local relaxation; 0 = vertex-wise rounds */
    double unpack_ms;        /* sum over layers of the record decisions + table writes */
    int64_t singleton_rounds;/* relaxation rounds of the singleton shortest paths */
    int64_t block_dist_bits; /* 16 / 32: block visits after the first apply in-block distances
                                from the lowered border slots; 0: they sweep */
    double merge_split_ms;   /* part of merge_ms in the layers merged in parts (rows of >= 128 pairs) */
    int64_t relative_label_layers; /* layers relaxed on 32-bit labels relative to a per-row base:
                                      64-bit labels were needed by the bound, but every row's merged
                                      values span less than 2^32 - 1 - max w */
    int64_t compact_merge_layers;  /* int64 tables: layer batches merged from the int32 copy of
                                      the rows (cell - row base) instead of the int64 cells */
    int64_t relax_item_visits;     /* with EMV_COUNT_VISITS: (chunk, vertex) items that the
                                      vertex-wise relaxation rounds of layers 2..k processed (each
                                      visit reads relax_chunk_rows rows of its own and its
                                      neighbours' labels); -1 otherwise */
    int64_t relax_chunk_rows;      /* rows of a relaxation item (chunk, vertex) in layers 2..k:
                                      128 when 16-bit labels ride two per 32-bit word, else 64 */
} emv_result;

/* Context on HIP device `device`. */
int emv_open(emv_ctx **ctx, int device);
void emv_close(emv_ctx *ctx);
const char *emv_last_error(const emv_ctx *ctx);

/* Graph in the reference's steinerq_t layout (reader-el.c:696-704, built by
 * root_build 706-899): for u < n, adj[pos[u]] = degree, followed by degree
 * (neighbour, weight) pairs.  Rows u >= n (virtual sources) are ignored.
 * Replaces the pos/adj hand-over at reader-el.c:2269. */
int emv_load_graph(emv_ctx *ctx, int64_t n, int64_t m, const int64_t *pos,
                   const int64_t *adj);

/* Graph as the reference's edge list graph_t.edges (reader-el.c:358-373,
 * graph_add_edge 421-438): 3*m int64 (u, v, w), 0-based. */
int emv_load_edges(emv_ctx *ctx, int64_t n, int64_t m, const int64_t *edges);

/* Full EMV DP for terminals kk[0..k-1] on the loaded graph: replaces
 * emv_kernel + build_tree (reader-el.c:2212-2310).  1 <= k <= 30. */
int emv_solve(emv_ctx *ctx, int64_t k, const int64_t *kk, uint32_t flags,
              emv_result *res);

/* After a solve with EMV_KEEP_TABLE: rows X0 .. X0+nrows-1 of the table as
 * int64 (INF = INT64_MAX, reader-el.c:69) and of the record codes (-1 unset,
 * Y for a merge record, 2^k + u for an edge record).  Either output may be
 * NULL.  Each output holds nrows*n values. */
int emv_read_rows(emv_ctx *ctx, int64_t X0, int64_t nrows, int64_t *f_out,
                  int64_t *code_out);

/* After a solve with EMV_KEEP_TABLE: the table as a canonical file (the
 * reference keeps its table in memory only; this is the dump format of the
 * path's table, for comparing solves byte for byte).  Little-endian:
 *   64-byte header: char magic[8] = "EMVTABLE"; uint32 version = 1;
 *   uint32 cell_bytes = 8; int64 k; int64 n; int64 inf = INT64_MAX
 *   (reader-el.c:69); uint32 records (0/1); uint32 reserved = 0;
 *   int64 cost (the answer cell, reader-el.c:2137); 8 bytes of zeros;
 *   then f[X][v] as int64 for X = 0 .. 2^k-1, v = 0 .. n-1 (row 0 is INF);
 *   then, with EMV_DUMP_RECORDS, the record codes of every cell as int32 in
 *   the same order (as emv_read_rows).
 * EMV_ERR_STATE without a kept table; EMV_ERR_ARG if the file cannot be
 * written. */
#define EMV_DUMP_RECORDS 0x1u
int emv_write_table(emv_ctx *ctx, const char *path, uint32_t flags);

/* Releases the table kept by EMV_KEEP_TABLE. */
int emv_release_table(emv_ctx *ctx);

/* Single-source shortest paths on the loaded graph (reader-el.c:1762-1880;
 * the -dijkstra command 2452-2478 and the k == 2 branch 2159-2209).
 * d_out[v] = distance (INT64_MAX if unreachable); pred_out[v] = smallest
 * neighbour u with d[u] + w(u,v) == d[v], -1 for s.  Either may be NULL. */
int emv_sssp(emv_ctx *ctx, int64_t s, int64_t *d_out, int64_t *pred_out,
             double *ms);

/* emv_sssp without any allocation.  No counterpart in the reference; it is
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid it).  Same results as
 * emv_sssp, but d_dist and d_pred are DEVICE pointers the caller owns (n
 * int64 each; either may be NULL), and the work memory is a fixed array of
 * 32 MiB in the library, one per GPU, held until the process ends: graphs of
 * more than 2,097,136 vertices are refused with EMV_ERR_CAPACITY.  The call
 * makes no hipMalloc/hipFree; it uses the context's stream and pinned word
 * and returns after the stream has drained.  Every call on one GPU shares
 * that GPU's array: calls that overlap on one GPU (two host threads, two
 * contexts) race and must be serialised by the caller. */
int emv_sssp_noalloc(emv_ctx *ctx, int64_t s, int64_t *d_dist, int64_t *d_pred);

void emv_result_free(emv_result *res);

/* Certificate of a table: counts of the cells that break the EMV fixed point
 * or the record rule.  A connected graph with positive weights has exactly
 * one table that satisfies, for every row X != 0 and vertex v,
 *   f[X][v] = least of: 0 (X = {t}, v = kk[t]); f[X][u] + w(u,v) over the
 *             edges; f[Y][v] + f[X\Y][v] over the splits of X,
 * so a table with no failing cell equals the reference's (reader-el.c:1964-
 * 2139) in every cell.  Records: row 0 holds -1; a singleton row holds {t}
 * at kk[t] and 2^k + kk[t] elsewhere (reader-el.c:2015-2027); other rows
 * the smallest least split Y (taken without X's top element) when the cell
 * equals the least split sum (reader-el.c:2070-2085), else 2^k + the
 * smallest neighbour u with f[u] + w == f[v] (2117-2126).  The reference has
 * no such check; its only self-check compares the cost with the STP file's
 * cost line (reader-el.c:2489-2491).  The check runs on the device, one
 * thread per cell, independently of the solver's kernels. */
typedef struct {
    int64_t cells;       /* cells checked: 2^k * n, row 0 included */
    int64_t bad_value;   /* cells that are not the least of their candidates */
    int64_t not_finite;  /* cells of rows X != 0 holding INF */
    int64_t bad_record;  /* record codes other than the rule's */
    int64_t first_X;     /* smallest failing cell (X, then v); -1 when none fails */
    int64_t first_v;
    double ms;           /* device time of the check */
} emv_certificate;

/* Certificate of the table kept by the last emv_solve (EMV_KEEP_TABLE).
 * EMV_ERR_STATE without a kept table. */
int emv_certify(emv_ctx *ctx, emv_certificate *out);

/* Certificate of a table held by the caller, for terminals kk[0..k-1] on the
 * loaded graph: f as int64 (INF = INT64_MAX), code as int32, each 2^k x n
 * row-major -- a table file (emv_write_table), the output of another solver,
 * or a table with a planted change.  The device copy is freed on return. */
int emv_certify_table(emv_ctx *ctx, int64_t k, const int64_t *kk, const int64_t *f,
                      const int32_t *code, emv_certificate *out);

/* Device memory the solve of a k-terminal query on the loaded graph needs,
 * in bytes (tables + relaxation staging). */
int64_t emv_solve_footprint(const emv_ctx *ctx, int64_t k, uint32_t flags);

#ifdef __cplusplus
}
#endif

#endif /* EMV_HIP_H */
