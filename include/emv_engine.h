/*
 * emv_engine.h -- C ABI of the multi-instance EMV engine: batched solves of
 * many independent instances and a continuous-batching query server on a
 * resident graph.  Plain C types only.
 *
 * The reference solves one instance per process (reader-el.c:2355-2536,
 * experiments/report.py:693 spawns one reader-el per instance) and has no
 * batching or serving interface.  These entry points have no counterpart
 * there: they run many reference-equivalent solves (every DP cell and cost
 * identical to erickson_monma_veinott, reader-el.c:2141) in shared launches.
 *
 * Conventions as in emv_hip.h: EMV_OK or a negative EMV_ERR_* status,
 * emv_engine_last_error() for the text, 0-based vertex ids, inputs owned by
 * the caller.  k must be in [2, 30].  Engine instances keep int32 tables, or
 * 16-bit cells when an instance's distance bound B (the sum over its
 * singleton rows of the row maximum) is below 0xFFFF.  A graph whose
 * n * max w reaches 2^32 - 1 is refused with EMV_ERR_RANGE at submission (the
 * singleton stage relaxes on 32-bit labels); an instance whose bound B reaches
 * 2^31 - 1 ends with status EMV_ERR_RANGE after its singleton stage
 * (emv_solve, with its int64 table, handles both).
 */
#ifndef EMV_ENGINE_H
#define EMV_ENGINE_H

#include <stdint.h>

#include "emv_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct emv_engine emv_engine;

/* One batched instance: its own graph (edge list as graph_t.edges,
 * reader-el.c:358-373) and terminals. */
typedef struct {
    int64_t n;
    int64_t m;
    const int64_t *edges; /* 3*m: u, v, w (0-based) */
    int64_t k;
    const int64_t *kk;
} emv_instance;

/* Tunables (emv_engine_default_params fills the defaults). */
typedef struct {
    int32_t relax_blocks_per_cu;  /* relaxation grid = CUs * this (1..256), default 64 */
    int32_t merge_rows_per_wave;  /* 1, 2, 4, 8, or 0 = by layer size */
    int32_t use_graph;            /* batch: replay steps 2..kmax from a hipGraph */
    int32_t graph_round_margin;   /* rounds added to each replayed step's recorded count
                                     (negative values force short replays, for tests) */
    int32_t dense_rounds;         /* vertex-wise relaxation: rounds 1..R of a layer step take
                                     every item and rounds 0..R-1 write no flags; 0 = default 6 */
    int32_t split_items;          /* split merge of the upper layers: items per merge tile aimed
                                     at (parts per row = items / rows); 0 = default 32 */
} emv_engine_params;

typedef struct {
    double total_ms;
    int64_t steps;
    int64_t relax_rounds;
    int64_t dp_cells;       /* sum over instances of 2^k * n */
    int32_t graph_used;     /* the solve was a graph replay */
    int32_t graph_fallback; /* a replay ran out of rounds and the solve ran eagerly */
} emv_batch_stats;

typedef struct {
    int32_t state;      /* 0 queued, 1 running, 2 done */
    int32_t status;     /* EMV_OK or the error that ended the request */
    int64_t cost;       /* f[C][q] once done */
    int32_t k;
    int32_t stage;      /* next DP stage: 1 singletons, m >= 2 layer m */
    double queue_ms;    /* submit -> admission */
    double latency_ms;  /* submit -> done */
    int64_t npairs;     /* traceback pairs held (EMV_LIST) */
    int64_t reduced_cost; /* group query: cost + groups * M, the cost of the reduced instance
                             (virtual terminals joined to their members by weight M =
                             emv_server_group_weight()); terminal query: cost */
    int32_t groups;     /* group query: number of groups (= k); 0 for a terminal query */
    int32_t reserved;
    int64_t root;       /* vertex the traceback starts at: kk[k-1], or for a group query the
                           smallest member of the last group at least cost */
} emv_request_info;

int emv_engine_default_params(emv_engine_params *p);
int emv_engine_set_params(emv_engine *e, const emv_engine_params *p);
const char *emv_engine_last_error(const emv_engine *e);
void emv_engine_destroy(emv_engine *e);

/* ---- batch: instances admitted together, stepped in lockstep ---------- */
int emv_batch_create(int device, int64_t ninst, const emv_instance *inst, emv_engine **out);
/* Solves every instance; costs[i] for instance i.  flags: EMV_LIST,
 * EMV_KEEP_TABLE (emv_hip.h).  The first solve runs eagerly and records the
 * relaxation rounds each step needed; later solves replay a hipGraph when
 * use_graph is set. */
int emv_batch_solve(emv_engine *e, uint32_t flags, int64_t *costs, emv_batch_stats *st);

/* ---- server: one resident graph, requests admitted as memory frees ---- */
/* arena_bytes: device memory for request tables (0 = 60% of free memory);
 * max_inflight: requests that may run at once. */
int emv_server_create(int device, int64_t n, int64_t m, const int64_t *edges,
                      int64_t arena_bytes, int max_inflight, emv_engine **out);
/* Queues a terminal-set query.  EMV_ERR_ADMISSION if its table footprint
 * 2 * 2^k * np * 4 bytes can never fit the arena. */
int emv_server_submit(emv_engine *e, int64_t k, const int64_t *kk, uint32_t flags,
                      int64_t *ticket);
/* Queues a group Steiner query on the resident graph: g groups (2..30),
 * group j = members[offsets[j] .. offsets[j+1]).  The reference solves such
 * a query through a reduction (README.txt:4-5): one virtual terminal per
 * group, joined to each member by an edge of weight M = 1 + the sum of all
 * edge weights, then an ordinary solve on the reduced STP.  Here the DP runs
 * on the resident graph itself, with singleton row {j} = the distance to the
 * nearest member of group j: M exceeds every tree of the graph, so an optimal
 * reduced tree uses each virtual terminal as a leaf and the reduced table at
 * a graph vertex is |X| * M + this table.  info.cost is the group Steiner
 * cost, info.reduced_cost = cost + g * M the reduced instance's cost; the
 * traceback lists graph edges only (the tree without its virtual leaves). */
int emv_server_submit_groups(emv_engine *e, int64_t g, const int64_t *offsets,
                             const int64_t *members, uint32_t flags, int64_t *ticket);
/* M = 1 + the sum of the resident graph's edge weights. */
int64_t emv_server_group_weight(const emv_engine *e);
/* Admits queued requests that fit (FIFO), then advances every running
 * request by one DP stage in shared launches. */
int emv_server_step(emv_engine *e, int64_t *running, int64_t *queued);
int emv_server_poll(emv_engine *e, int64_t ticket, emv_request_info *info);
/* Cancels a queued or running request: it leaves the queue or the running
 * set, its table region and slot are freed for the next admission, and it
 * ends with state 2 and status EMV_ERR_CANCELLED.  EMV_ERR_STATE for a
 * finished request or an engine made by emv_batch_create.  (The reference
 * solves one instance per process and has no request to cancel.) */
int emv_server_cancel(emv_engine *e, int64_t ticket);
int64_t emv_server_free_bytes(const emv_engine *e);
int64_t emv_request_footprint(const emv_engine *e, int64_t k);

/* ---- results of batch instances (ticket = index) and server requests --- */
int emv_engine_read_rows(emv_engine *e, int64_t ticket, int64_t X0, int64_t nrows,
                         int64_t *f_out, int64_t *code_out);
int emv_engine_pairs(emv_engine *e, int64_t ticket, int64_t *pairs, int64_t cap,
                     int64_t *npairs);
/* Frees a finished request's table kept with EMV_KEEP_TABLE. */
int emv_engine_release(emv_engine *e, int64_t ticket);
/* Certificate of a finished instance's kept table (EMV_KEEP_TABLE), as
 * emv_certify (emv_hip.h); a group query's singleton rows carry the distance
 * to the nearest member and edge records (2^k + the smallest neighbour on a
 * shortest path), {j} at the members. */
int emv_engine_certify(emv_engine *e, int64_t ticket, emv_certificate *out);

/* Test hook: the step allocation `after` allocations from now (0 = the next
 * one) fails as if the device were out of memory, without calling hipMalloc;
 * -1 disarms it.  Step allocations (staging, flags, merge accumulators,
 * descriptors) all happen before a step launches anything, so the failing
 * emv_server_step / emv_batch_solve returns EMV_ERR_NOMEM with every request
 * still at its stage; a later call carries on (a batch solve starts over). */
int emv_engine_fail_alloc(emv_engine *e, int64_t after);

#ifdef __cplusplus
}
#endif

#endif /* EMV_ENGINE_H */
