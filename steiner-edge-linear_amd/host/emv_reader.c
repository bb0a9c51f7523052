/*
 * emv_reader.c -- the reader-el command line with the EMV dynamic program on
 * an MI355X.  Same flags, same input format and same status lines as the
 * reference (reader/reader-el.c:2345-2536), so harnesses that parse the
 * reference's stdout (verify.py, report.py) read this one unchanged.  The
 * erickson_monma_veinott call (reader-el.c:2141) is replaced by the C ABI of
 * libemvhip.so (include/emv_hip.h); new information goes on lines with a new
 * prefix ("gpu:", "group:").
 *
 *   emv-reader -in <file.stp> -el [-list] [-seed s] [-dijkstra]
 *              [-groups <file>] [-device d] [-dump <table file>]
 */
#define _GNU_SOURCE
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/utsname.h>

#include "emv_hip.h"
#include "stp.h"

#define CMD_NOP 0
#define CMD_DIJKSTRA 1
#define CMD_EDGE_LINEAR 2

static const char *cmd_legend[] = {"no operation", "Dijkstra Single-Source-Shortest-Path",
                                   "Erickson-Monma-Veinott"};

static emv_ctx *ctx = NULL;
static const char *dumpfile = NULL; /* -dump: canonical table file after the solve */
static double ctx_init_ms = 0.0;
static int device_id = 0;

static void gpu_check(int rc, const char *what)
{
    if (rc != EMV_OK) FAIL("%s failed (%d): %s", what, rc, ctx ? emv_last_error(ctx) : "no context");
}

static void gpu_open(void)
{
    if (ctx) return;
    time_push();
    int rc = emv_open(&ctx, device_id);
    if (rc != EMV_OK) FAIL("emv_open(device %d) failed (%d)", device_id, rc);
    ctx_init_ms = time_pop();
}

static void print_pairs(const char *prefix, const index_t *pairs, index_t np)
{
    printf("%s [", prefix);
    for (index_t i = 0; i < np; i++)
        printf("%s\"%ld %ld\"", i ? ", " : "", pairs[2 * i] + 1, pairs[2 * i + 1] + 1);
    printf("]\n");
}

/* Path from v back to the source s along predecessors, as (v, p[v]) pairs
 * (reader-el.c:1946-1960). */
static index_t tracepath(index_t s, index_t v, const index_t *pred, index_t *pairs, index_t cap)
{
    index_t np = 0;
    while (v != s && np < cap) {
        index_t u = pred[v];
        if (u < 0) break;
        pairs[2 * np] = v;
        pairs[2 * np + 1] = u;
        np++;
        v = u;
    }
    return np;
}

typedef struct {
    index_t cost;
    index_t npairs;
    index_t *pairs; /* tracked */
} el_out;

/* The drop-in for erickson_monma_veinott (reader-el.c:2141-2343). */
static el_out erickson_gpu(const stp_query *q, int list)
{
    el_out out = {0, 0, NULL};
    memtrack_push();
    time_push();
    gpu_open();
    time_push();
    gpu_check(emv_load_graph(ctx, q->n, q->m, q->pos, q->adj), "emv_load_graph");
    double upload_ms = time_pop();

    if (q->k == 2) {
        index_t *d = tmalloc(sizeof(index_t) * (size_t)q->n);
        index_t *p = tmalloc(sizeof(index_t) * (size_t)q->n);
        printf("erickson: ");
        double ms = 0;
        time_push();
        gpu_check(emv_sssp(ctx, q->kk[0], d, p, &ms), "emv_sssp");
        double wall = time_pop();
        /* bytes of one SSSP pass: labels read+written, CSR read once */
        double bytes = 16.0 * (double)q->n + 16.0 * (double)q->m;
        printf("[kernel: %.2lf ms %.2lfGiB/s] ", wall, bytes / (wall / 1e3) / (double)(1 << 30));
        out.cost = d[q->kk[1]];
        if (list) {
            out.pairs = tmalloc(sizeof(index_t) * 2 * (size_t)(q->n + 1));
            out.npairs = tracepath(q->kk[0], q->kk[1], p, out.pairs, q->n + 1);
        }
        tfree(d);
        tfree(p);
        printf("done. [%.2lf ms] [cost: %ld] ", time_pop(), out.cost);
        printf("{peak: %.2lfGiB} ", memtrack_pop_gib());
        printf("{curr: %.2lfGiB}\n", memtrack_curr_gib());
        if (list) print_pairs("solution:", out.pairs, out.npairs);
        printf("gpu: [device: %d] [init: %.2lf ms] [upload: %.2lf ms] [sssp: %.2lf ms]\n",
               device_id, ctx_init_ms, upload_ms, ms);
        fflush(stdout);
        return out;
    }

    emv_result res;
    gpu_check(emv_solve(ctx, q->k, q->kk, (list ? EMV_LIST : 0u) | (dumpfile ? EMV_KEEP_TABLE : 0u),
                        &res),
              "emv_solve");
    double bytes = res.merge_alg_bytes + res.relax_alg_bytes;
    printf("erickson: [zero: %.2lf ms] [kernel: %.2lf ms %.2lfGiB/s] ", res.zero_ms,
           res.kernel_ms, bytes / (res.kernel_ms / 1e3) / (double)(1 << 30));
    if (list) printf("[traceback: %.2lf ms] ", res.traceback_ms);
    out.cost = res.cost;
    if (list) {
        out.npairs = res.npairs;
        out.pairs = tmalloc(sizeof(index_t) * 2 * (size_t)(res.npairs + 1));
        memcpy(out.pairs, res.pairs, sizeof(index_t) * 2 * (size_t)res.npairs);
    }
    printf("done. [%.2lf ms] [cost: %ld] ", time_pop(), out.cost);
    printf("{peak: %.2lfGiB} ", memtrack_pop_gib());
    printf("{curr: %.2lfGiB}\n", memtrack_curr_gib());
    if (list) print_pairs("solution:", out.pairs, out.npairs);
    double cells = (double)res.dp_cells;
    printf("gpu: [device: %d] [init: %.2lf ms] [upload: %.2lf ms] [table: int%d] [merge acc: %d-bit] "
           "[singleton: %.2lf ms] [merge: %.2lf ms %.2lf GB/s] [relax: %.2lf ms %.2lf GB/s %ld rounds] "
           "[records: %.2lf ms] [dp cells: %.0lf] [cells/s: %.4e] [device mem: %.2lfGiB]\n",
           device_id, ctx_init_ms, upload_ms, 8 * res.table_bytes, res.merge_bits, res.singleton_ms,
           res.merge_ms, res.merge_ms > 0 ? res.merge_alg_bytes / (res.merge_ms * 1e6) : 0.0,
           res.relax_ms, res.relax_ms > 0 ? res.relax_alg_bytes / (res.relax_ms * 1e6) : 0.0,
           (long)res.relax_rounds, res.unpack_ms, cells, cells / (res.kernel_ms / 1e3),
           (double)res.device_bytes / (double)(1 << 30));
    if (dumpfile) {
        gpu_check(emv_write_table(ctx, dumpfile, EMV_DUMP_RECORDS), "emv_write_table");
        gpu_check(emv_release_table(ctx), "emv_release_table");
        printf("dump: [file: %s] [k: %ld] [n: %ld] [rows: %ld] [records: int32]\n", dumpfile,
               (long)q->k, (long)q->n, (long)(1l << q->k));
    }
    fflush(stdout);
    emv_result_free(&res);
    return out;
}

/* Group mode: replace every shortcut pair (v, virtual terminal t) by the
 * real shortest path from v to group t, drop the virtual edge, and report
 * the group tree with cost OPT - g*M. */
static void report_group(const stp_query *q, index_t n_orig, index_t g, index_t M,
                         const el_out *o)
{
    index_t gcost = o->cost - g * M;
    printf("group: [groups: %ld] [M: %ld] [cost: %ld]\n", g, M, gcost);
    if (!o->pairs) {
        fflush(stdout);
        return;
    }
    index_t cap = 2 * (o->npairs + q->n + 1);
    index_t *edges = tmalloc(sizeof(index_t) * 2 * (size_t)cap);
    index_t ne = 0;
    index_t *d = tmalloc(sizeof(index_t) * (size_t)q->n);
    index_t *p = tmalloc(sizeof(index_t) * (size_t)q->n);
    index_t last_src = -1;
    for (index_t i = 0; i < o->npairs; i++) {
        index_t v = o->pairs[2 * i], u = o->pairs[2 * i + 1];
        int v_virt = v >= n_orig, u_virt = u >= n_orig;
        if (!v_virt && !u_virt) {
            /* a real graph edge (edge record) or a shortcut between real
             * vertices cannot occur: shortcuts always end at a terminal */
            edges[2 * ne] = v;
            edges[2 * ne + 1] = u;
            ne++;
            continue;
        }
        /* a pair touching a virtual terminal: the shortest path from v to u
         * (an edge record member-terminal yields no real edge, a shortcut
         * (x, t) yields the real path from x into group t) */
        if (u != last_src) {
            gpu_check(emv_sssp(ctx, u, d, p, NULL), "emv_sssp");
            last_src = u;
        }
        index_t y = v;
        while (y != u) {
            index_t z = p[y];
            if (z < 0) FAIL("no path from %ld to %ld", y + 1, u + 1);
            if (z < n_orig && y < n_orig) {
                if (ne >= cap) FAIL("group tree overflow");
                edges[2 * ne] = y;
                edges[2 * ne + 1] = z;
                ne++;
            }
            y = z;
        }
    }
    print_pairs("group solution:", edges, ne);
    fflush(stdout);
    tfree(d);
    tfree(p);
    tfree(edges);
}

int main(int argc, char **argv)
{
    time_push();
    memtrack_push();
    long seed = 123456789;
    int have_seed = 0, cmd = CMD_NOP, list = 0;
    const char *filename = NULL, *groupfile = NULL;
    for (int f = 1; f < argc; f++) {
        if (argv[f][0] != '-') continue;
        if (!strcmp(argv[f], "-dijkstra")) cmd = CMD_DIJKSTRA;
        if (!strcmp(argv[f], "-el") || !strcmp(argv[f], "-erickson")) cmd = CMD_EDGE_LINEAR;
        if (!strcmp(argv[f], "-list")) list = 1;
        if (!strcmp(argv[f], "-in")) {
            if (f == argc - 1) FAIL("file name missing from command line");
            filename = argv[++f];
        }
        if (!strcmp(argv[f], "-seed")) {
            if (f == argc - 1) FAIL("random seed missing from command line");
            seed = atol(argv[++f]);
            have_seed = 1;
        }
        if (!strcmp(argv[f], "-groups")) {
            if (f == argc - 1) FAIL("groups file missing from command line");
            groupfile = argv[++f];
        }
        if (!strcmp(argv[f], "-dump")) {
            if (f == argc - 1) FAIL("table file missing from command line");
            dumpfile = argv[++f];
        }
        if (!strcmp(argv[f], "-device")) {
            if (f == argc - 1) FAIL("device id missing from command line");
            device_id = atoi(argv[++f]);
        }
        if (!strcmp(argv[f], "-h") || !strcmp(argv[f], "-help")) {
            printf("usage: %s -in <input graph> <arguments>\n"
                   "\n"
                   "arguments :\n"
                   "\t-seed : seed value\n"
                   "\t-el : Erickson-Monma-Veinott algorithm (on the GPU)\n"
                   "\t-dijkstra : single source shortest path (on the GPU)\n"
                   "\t-list : Output Steiner tree\n"
                   "\t-groups <file> : group Steiner tree by reduction (lines \"g v1 v2 ...\")\n"
                   "\t-device <id> : HIP device\n"
                   "\t-dump <file> : write the DP table and records as a table file (with -el)\n"
                   "\n",
                   argv[0]);
            return 0;
        }
    }
    printf("invoked as:");
    for (int f = 0; f < argc; f++) printf(" %s", argv[f]);
    printf("\n");
    FILE *in = stdin;
    if (filename) {
        in = fopen(filename, "r");
        if (!in) FAIL("unable to open file '%s'", filename);
    } else {
        printf("no input file specified, defaulting to stdin\n");
    }
    if (!have_seed) printf("no random seed given, defaulting to %ld\n", seed);
    printf("random seed = %ld\n", seed);

    stp_graph *g = stp_load(in);
    if (in != stdin) fclose(in);
    index_t n_orig = g->n, ngroups = 0, M = 0;
    if (groupfile) {
        FILE *gf = fopen(groupfile, "r");
        if (!gf) FAIL("unable to open groups file '%s'", groupfile);
        stp_groups *gr = groups_load(gf, g->n);
        fclose(gf);
        stp_graph *r = group_reduce(g, gr, &M);
        ngroups = gr->g;
        printf("group: reduced [n: %ld] [m: %ld] [k: %ld] [M: %ld]\n", r->n, r->m, r->k, M);
        groups_free(gr);
        stp_free(g);
        g = r;
    }
    stp_query *q = query_build(g);
    index_t min_cost = g->cost;
    stp_free(g);

    printf("command: %s\n", cmd_legend[cmd]);
    fflush(stdout);
    time_push();
    switch (cmd) {
    case CMD_DIJKSTRA: {
        srand((unsigned)seed);
        index_t s = rand() % q->n;
        gpu_open();
        gpu_check(emv_load_graph(ctx, q->n, q->m, q->pos, q->adj), "emv_load_graph");
        double ms = 0;
        gpu_check(emv_sssp(ctx, s, NULL, NULL, &ms), "emv_sssp");
        printf("gpu: [device: %d] [dijkstra source: %ld] [sssp: %.2lf ms]\n", device_id, s + 1, ms);
        break;
    }
    case CMD_EDGE_LINEAR: {
        el_out o = erickson_gpu(q, list);
        if (groupfile) report_group(q, n_orig, ngroups, M, &o);
        if (min_cost != -1 && min_cost != o.cost)
            FAIL("min_cost != cost: minimum cost = %ld, cost = %ld", min_cost, o.cost);
        tfree(o.pairs);
        break;
    }
    default:
        break;
    }
    query_free(q);
    if (ctx) emv_close(ctx);

    printf("command done [%.2lf ms]\n", time_pop());
    printf("grand total [%.2lf ms] ", time_pop());
    printf("{peak: %.2lfGiB}\n", memtrack_pop_gib());
    struct utsname un;
    uname(&un);
    printf("host: %s\n", un.nodename);
    printf("build: edge-linear kernel, gfx950 device, frontier shortest paths\n");
    printf("list solution: %s\n", list ? "true" : "false");
    printf("num threads: %d\n", 1);
    printf("compiler: gcc %d.%d.%d\n", __GNUC__, __GNUC_MINOR__, __GNUC_PATCHLEVEL__);
    fflush(stdout);
    if (tmalloc_balance() != 0) FAIL("host allocation balance %ld != 0", tmalloc_balance());
    if (time_depth() != 0) FAIL("timer stack not empty");
    return 0;
}
