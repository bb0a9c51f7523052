/*
 * ref_dump.c -- golden-vector harness around the reference reader-el.
 *
 * ORACLE / TEST INFRASTRUCTURE.  Linked (by oracle/Makefile) against the
 * reference source reader/reader-el.c as it lies under the reference tree,
 * compiled with -Dmain=reader_el_main.  It loads an STP file with the
 * reference's own graph_load/root_build (reader-el.c:473,706), allocates and
 * initialises the tables as the reference driver does (reader-el.c:2212-2259)
 * and calls the reference emv_kernel (reader-el.c:1966) directly, then writes
 * the full DP table and the argmin records to a binary file:
 *
 *   int64 magic 0x454d5652454631 ("EMVREF1"), k, n, cost
 *   int64 f[2^k][n]       (INF = INT64_MAX, row X = 0 included)
 *   int64 code[2^k][n]    record (u, Y) re-encoded: -1 when unset,
 *                         Y when u == v (merge), 2^k + u otherwise (edge)
 *
 * Usage: ref_dump <in.stp> <out.bin>
 */
#include <stdio.h>
#include <stdlib.h>
#include <stdint.h>
#include <omp.h>

typedef long int index_t;

/* Layout of the reference's query struct (reader-el.c:696-704). */
typedef struct {
    index_t n, m, k;
    index_t *kk, *pos, *adj;
} ref_query;

struct graph;
struct graph *graph_load(FILE *in);
ref_query *root_build(struct graph *g);
index_t emv_kernel(index_t n, index_t m, index_t k, index_t c, index_t C,
                   index_t q, index_t *kk, index_t *f_v, index_t *pos,
                   index_t *adj, index_t *d, index_t *visit, index_t nt,
                   index_t *p, index_t *b_v);

int main(int argc, char **argv)
{
    if (argc != 3) { fprintf(stderr, "usage: %s in.stp out.bin\n", argv[0]); return 2; }
    FILE *in = fopen(argv[1], "r");
    if (!in) { perror(argv[1]); return 1; }
    /* The reference prints its status lines to stdout; keep them off ours. */
    FILE *devnull = freopen("/dev/null", "w", stdout);
    (void)devnull;
    struct graph *g = graph_load(in);
    fclose(in);
    ref_query *r = root_build(g);
    index_t n = r->n, m = r->m, k = r->k;
    index_t nt = omp_get_max_threads();
    index_t S = (index_t)1 << k;
    index_t *f = malloc(sizeof(index_t) * (size_t)(n * S));
    index_t *b = malloc(sizeof(index_t) * (size_t)(2 * n * S));
    index_t *d = malloc(sizeof(index_t) * (size_t)(nt * (n + nt)));
    index_t *visit = malloc(sizeof(index_t) * (size_t)(nt * (n + nt)));
    index_t *p = malloc(sizeof(index_t) * (size_t)(nt * (n + nt)));
    if (!f || !b || !d || !visit || !p) { fprintf(stderr, "out of memory\n"); return 1; }
    for (index_t i = 0; i < n * S; i++) { f[i] = INT64_MAX; b[2 * i] = -1; b[2 * i + 1] = 0; }
    index_t c = k - 1, C = ((index_t)1 << c) - 1, q = r->kk[k - 1];
    index_t cost = emv_kernel(n, m, k, c, C, q, r->kk, f, r->pos, r->adj, d, visit, nt, p, b);

    FILE *out = fopen(argv[2], "wb");
    if (!out) { perror(argv[2]); return 1; }
    int64_t hdr[4] = { 0x454d5652454631LL, k, n, cost };
    fwrite(hdr, sizeof(int64_t), 4, out);
    fwrite(f, sizeof(int64_t), (size_t)(n * S), out);
    int64_t *code = malloc(sizeof(int64_t) * (size_t)n);
    for (index_t X = 0; X < S; X++) {
        for (index_t v = 0; v < n; v++) {
            index_t u = b[2 * (X * n + v)], Y = b[2 * (X * n + v) + 1];
            if (X == 0 || u < 0) code[v] = -1;
            else if (u == v) code[v] = Y;
            else code[v] = S + u;
        }
        fwrite(code, sizeof(int64_t), (size_t)n, out);
    }
    fclose(out);
    return 0;
}
