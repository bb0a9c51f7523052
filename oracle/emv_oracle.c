/*
 * emv_oracle.c -- CPU restatement of the Erickson-Monma-Veinott (EMV) Steiner
 * tree dynamic program of reader-el (test infrastructure only).
 *
 * ORACLE / TEST INFRASTRUCTURE.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load this library, and only as the checker.
 * The product path (libemvhip.so, emv-reader) never links or calls it.
 *
 * What it restates (semantics, written independently):
 *   reader/reader-el.c:1964-2139  emv_kernel: singleton SSSPs, then layers
 *                                 m = 2..k; per X a subset merge over Y in
 *                                 increasing order with strict '<', then a
 *                                 multi-source shortest-path relaxation
 *                                 seeded with the merged row.
 *   reader/reader-el.c:1762-1880  dijkstra (positive integer weights).
 *   reader/reader-el.c:1907-1944  backtrack / build_tree (traceback order).
 *   reader/reader-el.c:2137,2263  answer cell f[C][q], C = 2^(k-1)-1,
 *                                 q = kk[k-1].
 *   reader/reader-el.c:2159-2209  k == 2: one SSSP from kk[0], cost d[kk[1]].
 *
 * Record ("argmin") encoding shared with the GPU library, one int64 per cell:
 *   row X = 0            : -1
 *   merge record (v, Y)  : Y                  (0 < Y <= X, Y == X terminates)
 *   edge record  (u, X)  : 2^k + u            (u = predecessor vertex)
 * The reference stores the pair (u, Y) per cell (reader-el.c:692,2084-2085,
 * 2121-2125); a merge record has u == v, an edge record has Y == X.
 *
 * Tie rules.  The merge keeps the first strict minimum over increasing Y
 * (reader-el.c:2070-2078), i.e. the numerically smallest minimising Y.  The
 * reference keeps the merge record exactly when the relaxed distance equals
 * the merged value (its predecessor is the virtual source, 2122); otherwise
 * its predecessor is whichever neighbour the binary heap popped first, which
 * depends on heap order and thread count.  This oracle (and the GPU) pick the
 * smallest neighbour id u with f[X][u] + w(u,v) == f[X][v] instead -- the
 * only place where records may legitimately differ from the reference's.
 *
 * The DP table itself is unique: for a connected graph with positive weights
 * f[X][v] is the cost of a minimum Steiner tree spanning {kk[t]: t in X} and v.
 *
 * Parity pins (tests/test_oracle.py): golden table digests and costs produced
 * by the reference itself (oracle/_ref), SteinLib cost lines, scipy Dijkstra
 * for the singleton rows, brute-force Steiner trees on tiny graphs, the
 * fixed-point certificate and a tree validator.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <stdio.h>

#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_INF INT64_MAX

typedef struct {
    int64_t n;
    int64_t *rp;   /* n+1 row offsets */
    int64_t *nb;   /* neighbour ids */
    int64_t *wt;   /* weights */
} orc_csr;

/* Adjacency of both directions in edge-list order, as reader-el.c:798-826
 * builds it (multi-edges and self-loops kept as given). */
static int csr_build(orc_csr *g, int64_t n, int64_t m, const int64_t *e)
{
    g->n = n;
    g->rp = calloc((size_t)n + 1, sizeof(int64_t));
    g->nb = malloc(sizeof(int64_t) * (size_t)(2 * m + 1));
    g->wt = malloc(sizeof(int64_t) * (size_t)(2 * m + 1));
    if (!g->rp || !g->nb || !g->wt) return -1;
    for (int64_t j = 0; j < m; j++) {
        int64_t u = e[3 * j], v = e[3 * j + 1];
        if (u < 0 || v < 0 || u >= n || v >= n) return -2;
        g->rp[u + 1]++;
        g->rp[v + 1]++;
    }
    for (int64_t u = 0; u < n; u++) g->rp[u + 1] += g->rp[u];
    int64_t *fill = malloc(sizeof(int64_t) * (size_t)n);
    if (!fill) return -1;
    memcpy(fill, g->rp, sizeof(int64_t) * (size_t)n);
    for (int64_t j = 0; j < m; j++) {
        int64_t u = e[3 * j], v = e[3 * j + 1], w = e[3 * j + 2];
        g->nb[fill[u]] = v; g->wt[fill[u]++] = w;
        g->nb[fill[v]] = u; g->wt[fill[v]++] = w;
    }
    free(fill);
    return 0;
}

static void csr_free(orc_csr *g)
{
    free(g->rp); free(g->nb); free(g->wt);
}

/* ------------------------------------------------------------------------
 * Indexed binary min-heap keyed by int64 distance (textbook, decrease-key).
 * ------------------------------------------------------------------------ */
typedef struct {
    int64_t size;
    int64_t *item;  /* heap slot -> vertex (1-based slots) */
    int64_t *slot;  /* vertex -> heap slot, 0 when absent */
    const int64_t *key;
} orc_heap;

static void heap_up(orc_heap *h, int64_t i)
{
    int64_t x = h->item[i];
    while (i > 1) {
        int64_t p = i >> 1;
        int64_t y = h->item[p];
        if (h->key[y] <= h->key[x]) break;
        h->item[i] = y; h->slot[y] = i;
        i = p;
    }
    h->item[i] = x; h->slot[x] = i;
}

static void heap_down(orc_heap *h, int64_t i)
{
    int64_t x = h->item[i];
    for (;;) {
        int64_t c = 2 * i;
        if (c > h->size) break;
        if (c < h->size && h->key[h->item[c + 1]] < h->key[h->item[c]]) c++;
        if (h->key[h->item[c]] >= h->key[x]) break;
        h->item[i] = h->item[c]; h->slot[h->item[i]] = i;
        i = c;
    }
    h->item[i] = x; h->slot[x] = i;
}

static void heap_push_or_decrease(orc_heap *h, int64_t v)
{
    if (h->slot[v] == 0) {
        h->item[++h->size] = v;
        h->slot[v] = h->size;
    }
    heap_up(h, h->slot[v]);
}

static int64_t heap_pop(orc_heap *h)
{
    int64_t top = h->item[1];
    h->slot[top] = 0;
    int64_t last = h->item[h->size--];
    if (h->size > 0) {
        h->item[1] = last; h->slot[last] = 1;
        heap_down(h, 1);
    }
    return top;
}

static inline int64_t sat_add(int64_t a, int64_t b)
{
    if (a == ORC_INF || b == ORC_INF) return ORC_INF;
    return a + b;
}

/* Multi-source Dijkstra: d[] holds the initial labels (the virtual source's
 * edge weights, reader-el.c:2096-2109); on return d[] holds the shortest
 * distances.  Scratch: item/slot arrays of n+1 entries. */
static void sssp_multi(const orc_csr *g, int64_t *d, int64_t *item, int64_t *slot)
{
    orc_heap h = { 0, item, slot, d };
    memset(slot, 0, sizeof(int64_t) * (size_t)g->n);
    for (int64_t v = 0; v < g->n; v++)
        if (d[v] != ORC_INF) heap_push_or_decrease(&h, v);
    while (h.size > 0) {
        int64_t u = heap_pop(&h);
        int64_t du = d[u];
        for (int64_t i = g->rp[u]; i < g->rp[u + 1]; i++) {
            int64_t v = g->nb[i];
            int64_t c = du + g->wt[i];
            if (c < d[v]) {
                d[v] = c;
                heap_push_or_decrease(&h, v);
            }
        }
    }
}

/* Smallest neighbour u with d[u] + w(u,v) == d[v], or -1. */
static int64_t min_pred(const orc_csr *g, const int64_t *d, int64_t v)
{
    int64_t best = -1;
    for (int64_t i = g->rp[v]; i < g->rp[v + 1]; i++) {
        int64_t u = g->nb[i];
        if (sat_add(d[u], g->wt[i]) == d[v] && (best < 0 || u < best)) best = u;
    }
    return best;
}

/* ------------------------------------------------------------------------ */

int oracle_sssp(int64_t n, int64_t m, const int64_t *edges, int64_t s,
                int64_t *d, int64_t *pred)
{
    orc_csr g;
    if (csr_build(&g, n, m, edges)) { csr_free(&g); return -1; }
    int64_t *item = malloc(sizeof(int64_t) * (size_t)(n + 1));
    int64_t *slot = malloc(sizeof(int64_t) * (size_t)(n + 1));
    for (int64_t v = 0; v < n; v++) d[v] = ORC_INF;
    d[s] = 0;
    sssp_multi(&g, d, item, slot);
    if (pred)
        for (int64_t v = 0; v < n; v++)
            pred[v] = (v == s) ? -1 : min_pred(&g, d, v);
    free(item); free(slot);
    csr_free(&g);
    return 0;
}

/* Full EMV table.  f and b are 2^k * n int64 arrays, subset-major
 * (reader-el.c:691-692).  Returns 0 and writes the answer f[C][q] to *cost. */
int oracle_emv(int64_t n, int64_t m, const int64_t *edges, int64_t k,
               const int64_t *kk, int64_t *f, int64_t *b, int64_t *cost,
               int nthreads)
{
    if (k < 1 || k > 30 || n < 1) return -1;
    orc_csr g;
    if (csr_build(&g, n, m, edges)) { csr_free(&g); return -2; }
    const int64_t S = (int64_t)1 << k;
    const int64_t PRED = S;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif

    for (int64_t X = 0; X < S; X++) {
        for (int64_t v = 0; v < n; v++) {
            f[X * n + v] = ORC_INF;
            b[X * n + v] = -1;
        }
    }

    int err = 0;
    /* Singleton rows: distances from each terminal; record (kk[t], {t})
     * (reader-el.c:2015-2027). */
#pragma omp parallel
    {
        int64_t *item = malloc(sizeof(int64_t) * (size_t)(n + 1));
        int64_t *slot = malloc(sizeof(int64_t) * (size_t)(n + 1));
#pragma omp for schedule(dynamic, 1)
        for (int64_t t = 0; t < k; t++) {
            int64_t X = (int64_t)1 << t;
            int64_t *row = f + X * n;
            for (int64_t v = 0; v < n; v++) row[v] = ORC_INF;
            row[kk[t]] = 0;
            sssp_multi(&g, row, item, slot);
            for (int64_t v = 0; v < n; v++)
                b[X * n + v] = (v == kk[t]) ? X : PRED + kk[t];
        }
        free(item); free(slot);
    }

    /* Layers by subset size (reader-el.c:2033-2134).  Rows of one layer only
     * read rows of smaller layers, so X order inside a layer is free. */
    for (int64_t sz = 2; sz <= k && !err; sz++) {
#pragma omp parallel
        {
            int64_t *item = malloc(sizeof(int64_t) * (size_t)(n + 1));
            int64_t *slot = malloc(sizeof(int64_t) * (size_t)(n + 1));
            int64_t *mrg = malloc(sizeof(int64_t) * (size_t)n);
#pragma omp for schedule(dynamic, 1)
            for (int64_t X = 0; X < S; X++) {
                if (__builtin_popcountll((unsigned long long)X) != sz) continue;
                int64_t *fx = f + X * n;
                int64_t *bx = b + X * n;
                /* merge: Y over nonempty proper subsets in increasing order,
                 * strict '<' keeps the first minimum (2070-2088). */
                for (int64_t Y = X & (-X); Y != X; Y = X & (Y - X)) {
                    const int64_t *fy = f + Y * n;
                    const int64_t *fz = f + (X ^ Y) * n;
                    for (int64_t v = 0; v < n; v++) {
                        int64_t s = sat_add(fy[v], fz[v]);
                        if (s < fx[v]) { fx[v] = s; bx[v] = Y; }
                    }
                }
                /* relax: multi-source SSSP seeded with the merged row
                 * (2090-2116); the terminal override of 2100-2107 writes
                 * f[X-{t}][kk[t]], which equals the merged value there. */
                memcpy(mrg, fx, sizeof(int64_t) * (size_t)n);
                sssp_multi(&g, fx, item, slot);
                for (int64_t v = 0; v < n; v++) {
                    if (fx[v] < mrg[v]) bx[v] = PRED + min_pred(&g, fx, v);
                }
            }
            free(item); free(slot); free(mrg);
        }
    }

    if (k == 1) *cost = f[0 * n + kk[0]];          /* row C = 0 is never written */
    else *cost = f[(((int64_t)1 << (k - 1)) - 1) * n + kk[k - 1]];
    csr_free(&g);
    return err;
}

/* Traceback over the record table, in the reference's depth-first order
 * (reader-el.c:1907-1944).  Writes 0-based (v,u) pairs; returns their count,
 * or -1 when cap is exceeded. */
int64_t oracle_traceback(int64_t n, int64_t k, const int64_t *b, int64_t q,
                         int64_t C, int64_t *pairs, int64_t cap)
{
    const int64_t S = (int64_t)1 << k;
    int64_t sp = 0, np = 0, stcap = 64 + 4 * k + 2 * n;
    int64_t *st = malloc(sizeof(int64_t) * 2 * (size_t)stcap);
    st[0] = q; st[1] = C; sp = 1;
    while (sp > 0) {
        sp--;
        int64_t v = st[2 * sp], X = st[2 * sp + 1];
        if (X == 0 || v < 0) continue;
        int64_t r = b[X * n + v];
        if (r < 0) continue;
        if (r >= S) {
            int64_t u = r - S;
            if (np >= cap) { free(st); return -1; }
            pairs[2 * np] = v; pairs[2 * np + 1] = u; np++;
            st[2 * sp] = u; st[2 * sp + 1] = X; sp++;
        } else {
            int64_t Y = r;
            if (Y == X) continue;
            if (sp + 2 > stcap) { free(st); return -1; }
            /* (v, Y) is visited before (v, X\Y): push it last. */
            st[2 * sp] = v; st[2 * sp + 1] = X & ~Y; sp++;
            st[2 * sp] = v; st[2 * sp + 1] = Y; sp++;
        }
    }
    free(st);
    return np;
}
