/*
 * stp.c -- host utilities of emv-reader: error exit, tracked allocation,
 * timers, the DIMACS STP reader, the steinerq-layout query build and the
 * group Steiner reduction.
 *
 * The STP reader keeps the reference's accepted format and quirks
 * (reader-el.c:473-618): first token of each line decides; lower-case
 * keywords only; an "end" line must be exactly "end\n"; "section
 * coordinates", "dd" and unknown lines are ignored; ids are 1-based in the
 * file and 0-based in memory; the optional "cost" line is kept for the
 * self-check.
 */
#define _GNU_SOURCE
#include "stp.h"

#include <stdarg.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

void emv_fail(const char *file, int line, const char *func, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    fflush(stdout);
    fprintf(stderr, "ERROR [file = %s, line = %d] %s: ", file, line, func);
    vfprintf(stderr, fmt, ap);
    fprintf(stderr, "\n");
    va_end(ap);
    abort();
}

/* ------------------------------------------------------- tracked malloc */

typedef struct {
    size_t size;
    size_t pad; /* keeps the payload 16-byte aligned */
} alloc_hdr;

static size_t mem_total = 0;
static long mem_balance = 0;
#define MEMSTACK 256
static size_t mem_peak[MEMSTACK];
static int mem_top = -1;

void *tmalloc(size_t size)
{
    alloc_hdr *h = malloc(sizeof(alloc_hdr) + size);
    if (!h) FAIL("malloc fails (%zu bytes)", size);
    h->size = size;
    mem_total += size;
    mem_balance++;
    for (int i = 0; i <= mem_top; i++)
        if (mem_peak[i] < mem_total) mem_peak[i] = mem_total;
    return h + 1;
}

void tfree(void *p)
{
    if (!p) return;
    alloc_hdr *h = (alloc_hdr *)p - 1;
    mem_total -= h->size;
    mem_balance--;
    free(h);
}

long tmalloc_balance(void) { return mem_balance; }

void memtrack_push(void)
{
    if (mem_top + 1 >= MEMSTACK) FAIL("memory tracking stack overflow");
    mem_peak[++mem_top] = mem_total;
}

double memtrack_pop_gib(void)
{
    if (mem_top < 0) FAIL("memory tracking stack underflow");
    return (double)mem_peak[mem_top--] / (double)(1 << 30);
}

double memtrack_curr_gib(void) { return (double)mem_total / (double)(1 << 30); }

/* ------------------------------------------------------------- timers */

#define TIMESTACK 256
static double t_start[TIMESTACK];
static int t_top = -1;

static double now_ms(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec * 1e3 + (double)ts.tv_nsec * 1e-6;
}

void time_push(void)
{
    if (t_top + 1 >= TIMESTACK) FAIL("timer stack overflow");
    t_start[++t_top] = now_ms();
}

double time_pop(void)
{
    double t = now_ms();
    if (t_top < 0) FAIL("timer stack underflow");
    return t - t_start[t_top--];
}

int time_depth(void) { return t_top + 1; }

/* ---------------------------------------------------------- STP graph */

#define SEC_GRAPH 0x2u
#define SEC_TERMINALS 0x4u

stp_graph *stp_alloc(void)
{
    stp_graph *g = tmalloc(sizeof(stp_graph));
    memset(g, 0, sizeof(*g));
    g->cost = -1;
    g->edge_capacity = 128;
    g->edges = tmalloc(sizeof(index_t) * 3 * (size_t)g->edge_capacity);
    return g;
}

void stp_free(stp_graph *g)
{
    if (!g) return;
    tfree(g->edges);
    tfree(g->terminals);
    tfree(g);
}

void stp_add_edge(stp_graph *g, index_t u, index_t v, index_t w)
{
    if (u < 0 || v < 0 || u >= g->n || v >= g->n)
        FAIL("edge (%ld, %ld) outside 1..%ld", u + 1, v + 1, g->n);
    if (g->num_edges == g->edge_capacity) {
        index_t cap = 2 * g->edge_capacity;
        index_t *e = tmalloc(sizeof(index_t) * 3 * (size_t)cap);
        memcpy(e, g->edges, sizeof(index_t) * 3 * (size_t)g->num_edges);
        tfree(g->edges);
        g->edges = e;
        g->edge_capacity = cap;
    }
    index_t *e = g->edges + 3 * g->num_edges++;
    e[0] = u;
    e[1] = v;
    e[2] = w;
}

#define LINE_MAX_STP 1024

stp_graph *stp_load(FILE *in)
{
    time_push();
    memtrack_push();
    char line[LINE_MAX_STP], tok[LINE_MAX_STP];
    stp_graph *g = stp_alloc();
    int in_section = 0;
    while (fgets(line, sizeof(line), in)) {
        strcpy(tok, line);
        char *c = strtok(tok, " ");
        if (!c) continue;
        long a, b, w;
        if (!strcmp(c, "section")) {
            char name[LINE_MAX_STP];
            if (in_section) FAIL("nested sections");
            in_section = 1;
            if (sscanf(line, "section %1023s", name) != 1) FAIL("invalid section line");
            if (!strcmp(name, "comment")) {
            } else if (!strcmp(name, "graph")) {
                g->flags |= SEC_GRAPH;
            } else if (!strcmp(name, "terminals")) {
                g->flags |= SEC_TERMINALS;
            } else if (!strcmp(name, "coordinates")) {
            } else {
                FAIL("invalid section");
            }
        } else if (!strcmp(c, "end\n")) {
            if (!in_section) FAIL("no section to end");
            in_section = 0;
        } else if (!strcmp(c, "nodes")) {
            if (sscanf(line, "nodes %ld", &a) != 1) FAIL("invalid nodes line");
            g->n = a;
        } else if (!strcmp(c, "edges")) {
            if (sscanf(line, "edges %ld", &a) != 1) FAIL("invalid edges line");
            g->m = a;
        } else if (!strcmp(c, "terminals")) {
            if (sscanf(line, "terminals %ld", &a) != 1) FAIL("invalid terminals line");
            if (a < 0) FAIL("invalid terminals line");
            g->k = a;
            tfree(g->terminals);
            g->terminals = tmalloc(sizeof(index_t) * (size_t)(a > 0 ? a : 1));
            g->num_terminals = 0;
        } else if (!strcmp(c, "e")) {
            if (sscanf(line, "e %ld %ld %ld", &a, &b, &w) != 3) FAIL("invalid edge line %s", line);
            stp_add_edge(g, a - 1, b - 1, w);
        } else if (!strcmp(c, "t")) {
            if (sscanf(line, "t %ld", &a) != 1) FAIL("invalid terminal line %s", line);
            if (!g->terminals) FAIL("section terminals not initialised");
            if (a < 1 || a > g->n) FAIL("terminal %ld outside 1..%ld", a, g->n);
            if (g->num_terminals >= g->k) FAIL("more terminals than declared (%ld)", g->k);
            g->terminals[g->num_terminals++] = a - 1;
        } else if (!strcmp(c, "cost")) {
            if (sscanf(line, "cost %ld", &a) != 1) FAIL("invalid cost line %s", line);
            g->cost = a;
        }
        /* "dd", "eof", "coordinates" and anything else: ignored */
    }
    if (g->n == 0) FAIL("input declares no nodes");
    if (g->m != g->num_edges || g->m == 0)
        FAIL("edge count %ld does not match the %ld edge lines", g->m, g->num_edges);
    if (g->k != g->num_terminals || g->k == 0)
        FAIL("terminal count %ld does not match the %ld terminal lines", g->k, g->num_terminals);
    if (!(g->flags & SEC_GRAPH) || !(g->flags & SEC_TERMINALS))
        FAIL("input lacks a graph or terminals section");
    double ms = time_pop();
    printf("input: n = %ld, m = %ld, k = %ld, cost = %ld [%.2lf ms] ", g->n, g->m, g->k, g->cost,
           ms);
    printf("{peak: %.2lfGiB} ", memtrack_pop_gib());
    printf("{curr: %.2lfGiB}\n", memtrack_curr_gib());
    printf("terminals:");
    for (index_t i = 0; i < g->k; i++) printf(" %ld", g->terminals[i] + 1);
    printf("\n");
    fflush(stdout);
    return g;
}

/* ------------------------------------------------------------ query build */

stp_query *query_build(const stp_graph *g)
{
    memtrack_push();
    time_push();
    index_t n = g->n, m = g->m, k = g->k;
    stp_query *q = tmalloc(sizeof(stp_query));
    q->n = n;
    q->m = m;
    q->k = k;
    q->kk = tmalloc(sizeof(index_t) * (size_t)k);
    q->pos = tmalloc(sizeof(index_t) * (size_t)(n + 1));
    q->adj = tmalloc(sizeof(index_t) * (size_t)(n + 4 * m + 1));
    printf("root build: ");

    time_push();
    memset(q->pos, 0, sizeof(index_t) * (size_t)(n + 1));
    printf("[zero: %.2lf ms] ", time_pop());
    fflush(stdout);

    /* row u occupies 1 + 2*deg(u) slots: the degree, then (v, w) pairs */
    time_push();
    const index_t *e = g->edges;
    for (index_t j = 0; j < m; j++) {
        q->pos[e[3 * j]] += 2;
        q->pos[e[3 * j + 1]] += 2;
    }
    index_t run = 0;
    for (index_t u = 0; u < n; u++) {
        index_t len = q->pos[u] + 1;
        q->pos[u] = run;
        run += len;
    }
    q->pos[n] = run;
    if (run != n + 4 * m) FAIL("adjacency size %ld != n + 4m", run);
    printf("[pos: %.2lf ms] ", time_pop());
    fflush(stdout);

    time_push();
    for (index_t u = 0; u < n; u++) q->adj[q->pos[u]] = 0;
    for (index_t j = 0; j < m; j++) {
        index_t u = e[3 * j], v = e[3 * j + 1], w = e[3 * j + 2];
        index_t *au = q->adj + q->pos[u];
        au[1 + 2 * au[0]] = v;
        au[2 + 2 * au[0]] = w;
        au[0]++;
        index_t *av = q->adj + q->pos[v];
        av[1 + 2 * av[0]] = u;
        av[2 + 2 * av[0]] = w;
        av[0]++;
    }
    printf("[adj: %.2lf ms] ", time_pop());
    fflush(stdout);

    time_push();
    memcpy(q->kk, g->terminals, sizeof(index_t) * (size_t)k);
    printf("[term: %.2lf ms] ", time_pop());

    printf("done. [%.2lf ms] ", time_pop());
    printf("{peak: %.2lfGiB} ", memtrack_pop_gib());
    printf("{curr: %.2lfGiB}\n", memtrack_curr_gib());
    fflush(stdout);
    return q;
}

void query_free(stp_query *q)
{
    if (!q) return;
    tfree(q->kk);
    tfree(q->pos);
    tfree(q->adj);
    tfree(q);
}

/* --------------------------------------------------------------- groups */

/* Groups file: one group per line, "g v1 v2 ..." with 1-based vertex ids;
 * other lines are ignored. */
stp_groups *groups_load(FILE *in, index_t n)
{
    stp_groups *gr = tmalloc(sizeof(stp_groups));
    index_t cap_g = 16, cap_m = 256;
    gr->g = 0;
    gr->start = tmalloc(sizeof(index_t) * (size_t)(cap_g + 1));
    gr->members = tmalloc(sizeof(index_t) * (size_t)cap_m);
    gr->start[0] = 0;
    char *line = NULL;
    size_t len = 0;
    while (getline(&line, &len, in) > 0) {
        if (line[0] != 'g' || (line[1] != ' ' && line[1] != '\t')) continue;
        char *p = line + 1, *end;
        if (gr->g == cap_g) {
            index_t *s = tmalloc(sizeof(index_t) * (size_t)(2 * cap_g + 1));
            memcpy(s, gr->start, sizeof(index_t) * (size_t)(cap_g + 1));
            tfree(gr->start);
            gr->start = s;
            cap_g *= 2;
        }
        index_t cnt = gr->start[gr->g];
        for (;;) {
            long v = strtol(p, &end, 10);
            if (end == p) break;
            p = end;
            if (v < 1 || v > n) FAIL("group member %ld outside 1..%ld", v, n);
            if (cnt == cap_m) {
                index_t *mm = tmalloc(sizeof(index_t) * (size_t)(2 * cap_m));
                memcpy(mm, gr->members, sizeof(index_t) * (size_t)cap_m);
                tfree(gr->members);
                gr->members = mm;
                cap_m *= 2;
            }
            gr->members[cnt++] = v - 1;
        }
        if (cnt == gr->start[gr->g]) FAIL("empty group on line: %s", line);
        gr->start[++gr->g] = cnt;
    }
    free(line);
    if (gr->g == 0) FAIL("no groups in the groups file");
    return gr;
}

void groups_free(stp_groups *gr)
{
    if (!gr) return;
    tfree(gr->start);
    tfree(gr->members);
    tfree(gr);
}

/* Reduced instance: vertices n..n+g-1 are the virtual terminals; virtual
 * terminal i is joined to every member of group i by an edge of weight
 * M = 1 + sum of all edge weights.  Since M exceeds every tree of real
 * edges, an optimal tree uses exactly one virtual edge per group, so
 * OPT_group = OPT(reduced) - g*M. */
stp_graph *group_reduce(const stp_graph *g, const stp_groups *gr, index_t *M_out)
{
    index_t M = 1;
    for (index_t j = 0; j < g->m; j++) {
        index_t w = g->edges[3 * j + 2];
        if (w < 0 || M > EMV_INF - w) FAIL("edge weights overflow the reduction weight");
        M += w;
    }
    stp_graph *r = stp_alloc();
    r->n = g->n + gr->g;
    for (index_t j = 0; j < g->m; j++)
        stp_add_edge(r, g->edges[3 * j], g->edges[3 * j + 1], g->edges[3 * j + 2]);
    for (index_t i = 0; i < gr->g; i++)
        for (index_t p = gr->start[i]; p < gr->start[i + 1]; p++)
            stp_add_edge(r, g->n + i, gr->members[p], M);
    r->m = r->num_edges;
    r->k = gr->g;
    r->terminals = tmalloc(sizeof(index_t) * (size_t)gr->g);
    for (index_t i = 0; i < gr->g; i++) r->terminals[i] = g->n + i;
    r->num_terminals = gr->g;
    r->cost = -1;
    r->flags = SEC_GRAPH | SEC_TERMINALS;
    *M_out = M;
    return r;
}
