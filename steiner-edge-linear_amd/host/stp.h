/*
 * stp.h -- host side of emv-reader: tracked allocation, wall-clock stack,
 * DIMACS STP loading and the query build handed to the device library.
 */
#ifndef EMV_STP_H
#define EMV_STP_H

#include <stdint.h>
#include <stdio.h>

typedef long int index_t; /* 64-bit, as reader-el.c:64 */

#define EMV_INF ((index_t)0x7FFFFFFFFFFFFFFF)

/* Abort with the reference's message shape: "ERROR [file = F, line = L] fn: msg"
 * (reader-el.c:87-104); harnesses grep for "ERROR". */
#define FAIL(...) emv_fail(__FILE__, __LINE__, __func__, __VA_ARGS__)
void emv_fail(const char *file, int line, const char *func, const char *fmt, ...)
    __attribute__((noreturn, format(printf, 4, 5)));

/* Tracked host allocation: current total, balance and a peak stack. */
void *tmalloc(size_t size);
void tfree(void *p);
long tmalloc_balance(void);
void memtrack_push(void);
double memtrack_pop_gib(void);
double memtrack_curr_gib(void);

/* Wall-clock stack in milliseconds. */
void time_push(void);
double time_pop(void);
int time_depth(void);

/* Edge-list graph as read from STP (0-based ids). */
typedef struct {
    index_t n, m, k;
    index_t num_edges, num_terminals, edge_capacity;
    index_t *edges;     /* 3 per edge: u, v, w */
    index_t *terminals; /* k, file order */
    index_t cost;       /* "cost" line, -1 when absent */
    unsigned flags;
} stp_graph;

stp_graph *stp_alloc(void);
void stp_free(stp_graph *g);
void stp_add_edge(stp_graph *g, index_t u, index_t v, index_t w);
stp_graph *stp_load(FILE *in);

/* Query in the reference's steinerq layout (reader-el.c:696-704): pos[u]
 * indexes adj, adj[pos[u]] = degree, then (neighbour, weight) pairs. */
typedef struct {
    index_t n, m, k;
    index_t *kk, *pos, *adj;
} stp_query;

stp_query *query_build(const stp_graph *g);
void query_free(stp_query *q);

/* Group Steiner reduction: one virtual terminal per group joined to its
 * members with weight M = 1 + sum of all edge weights. */
typedef struct {
    index_t g;        /* number of groups */
    index_t *start;   /* g+1 offsets into members */
    index_t *members; /* 0-based vertex ids */
} stp_groups;

stp_groups *groups_load(FILE *in, index_t n);
void groups_free(stp_groups *gr);
stp_graph *group_reduce(const stp_graph *g, const stp_groups *gr, index_t *M_out);

#endif
