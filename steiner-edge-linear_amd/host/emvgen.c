/*
 * emvgen.c -- seeded synthetic Steiner instances (library + CLI).
 *
 * Families:
 *   regular  random connected d-regular simple graph from the configuration
 *            model: stubs shuffled and paired; "reject" resamples until the
 *            pairing is simple and connected, "repair" removes self-loops
 *            and multi-edges by random double-edge swaps (degree preserving)
 *            and resamples only if the result is disconnected.
 *   grid     R x C grid, id = r*C + c, 4-neighbour edges in row-major order
 *            (right edge, then down edge, per vertex).
 * Weights are uniform integers in [1, wmax]; terminals are distinct uniform
 * vertices in draw order; groups are sets of distinct uniform vertices.
 *
 * Randomness: SplitMix64 (Steele, Lea, Flood 2014) with Lemire's unbiased
 * bounded draw (Lemire 2019).  Independent streams for structure, weights,
 * terminals and groups are seeded from (seed, stream id), so the graph does
 * not change when k changes.
 *
 * CLI (writes lower-case DIMACS STP on stdout, readable by reader-el):
 *   emv-gen regular <n> <d> <wmax> <k> <seed> [reject|repair]
 *   emv-gen grid <rows> <cols> <wmax> <k> <seed>
 *   emv-gen groups <n> <g> <size> <seed>          (groups file, 1-based)
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define GEN_STREAM_GRAPH 0
#define GEN_STREAM_WEIGHT 1
#define GEN_STREAM_TERM 2
#define GEN_STREAM_GROUP 3

static uint64_t sm_next(uint64_t *s)
{
    uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

/* uniform in [0, range), range >= 1 */
static uint64_t sm_below(uint64_t *s, uint64_t range)
{
    uint64_t x = sm_next(s);
    __uint128_t m = (__uint128_t)x * range;
    uint64_t lo = (uint64_t)m;
    if (lo < range) {
        uint64_t thr = (0 - range) % range;
        while (lo < thr) {
            x = sm_next(s);
            m = (__uint128_t)x * range;
            lo = (uint64_t)m;
        }
    }
    return (uint64_t)(m >> 64);
}

static uint64_t stream_seed(uint64_t seed, int id)
{
    uint64_t s = seed * 0xD1B54A32D192ED03ull + (uint64_t)id * 0x8CB92BA72F3D8DD7ull;
    return sm_next(&s);
}

/* ------------------------------------------------------------ regular */

typedef struct {
    int64_t n;
    int d;
    int32_t *nb;  /* n*d neighbour slots */
    int32_t *cnt; /* used slots per vertex */
} slots;

static int64_t mult(const slots *a, int32_t u, int32_t v)
{
    int64_t c = 0;
    const int32_t *p = a->nb + (int64_t)u * a->d;
    for (int i = 0; i < a->cnt[u]; i++) c += (p[i] == v);
    return (u == v) ? c / 2 : c;
}

static void slot_add(slots *a, int32_t u, int32_t v)
{
    a->nb[(int64_t)u * a->d + a->cnt[u]++] = v;
    a->nb[(int64_t)v * a->d + a->cnt[v]++] = u;
}

static void slot_del1(slots *a, int32_t u, int32_t v)
{
    int32_t *p = a->nb + (int64_t)u * a->d;
    for (int i = 0; i < a->cnt[u]; i++)
        if (p[i] == v) {
            p[i] = p[--a->cnt[u]];
            return;
        }
}

static void slot_del(slots *a, int32_t u, int32_t v)
{
    slot_del1(a, u, v);
    slot_del1(a, v, u);
}

static int is_bad(const slots *a, int32_t u, int32_t v)
{
    return u == v || mult(a, u, v) > 1;
}

static int connected(const slots *a)
{
    int64_t n = a->n;
    char *seen = calloc((size_t)n, 1);
    int32_t *st = malloc(sizeof(int32_t) * (size_t)n);
    int64_t sp = 0, cnt = 1;
    seen[0] = 1;
    st[sp++] = 0;
    while (sp) {
        int32_t u = st[--sp];
        for (int i = 0; i < a->cnt[u]; i++) {
            int32_t v = a->nb[(int64_t)u * a->d + i];
            if (!seen[v]) {
                seen[v] = 1;
                cnt++;
                st[sp++] = v;
            }
        }
    }
    free(seen);
    free(st);
    return cnt == n;
}

/* edges: 3*(n*d/2) int64 (u, v, w), 0-based.  method 0 = reject, 1 = repair.
 * Returns the number of pairings sampled, or -1 on bad arguments. */
int64_t emvgen_regular(int64_t n, int d, int64_t wmax, uint64_t seed, int method,
                       int64_t *edges)
{
    if (n < 2 || d < 1 || d >= n || (n * d) % 2 || wmax < 1) return -1;
    int64_t m = n * d / 2;
    uint64_t sg = stream_seed(seed, GEN_STREAM_GRAPH);
    uint64_t sw = stream_seed(seed, GEN_STREAM_WEIGHT);
    int32_t *stub = malloc(sizeof(int32_t) * (size_t)(2 * m));
    int32_t *eu = malloc(sizeof(int32_t) * (size_t)m), *ev = malloc(sizeof(int32_t) * (size_t)m);
    slots a = {n, d, malloc(sizeof(int32_t) * (size_t)(n * d)), malloc(sizeof(int32_t) * (size_t)n)};
    int64_t tries = 0;
    for (;;) {
        tries++;
        for (int64_t i = 0; i < 2 * m; i++) stub[i] = (int32_t)(i / d);
        for (int64_t i = 2 * m - 1; i > 0; i--) {
            int64_t j = (int64_t)sm_below(&sg, (uint64_t)i + 1);
            int32_t t = stub[i];
            stub[i] = stub[j];
            stub[j] = t;
        }
        memset(a.cnt, 0, sizeof(int32_t) * (size_t)n);
        for (int64_t j = 0; j < m; j++) {
            eu[j] = stub[2 * j];
            ev[j] = stub[2 * j + 1];
            slot_add(&a, eu[j], ev[j]);
        }
        int64_t bad = 0;
        for (int64_t j = 0; j < m; j++) bad += is_bad(&a, eu[j], ev[j]);
        if (bad && method == 0) continue;
        /* repair passes: swap each bad edge with a random partner */
        int64_t guard = 0;
        while (bad && guard < 1000) {
            guard++;
            for (int64_t j = 0; j < m; j++) {
                while (is_bad(&a, eu[j], ev[j])) {
                    int64_t r = (int64_t)sm_below(&sg, (uint64_t)m);
                    if (r == j) continue;
                    int32_t x = eu[j], y = ev[j], p = eu[r], q = ev[r];
                    if (sm_next(&sg) & 1) {
                        int32_t t = p;
                        p = q;
                        q = t;
                    }
                    slot_del(&a, x, y);
                    slot_del(&a, p, q);
                    int ok = x != p && y != q && mult(&a, x, p) == 0 && mult(&a, y, q) == 0 &&
                             !((x == q && y == p) || (x == y && p == q));
                    if (ok) {
                        slot_add(&a, x, p);
                        slot_add(&a, y, q);
                        eu[j] = x;
                        ev[j] = p;
                        eu[r] = y;
                        ev[r] = q;
                    } else {
                        slot_add(&a, x, y);
                        slot_add(&a, p, q);
                    }
                }
            }
            bad = 0;
            for (int64_t j = 0; j < m; j++) bad += is_bad(&a, eu[j], ev[j]);
        }
        if (bad) continue;
        if (!connected(&a)) continue;
        break;
    }
    for (int64_t j = 0; j < m; j++) {
        edges[3 * j] = eu[j];
        edges[3 * j + 1] = ev[j];
        edges[3 * j + 2] = 1 + (int64_t)sm_below(&sw, (uint64_t)wmax);
    }
    free(stub);
    free(eu);
    free(ev);
    free(a.nb);
    free(a.cnt);
    return tries;
}

/* edges: 3*(R*(C-1) + (R-1)*C) int64. Returns the edge count. */
int64_t emvgen_grid(int64_t R, int64_t C, int64_t wmax, uint64_t seed, int64_t *edges)
{
    if (R < 1 || C < 1 || wmax < 1) return -1;
    uint64_t sw = stream_seed(seed, GEN_STREAM_WEIGHT);
    int64_t m = 0;
    for (int64_t r = 0; r < R; r++)
        for (int64_t c = 0; c < C; c++) {
            int64_t id = r * C + c;
            if (c + 1 < C) {
                edges[3 * m] = id;
                edges[3 * m + 1] = id + 1;
                edges[3 * m + 2] = 1 + (int64_t)sm_below(&sw, (uint64_t)wmax);
                m++;
            }
            if (r + 1 < R) {
                edges[3 * m] = id;
                edges[3 * m + 1] = id + C;
                edges[3 * m + 2] = 1 + (int64_t)sm_below(&sw, (uint64_t)wmax);
                m++;
            }
        }
    return m;
}

/* k distinct vertices of [0, n), uniform, in draw order; stream `stream`
 * lets callers draw several independent sets from one seed. */
int emvgen_distinct(int64_t n, int64_t k, uint64_t seed, int stream, int64_t *out)
{
    if (k < 0 || k > n) return -1;
    uint64_t s = stream_seed(seed, stream);
    for (int64_t i = 0; i < k; i++) {
        for (;;) {
            int64_t v = (int64_t)sm_below(&s, (uint64_t)n);
            int dup = 0;
            for (int64_t j = 0; j < i; j++) dup |= (out[j] == v);
            if (!dup) {
                out[i] = v;
                break;
            }
        }
    }
    return 0;
}

int emvgen_terminals(int64_t n, int64_t k, uint64_t seed, int64_t *kk)
{
    return emvgen_distinct(n, k, seed, GEN_STREAM_TERM, kk);
}

/* g groups of `size` distinct vertices each, drawn independently. */
int emvgen_groups(int64_t n, int64_t g, int64_t size, uint64_t seed, int64_t *members)
{
    for (int64_t i = 0; i < g; i++)
        if (emvgen_distinct(n, size, seed, GEN_STREAM_GROUP + 16 + (int)i, members + i * size))
            return -1;
    return 0;
}

/* Uniform draws for workload parameters: u64 stream values. */
uint64_t emvgen_u64(uint64_t seed, int stream, int64_t index)
{
    uint64_t s = stream_seed(seed, stream);
    uint64_t x = 0;
    for (int64_t i = 0; i <= index; i++) x = sm_next(&s);
    return x;
}

#ifdef EMVGEN_MAIN
static void write_stp(const char *name, int64_t n, int64_t m, const int64_t *e, int64_t k,
                      const int64_t *kk)
{
    printf("33d32945 stp file, stp format version 1.0\n\n");
    printf("section comment\nname    \"%s\"\ncreator \"emv-gen\"\nend\n\n", name);
    printf("section graph\nnodes %ld\nedges %ld\n", (long)n, (long)m);
    for (int64_t j = 0; j < m; j++)
        printf("e %ld %ld %ld\n", (long)e[3 * j] + 1, (long)e[3 * j + 1] + 1, (long)e[3 * j + 2]);
    printf("end\n\nsection terminals\nterminals %ld\n", (long)k);
    for (int64_t i = 0; i < k; i++) printf("t %ld\n", (long)kk[i] + 1);
    printf("end\n\neof\n");
}

int main(int argc, char **argv)
{
    if (argc >= 7 && !strcmp(argv[1], "regular")) {
        int64_t n = atoll(argv[2]), wmax = atoll(argv[4]), k = atoll(argv[5]);
        int d = atoi(argv[3]);
        uint64_t seed = strtoull(argv[6], NULL, 10);
        int method = (argc >= 8 && !strcmp(argv[7], "reject")) ? 0 : 1;
        int64_t m = n * d / 2;
        int64_t *e = malloc(sizeof(int64_t) * 3 * (size_t)m), *kk = malloc(sizeof(int64_t) * (size_t)k);
        if (emvgen_regular(n, d, wmax, seed, method, e) < 0 || emvgen_terminals(n, k, seed, kk)) {
            fprintf(stderr, "bad arguments\n");
            return 2;
        }
        write_stp("regular", n, m, e, k, kk);
        return 0;
    }
    if (argc >= 7 && !strcmp(argv[1], "grid")) {
        int64_t R = atoll(argv[2]), C = atoll(argv[3]), wmax = atoll(argv[4]), k = atoll(argv[5]);
        uint64_t seed = strtoull(argv[6], NULL, 10);
        int64_t mcap = 2 * R * C;
        int64_t *e = malloc(sizeof(int64_t) * 3 * (size_t)mcap), *kk = malloc(sizeof(int64_t) * (size_t)k);
        int64_t m = emvgen_grid(R, C, wmax, seed, e);
        if (m < 0 || emvgen_terminals(R * C, k, seed, kk)) {
            fprintf(stderr, "bad arguments\n");
            return 2;
        }
        write_stp("grid", R * C, m, e, k, kk);
        return 0;
    }
    if (argc >= 6 && !strcmp(argv[1], "groups")) {
        int64_t n = atoll(argv[2]), g = atoll(argv[3]), s = atoll(argv[4]);
        uint64_t seed = strtoull(argv[5], NULL, 10);
        int64_t *mem = malloc(sizeof(int64_t) * (size_t)(g * s));
        if (emvgen_groups(n, g, s, seed, mem)) {
            fprintf(stderr, "bad arguments\n");
            return 2;
        }
        for (int64_t i = 0; i < g; i++) {
            printf("g");
            for (int64_t j = 0; j < s; j++) printf(" %ld", (long)mem[i * s + j] + 1);
            printf("\n");
        }
        return 0;
    }
    fprintf(stderr,
            "usage: %s regular <n> <d> <wmax> <k> <seed> [reject|repair]\n"
            "       %s grid <rows> <cols> <wmax> <k> <seed>\n"
            "       %s groups <n> <g> <size> <seed>\n",
            argv[0], argv[0], argv[0]);
    return 2;
}
#endif
