/*
 * oracle.c -- CPU oracle for the Bellman-Ford SSSP path.
 *
 * TEST INFRASTRUCTURE ONLY. Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker
 * (or the timed CPU baseline). The GPU product never calls it.
 *
 * Contents (each function names the reference code it restates):
 *   orc_bf_f32 / orc_bf_f64   -- cpu_bellman_ford, algorithm/sssp/sssp.hpp:3-71
 *   orc_bf_u32                -- same loop with an INF-guarded add (the
 *                                reference wraps 0xFFFFFFFF + w, sssp.hpp:21,56;
 *                                deliberate, documented deviation)
 *   orc_dijkstra_f32/_u32     -- textbook binary-heap Dijkstra in the same
 *                                arithmetic (independent pin: for w >= 0 the
 *                                greatest fixpoint of d[v] <= d[u] (+) w is unique
 *                                and Dijkstra reaches it exactly)
 *   orc_certify_f32/_u32      -- shortest-path certificate: every edge satisfied,
 *                                d[s] == 0, and a tight parent tree rooted at s
 *                                for every reachable vertex
 *   orc_bfs                   -- BFS::cpu_sequential_bfs, algorithm/bfs/bfs.hpp:3-45
 *                                (queue over the CSR, reachability from the
 *                                source), plus the hop level of every vertex
 *
 *   orc_bf_f32_capped         -- orc_bf_f32 stopped after a pass cap (the
 *                                bounded CPU baseline of the bench)
 *
 * Pins (tests/test_oracle.py): the reference's own cpu_bellman_ford compiled
 * from /root/reference into oracle/_ref (golden vectors in tests/golden/ for
 * float and double weights, and a direct cross-check where oracle/_ref is
 * built), Dijkstra bit-equality, brute force over all simple paths on tiny
 * graphs (float, double and unsigned sums), closed forms (chains, stars,
 * disconnected parts), edge-permutation invariance; the capped loop equals
 * the full loop at a cap >= its pass count and bounds the fixpoint below it.
 *
 * Build: see oracle/Makefile (gcc -O3 -fopenmp, never -ffast-math: fp32 adds
 * must be single IEEE RNE operations for bit-exact parity).
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>
#include <omp.h>

#define U32_INF 0xFFFFFFFFu

/* ---------------------------------------------------------------------------
 * cpu_bellman_ford restatement (algorithm/sssp/sssp.hpp:3-71).
 * Parallel init of d to the type's max (sssp.hpp:27-32), d[source] = 0, then
 * repeat an in-place relaxation pass over the edge list with
 * `omp for schedule(guided)` until a pass records no change (sssp.hpp:35-62).
 * Returns the number of passes (the reference's `iterations`, sssp.hpp:50).
 * ------------------------------------------------------------------------- */
int orc_bf_f32(const int32_t *src, const int32_t *dst, const float *w,
               int32_t V, int64_t E, int32_t source, int32_t threads,
               float *dist)
{
    if (threads <= 0) threads = omp_get_max_threads();
    volatile int changes = 1;
    int iterations = 0;
    #pragma omp parallel num_threads(threads) shared(changes, iterations)
    {
        #pragma omp for schedule(static)
        for (int32_t i = 0; i < V; i++) dist[i] = FLT_MAX;
        #pragma omp single
        dist[source] = 0.0f;
        while (changes) {
            #pragma omp barrier
            #pragma omp single
            { iterations++; changes = 0; }
            #pragma omp for schedule(guided)
            for (int64_t e = 0; e < E; e++) {
                float ds = dist[src[e]];
                float dd = dist[dst[e]];
                float cand = ds + w[e];
                if (dd > cand) { dist[dst[e]] = cand; changes = 1; }
            }
            /* implicit barrier at the end of omp for */
        }
    }
    return iterations;
}

/* The same loop stopped after at most `max_passes` passes (bounded CPU
 * baseline samples); returns the passes run, *converged = 1 when the last
 * pass changed nothing. */
int orc_bf_f32_capped(const int32_t *src, const int32_t *dst, const float *w,
                      int32_t V, int64_t E, int32_t source, int32_t threads,
                      int32_t max_passes, float *dist, int32_t *converged)
{
    if (threads <= 0) threads = omp_get_max_threads();
    volatile int changes = 1;
    int iterations = 0;
    #pragma omp parallel num_threads(threads) shared(changes, iterations)
    {
        #pragma omp for schedule(static)
        for (int32_t i = 0; i < V; i++) dist[i] = FLT_MAX;
        #pragma omp single
        dist[source] = 0.0f;
        while (changes && iterations < max_passes) {
            #pragma omp barrier
            #pragma omp single
            { iterations++; changes = 0; }
            #pragma omp for schedule(guided)
            for (int64_t e = 0; e < E; e++) {
                float ds = dist[src[e]];
                float dd = dist[dst[e]];
                float cand = ds + w[e];
                if (dd > cand) { dist[dst[e]] = cand; changes = 1; }
            }
        }
    }
    *converged = changes ? 0 : 1;
    return iterations;
}

int orc_bf_f64(const int32_t *src, const int32_t *dst, const double *w,
               int32_t V, int64_t E, int32_t source, int32_t threads,
               double *dist)
{
    if (threads <= 0) threads = omp_get_max_threads();
    volatile int changes = 1;
    int iterations = 0;
    #pragma omp parallel num_threads(threads) shared(changes, iterations)
    {
        #pragma omp for schedule(static)
        for (int32_t i = 0; i < V; i++) dist[i] = DBL_MAX;
        #pragma omp single
        dist[source] = 0.0;
        while (changes) {
            #pragma omp barrier
            #pragma omp single
            { iterations++; changes = 0; }
            #pragma omp for schedule(guided)
            for (int64_t e = 0; e < E; e++) {
                double ds = dist[src[e]];
                double dd = dist[dst[e]];
                double cand = ds + w[e];
                if (dd > cand) { dist[dst[e]] = cand; changes = 1; }
            }
        }
    }
    return iterations;
}

/* Same loop for uint32 weights, with the add guarded: an unreached source
 * (d == 0xFFFFFFFF) relaxes nothing and sums saturate at 0xFFFFFFFF. The
 * reference computes `src_distance + weight` unguarded (sssp.hpp:56) and so
 * wraps; that defect is not reproduced. */
int orc_bf_u32(const int32_t *src, const int32_t *dst, const uint32_t *w,
               int32_t V, int64_t E, int32_t source, int32_t threads,
               uint32_t *dist)
{
    if (threads <= 0) threads = omp_get_max_threads();
    volatile int changes = 1;
    int iterations = 0;
    #pragma omp parallel num_threads(threads) shared(changes, iterations)
    {
        #pragma omp for schedule(static)
        for (int32_t i = 0; i < V; i++) dist[i] = U32_INF;
        #pragma omp single
        dist[source] = 0u;
        while (changes) {
            #pragma omp barrier
            #pragma omp single
            { iterations++; changes = 0; }
            #pragma omp for schedule(guided)
            for (int64_t e = 0; e < E; e++) {
                uint32_t ds = dist[src[e]];
                if (ds == U32_INF) continue;
                uint64_t c = (uint64_t)ds + w[e];
                uint32_t cand = c >= U32_INF ? U32_INF : (uint32_t)c;
                if (dist[dst[e]] > cand) { dist[dst[e]] = cand; changes = 1; }
            }
        }
    }
    return iterations;
}

/* ---------------------------------------------------------------------------
 * CSR helper (host, serial): out-edges grouped by src, original order kept.
 * ------------------------------------------------------------------------- */
static int64_t *build_row_ptr(const int32_t *src, int32_t V, int64_t E, int64_t **perm_out)
{
    int64_t *rp = (int64_t *)calloc((size_t)V + 1, sizeof(int64_t));
    int64_t *perm = (int64_t *)malloc((size_t)(E > 0 ? E : 1) * sizeof(int64_t));
    int64_t *cur = (int64_t *)malloc((size_t)V * sizeof(int64_t) + 8);
    if (!rp || !perm || !cur) { free(rp); free(perm); free(cur); return NULL; }
    for (int64_t e = 0; e < E; e++) rp[src[e] + 1]++;
    for (int32_t v = 0; v < V; v++) rp[v + 1] += rp[v];
    for (int32_t v = 0; v < V; v++) cur[v] = rp[v];
    for (int64_t e = 0; e < E; e++) perm[cur[src[e]]++] = e;
    free(cur);
    *perm_out = perm;
    return rp;
}

/* ---------------------------------------------------------------------------
 * Binary-heap Dijkstra (lazy deletion) in the same arithmetic as the BF loop.
 * ------------------------------------------------------------------------- */
typedef struct { uint64_t key; int32_t v; } hitem;

static void heap_push(hitem *h, int64_t *n, hitem it)
{
    int64_t i = (*n)++;
    while (i > 0) {
        int64_t p = (i - 1) >> 1;
        if (h[p].key <= it.key) break;
        h[i] = h[p]; i = p;
    }
    h[i] = it;
}

static hitem heap_pop(hitem *h, int64_t *n)
{
    hitem top = h[0];
    hitem last = h[--(*n)];
    int64_t i = 0;
    for (;;) {
        int64_t l = 2 * i + 1, r = l + 1, m = i;
        uint64_t mk = last.key;
        if (l < *n && h[l].key < mk) { m = l; mk = h[l].key; }
        if (r < *n && h[r].key < mk) { m = r; }
        if (m == i) break;
        h[i] = h[m]; i = m;
    }
    if (*n > 0) h[i] = last;
    return top;
}

/* f32 keys: non-negative floats order like their bit patterns. */
static inline uint32_t fbits(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
static inline float bitsf(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }

int orc_dijkstra_f32(const int32_t *src, const int32_t *dst, const float *w,
                     int32_t V, int64_t E, int32_t source, float *dist)
{
    int64_t *perm = NULL;
    int64_t *rp = build_row_ptr(src, V, E, &perm);
    if (!rp) return -1;
    hitem *h = (hitem *)malloc((size_t)(E + 2) * sizeof(hitem));
    if (!h) { free(rp); free(perm); return -1; }
    for (int32_t v = 0; v < V; v++) dist[v] = FLT_MAX;
    dist[source] = 0.0f;
    int64_t n = 0;
    heap_push(h, &n, (hitem){ fbits(0.0f), source });
    while (n > 0) {
        hitem it = heap_pop(h, &n);
        if (it.key != fbits(dist[it.v])) continue;
        float du = dist[it.v];
        for (int64_t k = rp[it.v]; k < rp[it.v + 1]; k++) {
            int64_t e = perm[k];
            float cand = du + w[e];
            if (dist[dst[e]] > cand) {
                dist[dst[e]] = cand;
                heap_push(h, &n, (hitem){ fbits(cand), dst[e] });
            }
        }
    }
    free(h); free(rp); free(perm);
    return 0;
}

int orc_dijkstra_u32(const int32_t *src, const int32_t *dst, const uint32_t *w,
                     int32_t V, int64_t E, int32_t source, uint32_t *dist)
{
    int64_t *perm = NULL;
    int64_t *rp = build_row_ptr(src, V, E, &perm);
    if (!rp) return -1;
    hitem *h = (hitem *)malloc((size_t)(E + 2) * sizeof(hitem));
    if (!h) { free(rp); free(perm); return -1; }
    for (int32_t v = 0; v < V; v++) dist[v] = U32_INF;
    dist[source] = 0u;
    int64_t n = 0;
    heap_push(h, &n, (hitem){ 0u, source });
    while (n > 0) {
        hitem it = heap_pop(h, &n);
        if (it.key != dist[it.v]) continue;
        uint64_t du = dist[it.v];
        for (int64_t k = rp[it.v]; k < rp[it.v + 1]; k++) {
            int64_t e = perm[k];
            uint64_t c = du + w[e];
            uint32_t cand = c >= U32_INF ? U32_INF : (uint32_t)c;
            if (dist[dst[e]] > cand) {
                dist[dst[e]] = cand;
                heap_push(h, &n, (hitem){ cand, dst[e] });
            }
        }
    }
    free(h); free(rp); free(perm);
    return 0;
}

/* ---------------------------------------------------------------------------
 * Certificate. counts[] (8 int64):
 *   [0] edges violated (d[v] > d[u] (+) w with d[u] finite)
 *   [1] d[source] != 0
 *   [2] reachable vertices (finite d) without a tight parent path to source
 *   [3] vertices with finite d that a plain BFS from source does not reach
 *   [4] vertices BFS reaches whose d is infinite
 *   [5] reachable count (finite d)
 *   [6] if parent != NULL: parent entries that are invalid (not a tight edge,
 *       or not forming a tree rooted at source)
 *   [7] tight-edge tree size (vertices reached through tight edges)
 * Returns 0 when the certificate holds (all of [0],[1],[2],[3],[4],[6] zero).
 * ------------------------------------------------------------------------- */
static int certify_generic(const int32_t *src, const int32_t *dst, const void *wv,
                           int32_t V, int64_t E, int32_t source, const void *dv,
                           const int32_t *parent, int is_u32, int64_t *counts)
{
    const float *wf = (const float *)wv; const uint32_t *wu = (const uint32_t *)wv;
    const float *df = (const float *)dv; const uint32_t *du = (const uint32_t *)dv;
    memset(counts, 0, 8 * sizeof(int64_t));
    #define FIN(v) (is_u32 ? (du[v] != U32_INF) : (df[v] != FLT_MAX))
    /* [0] edge condition */
    int64_t viol = 0;
    #pragma omp parallel for reduction(+:viol) schedule(static)
    for (int64_t e = 0; e < E; e++) {
        int32_t u = src[e], v = dst[e];
        if (!FIN(u)) continue;
        if (is_u32) {
            uint64_t c = (uint64_t)du[u] + wu[e];
            if ((uint64_t)du[v] > c) viol++;
        } else {
            float c = df[u] + wf[e];
            if (df[v] > c) viol++;
        }
    }
    counts[0] = viol;
    counts[1] = is_u32 ? (du[source] != 0u) : (df[source] != 0.0f || fbits(df[source]) != 0u);
    int64_t *perm = NULL;
    int64_t *rp = build_row_ptr(src, V, E, &perm);
    if (!rp) return -1;
    int32_t *q = (int32_t *)malloc((size_t)V * sizeof(int32_t) + 4);
    uint8_t *seen = (uint8_t *)calloc((size_t)V + 1, 1);
    uint8_t *tseen = (uint8_t *)calloc((size_t)V + 1, 1);
    if (!q || !seen || !tseen) { free(rp); free(perm); free(q); free(seen); free(tseen); return -1; }
    /* plain BFS reachability */
    int64_t head = 0, tail = 0;
    q[tail++] = source; seen[source] = 1;
    while (head < tail) {
        int32_t u = q[head++];
        for (int64_t k = rp[u]; k < rp[u + 1]; k++) {
            int32_t v = dst[perm[k]];
            if (!seen[v]) { seen[v] = 1; q[tail++] = v; }
        }
    }
    /* tight-edge BFS */
    head = 0; tail = 0;
    q[tail++] = source; tseen[source] = 1;
    while (head < tail) {
        int32_t u = q[head++];
        for (int64_t k = rp[u]; k < rp[u + 1]; k++) {
            int64_t e = perm[k];
            int32_t v = dst[e];
            if (tseen[v]) continue;
            int tight;
            if (is_u32) tight = ((uint64_t)du[u] + wu[e] == (uint64_t)du[v]);
            else tight = (df[u] + wf[e] == df[v]);
            if (tight && FIN(v)) { tseen[v] = 1; q[tail++] = v; }
        }
    }
    counts[7] = tail;
    for (int32_t v = 0; v < V; v++) {
        int fin = FIN(v);
        counts[5] += fin;
        if (fin && !tseen[v]) counts[2]++;
        if (fin && !seen[v]) counts[3]++;
        if (!fin && seen[v]) counts[4]++;
    }
    /* parent array check: for every reachable v != source, parent[v] = u with
     * an edge (u,v) that is tight, and following parents reaches source
     * without a cycle. Unreachable vertices must carry -1. */
    if (parent) {
        int64_t bad = 0;
        int8_t *state = (int8_t *)calloc((size_t)V + 1, 1); /* 0 new, 1 on path, 2 ok, 3 bad */
        int32_t *stack = q;
        for (int32_t v = 0; v < V; v++) {
            int fin = FIN(v);
            if (!fin) { if (parent[v] != -1) bad++; continue; }
            if (v == source) { if (parent[v] != -1 && parent[v] != source) bad++; continue; }
            int32_t p = parent[v];
            if (p < 0 || p >= V) { bad++; continue; }
            int ok = 0;
            for (int64_t k = rp[p]; k < rp[p + 1] && !ok; k++) {
                int64_t e = perm[k];
                if (dst[e] != v) continue;
                if (is_u32) ok = ((uint64_t)du[p] + wu[e] == (uint64_t)du[v]);
                else ok = (FIN(p) && df[p] + wf[e] == df[v]);
            }
            if (!ok) bad++;
        }
        /* acyclicity: walk each chain, marking states */
        for (int32_t v = 0; v < V && bad == 0; v++) {
            if (!FIN(v) || state[v]) continue;
            int64_t sp = 0;
            int32_t x = v;
            while (x != source && state[x] == 0) {
                state[x] = 1; stack[sp++] = x;
                x = parent[x];
                if (x < 0) break;
            }
            int result = (x == source || (x >= 0 && state[x] == 2)) ? 2 : 3;
            if (x >= 0 && state[x] == 1) result = 3; /* cycle */
            if (x >= 0 && state[x] == 3) result = 3;
            for (int64_t i = 0; i < sp; i++) state[stack[i]] = (int8_t)result;
            if (result == 3) bad++;
        }
        free(state);
        counts[6] = bad;
    }
    free(rp); free(perm); free(q); free(seen); free(tseen);
    #undef FIN
    return (counts[0] || counts[1] || counts[2] || counts[3] || counts[4] || counts[6]) ? 1 : 0;
}

int orc_certify_f32(const int32_t *src, const int32_t *dst, const float *w,
                    int32_t V, int64_t E, int32_t source, const float *dist,
                    const int32_t *parent, int64_t *counts)
{
    return certify_generic(src, dst, w, V, E, source, dist, parent, 0, counts);
}

int orc_certify_u32(const int32_t *src, const int32_t *dst, const uint32_t *w,
                    int32_t V, int64_t E, int32_t source, const uint32_t *dist,
                    const int32_t *parent, int64_t *counts)
{
    return certify_generic(src, dst, w, V, E, source, dist, parent, 1, counts);
}

/* ---------------------------------------------------------------------------
 * BFS (BFS::cpu_sequential_bfs, algorithm/bfs/bfs.hpp:3-45): FIFO queue over
 * the out-edges in CSR order, a vertex is marked when first seen. visited[v]
 * is the reference's bool result (1/0); level[v] the hop count of the first
 * discovery (0xFFFFFFFF when unreached), which a FIFO BFS makes the shortest
 * hop count. Returns the number of reached vertices, -1 on bad input.
 * ------------------------------------------------------------------------- */
int64_t orc_bfs(const int32_t *src, const int32_t *dst, int32_t V, int64_t E, int32_t source,
                uint8_t *visited, uint32_t *level)
{
    if (V < 1 || source < 0 || source >= V || E < 0) return -1;
    int64_t *perm = NULL;
    int64_t *rp = build_row_ptr(src, V, E, &perm);
    int32_t *q = (int32_t *)malloc((size_t)V * sizeof(int32_t));
    if (!rp || !q) { free(rp); free(perm); free(q); return -1; }
    for (int32_t v = 0; v < V; v++) { visited[v] = 0; if (level) level[v] = 0xFFFFFFFFu; }
    int64_t head = 0, tail = 0;
    q[tail++] = source;
    visited[source] = 1;
    if (level) level[source] = 0;
    while (head < tail) {
        const int32_t u = q[head++];
        for (int64_t i = rp[u]; i < rp[u + 1]; i++) {
            const int32_t t = dst[perm[i]];
            if (!visited[t]) {
                visited[t] = 1;
                if (level) level[t] = level[u] + 1;
                q[tail++] = t;
            }
        }
    }
    free(rp); free(perm); free(q);
    return tail;
}

int orc_max_threads(void) { return omp_get_max_threads(); }
double orc_wtime(void) { return omp_get_wtime(); }
