// ref_harness.cpp -- extern "C" shim over the reference library's own code,
// compiled from the reference sources where they lie (-I/root/reference) into
// oracle/_ref/libref_sssp.so. TEST INFRASTRUCTURE ONLY: used in this container
// to write golden vectors (tests/golden/make_golden.py) and, where the .so
// exists, to cross-check the oracle and the host generators.
//
// include/architectures.h unconditionally defines __USE_GPU__ (line 6), which
// makes include/node_data/node_data.h:11-14 pull in cuda_runtime_api.h. This
// file includes architectures.h first and then #undefs the macro; the header's
// #pragma once keeps it from being redefined, so every reference header below
// compiles exactly as written, CPU paths only.
#include "include/architectures.h"
#undef __USE_GPU__
#include "include/test/generation/graph_generation_API.h"
#include "include/storage/graph_storage_API.h"
#include "algorithm/sssp/sssp.h"
#include "algorithm/bfs/bfs.h"

#include <cstdint>
#include <cstring>
#include <cstdlib>
#include <vector>

template <typename W>
static void copy_out(Graph<int, W> &g, int32_t *src, int32_t *dst, W *w)
{
    GraphContainerData<int, W> d = g.get_graph_data();
    long long E = g.get_edges_count();
    std::memcpy(src, d.edges_src_ids, sizeof(int) * E);
    std::memcpy(dst, d.edges_dst_ids, sizeof(int) * E);
    std::memcpy(w, d.edges_weights, sizeof(W) * E);
}

template <typename W>
static void fill_graph(Graph<int, W> &g, const int32_t *src, const int32_t *dst, const W *w,
                       int32_t V, int64_t E)
{
    // the same sequence load_from_edges_list_bin_file uses
    // (graph_storage_API.hpp:134-156): resize, add_edge x E, add_vertex x V
    g.resize(V, E);
    for (int64_t i = 0; i < E; i++) g.add_edge(src[i], dst[i], w[i]);
    for (int32_t i = 0; i < V; i++) g.add_vertex(i, 0);
}

extern "C" {

// GraphGenerationAPI::R_MAT_parallel (graph_generation_API.hpp:404-484).
// With one OpenMP thread the per-thread seed is time(NULL)*0 == 0, the only
// deterministic setting of the reference generator.
int64_t ref_rmat_parallel(int32_t vertices, int32_t avg_deg, int32_t a, int32_t b,
                          int32_t c, int32_t d, int32_t *src, int32_t *dst, float *w)
{
    Graph<int, float> g(0, 0, EDGES_LIST, true);
    GraphGenerationAPI<int, float>::R_MAT_parallel(g, vertices, avg_deg, a, b, c, d, 1, true);
    long long E = g.get_edges_count();
    if (src) copy_out(g, src, dst, w);
    return E;
}

// GraphGenerationAPI::random_uniform (graph_generation_API.hpp:4-71), directed,
// not optimized, global rand() seeded with srand(seed) first.
int64_t ref_random_uniform_f32(int32_t vertices, int32_t avg_deg, uint32_t seed,
                               int32_t *src, int32_t *dst, float *w)
{
    srand(seed);
    Graph<int, float> g(0, 0, EDGES_LIST, true);
    GraphGenerationAPI<int, float>::random_uniform(g, vertices, avg_deg, true, false);
    long long E = g.get_edges_count();
    if (src) copy_out(g, src, dst, w);
    return E;
}

// SingleSourceShortestPaths::cpu_bellman_ford (sssp.hpp:3-71), float weights.
int ref_cpu_bellman_ford_f32(const int32_t *src, const int32_t *dst, const float *w,
                             int32_t V, int64_t E, int32_t source, int32_t threads,
                             float *dist)
{
    try {
        Graph<int, float> g(0, 0, EDGES_LIST, true);
        fill_graph(g, src, dst, w, V, E);
        SingleSourceShortestPaths<int, float> op(threads);
        op.cpu_bellman_ford(g, dist, source);
    } catch (const char *) { return 1; }
    return 0;
}

// The same for double weights (the reference's other instantiated weight
// type, bellman_ford.cu:150-156).
int ref_cpu_bellman_ford_f64(const int32_t *src, const int32_t *dst, const double *w,
                             int32_t V, int64_t E, int32_t source, int32_t threads,
                             double *dist)
{
    try {
        Graph<int, double> g(0, 0, EDGES_LIST, true);
        fill_graph(g, src, dst, w, V, E);
        SingleSourceShortestPaths<int, double> op(threads);
        op.cpu_bellman_ford(g, dist, source);
    } catch (const char *) { return 1; }
    return 0;
}

// The same for unsigned int weights: the reference loop as written, including
// its unguarded 0xFFFFFFFF + w (sssp.hpp:21,56). Used only to document that
// defect next to the guarded oracle.
int ref_cpu_bellman_ford_u32(const int32_t *src, const int32_t *dst, const uint32_t *w,
                             int32_t V, int64_t E, int32_t source, int32_t threads,
                             uint32_t *dist)
{
    try {
        Graph<int, unsigned int> g(0, 0, EDGES_LIST, true);
        fill_graph(g, src, dst, (const unsigned int *)w, V, E);
        SingleSourceShortestPaths<int, unsigned int> op(threads);
        op.cpu_bellman_ford(g, (unsigned int *)dist, source);
    } catch (const char *) { return 1; }
    return 0;
}

// GraphStorageAPI::save_to_edges_list_bin_file (graph_storage_API.hpp:40-65).
int ref_save_bin_f32(const char *path, const int32_t *src, const int32_t *dst, const float *w,
                     int32_t V, int64_t E)
{
    Graph<int, float> g(0, 0, EDGES_LIST, true);
    fill_graph(g, src, dst, w, V, E);
    return GraphStorageAPI<int, float>::save_to_edges_list_bin_file(g, path) ? 0 : 1;
}

// GraphStorageAPI::load_from_edges_list_bin_file (graph_storage_API.hpp:115-160).
// Call with src == NULL to get (V, E) only.
int ref_load_bin_f32(const char *path, int32_t *V_out, int64_t *E_out,
                     int32_t *src, int32_t *dst, float *w)
{
    try {
        Graph<int, float> g(0, 0, EDGES_LIST, true);
        if (!GraphStorageAPI<int, float>::load_from_edges_list_bin_file(g, path)) return 1;
        *V_out = g.get_vertices_count();
        *E_out = g.get_edges_count();
        if (src) copy_out(g, src, dst, w);
    } catch (const char *) { return 2; }
    return 0;
}

// BFS::cpu_sequential_bfs (algorithm/bfs/bfs.hpp:3-45): the reference's bool
// reachability result, returned as 0/1 bytes. The graph is filled as
// load_from_edges_list_bin_file does (weights unused by BFS: 1).
int ref_cpu_sequential_bfs(const int32_t *src, const int32_t *dst, int32_t V, int64_t E,
                           int32_t source, uint8_t *visited)
{
    try {
        Graph<int, float> g(0, 0, EDGES_LIST, true);
        std::vector<float> w((size_t)(E > 0 ? E : 1), 1.0f);
        fill_graph(g, src, dst, w.data(), V, E);
        bool *res = new bool[(size_t)V];
        BFS<int, float> op(1);
        op.cpu_sequential_bfs(g, res, source);
        for (int32_t i = 0; i < V; i++) visited[i] = res[i] ? 1 : 0;
        delete[] res;
    } catch (const char *) { return 1; }
    return 0;
}

}  // extern "C"
