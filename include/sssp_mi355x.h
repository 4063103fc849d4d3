/*
 * sssp_mi355x.h -- C ABI of the MI355X (gfx950) Bellman-Ford SSSP library
 * (libsssp_mi355x.so). Plain pointers and sizes only.
 *
 * It replaces the reference's device boundary
 *   template<W> void init_distances_wrapper(W*, int, int)
 *   template<W> void bellman_ford_wrapper(int*, int*, W*, int, long long, int, W*)
 *   (algorithm/sssp/gpu/bellman_ford.cuh:13-18, instantiated for float/double
 *    at bellman_ford.cu:150-156)
 * and the orchestration in SingleSourceShortestPaths::gpu_bellman_ford
 *   (algorithm/sssp/sssp.hpp:219-278),
 * plus the host services that feed it (generators, binary edge-list I/O).
 *
 * Conventions
 *  - Every function returns 0 (SSSP_OK) or an error class; the message of the
 *    last failure on the calling thread is returned by sssp_last_error().
 *    (The reference printed and threw `const char*`,
 *    include/common/cuda_error_hadling.h:4-24; the C++ adapter in
 *    parallel-task1_amd/host/ turns non-zero codes back into that throw.)
 *  - All calls are synchronous on return. A graph/pack/mlp handle is bound to
 *    one device chosen explicitly at creation and must not be used from two
 *    host threads at once.
 *  - Unreached vertices: FLT_MAX (f32), 0xFFFFFFFF (u32), DBL_MAX (f64), as
 *    the reference's init kernels (bellman_ford.cu:10-34).
 *  - Weights must be >= 0 and not NaN (checked; SSSP_ERR_ARG otherwise).
 */
#ifndef SSSP_MI355X_H
#define SSSP_MI355X_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------- */
#define SSSP_OK                0
#define SSSP_ERR_ARG           1   /* invalid argument */
#define SSSP_ERR_HIP           2   /* HIP runtime / kernel error */
#define SSSP_ERR_OOM           3   /* device or host allocation failed */
#define SSSP_ERR_INTERNAL      4
#define SSSP_ERR_IO            5   /* file open / short read / bad header */
#define SSSP_ERR_NOT_CONVERGED 6   /* iteration cap reached */

/* ---- weight types ------------------------------------------------------- */
#define SSSP_W_F32 0
#define SSSP_W_U32 1
#define SSSP_W_F64 2

/* ---- schedules (how each relaxation round is executed) ------------------- */
#define SSSP_SCHED_PLAIN     0  /* all-edge passes over the src-sorted edge list until a pass changes nothing (the plain GPU Bellman-Ford) */
#define SSSP_SCHED_HEURISTIC 1  /* per-round choice push / all-edge / pull by a hand-written rule */
#define SSSP_SCHED_LEARNED   2  /* per-round choice by the MLP set with sssp_graph_set_mlp (heuristic fallback) */
#define SSSP_SCHED_PUSH      3  /* every round: sparse frontier push */
#define SSSP_SCHED_ALL       4  /* every round: all-edge sweep of the active sources */
#define SSSP_SCHED_PULL      5  /* every round: pull into unsettled vertices */

/* ---- round modes reported in traces -------------------------------------- */
#define SSSP_MODE_PUSH  0
#define SSSP_MODE_ALL   1  /* all-edge round with per-edge frontier appends */
#define SSSP_MODE_PULL  2
#define SSSP_MODE_CHECK 3  /* trace only: the final fixpoint check (run dense) */
#define SSSP_MODE_DENSE 4  /* trace only: an all-edge round run dense (see dense_all_frac) */
#define SSSP_MODE_TAIL  5  /* trace only: a push round run as a tail round (see tail_edges) */

/* ---- graph creation flags ------------------------------------------------ */
#define SSSP_INPUT_DEVICE 1u   /* src/dst/w are device pointers on `device` */
#define SSSP_NO_CSC       2u   /* skip the CSC (pull) layout */
#define SSSP_NO_RELABEL   4u   /* keep vertex ids as the layouts' order (no degree-ordered slots) */
#define SSSP_RELABEL      8u   /* degree-ordered slots whatever the graph (default: from 2^23
                                  vertices on skewed graphs, see sssp_graph_vertex_order) */

typedef struct sssp_graph sssp_graph;
typedef struct sssp_pack  sssp_pack;
typedef struct sssp_mlp   sssp_mlp;

typedef struct sssp_stats {
    int32_t iterations;      /* relaxation rounds run, verification pass included */
    int32_t n_push;          /* rounds executed as sparse push */
    int32_t n_all;           /* rounds executed as all-edge sweeps (incl. plain passes) */
    int32_t n_pull;          /* rounds executed as pull */
    int32_t verify_passes;   /* full all-edge fixpoint checks (must change nothing) */
    int32_t mlp_rounds;      /* rounds whose mode the MLP chose */
    int32_t reached;         /* vertices with a finite distance */
    int32_t reserved;
    int64_t edges_relaxed;   /* edge relaxations evaluated (sum over rounds) */
    double  t_solve_ms;      /* device time from distance init to the certified fixpoint */
    double  gteps;           /* E / t_solve / 1e9 (reference convention, sssp.hpp:274-276) */
} sssp_stats;

/* ---- library / device ---------------------------------------------------- */
const char *sssp_last_error(void);
int sssp_version(void);
/* Device count and properties (replaces NodeData/init_library,
 * include/node_data/node_data.cpp:15-147; the reference left the LAST device
 * current, here the device is always explicit). */
int sssp_device_count(int32_t *count);
int sssp_device_info(int32_t device, char *name, int32_t name_len, int32_t *cu_count,
                     int64_t *hbm_bytes, int32_t *l2_bytes);
/* What NodeData collects per device (include/node_data/node_data.cpp:120-147:
 * cudaGetDeviceProperties + cudaMemGetInfo): the device's product name, free
 * and total global memory, the largest workgroup and the largest grid
 * (max_grid: 3 values). Any output may be NULL. The calling thread's current
 * device is the same before and after the call. */
int sssp_device_props(int32_t device, char *name, int32_t name_len, int64_t *free_bytes,
                      int64_t *total_bytes, int32_t *max_threads_per_block, int32_t *max_grid);

/* ---- graph on the device ------------------------------------------------- */
/* Uploads an edge list (SoA, as EdgesList / GraphContainerData hold it:
 * include/graph_data_structures/containers/edges_list.h:16-23,
 * graph_container_data.h:12-26) and builds, on the device: the edge list
 * sorted by src (stable, = CSR order of compressed_adjacency_list.hpp:176-203
 * reached through graph.hpp:316-361) with its row offsets, and the CSC
 * (in-edges, stable by dst) unless SSSP_NO_CSC. The library owns the device
 * memory; the input arrays may be freed after the call.
 * Replaces the per-call upload in gpu_bellman_ford (sssp.hpp:245-253). */
int sssp_graph_create(int32_t weight_type, const int32_t *src, const int32_t *dst,
                      const void *weights, int32_t vertices, int64_t edges,
                      int32_t device, uint32_t flags, sssp_graph **out);
/* Generates a directed R-MAT graph on the device -- bit-identical to
 * sssp_rmat_generate with the same arguments (the LCG behind rand_r is jumped
 * ahead per edge) -- and builds the graph from it without a host round trip.
 * src_out/dst_out/w_out (host, 2^scale*edge_factor each, may be NULL) receive
 * the generated edge list in generation order. */
int sssp_graph_create_rmat(int32_t scale, int32_t edge_factor, int32_t a, int32_t b, int32_t c,
                           int32_t d, uint32_t base_seed, int32_t virtual_threads, int32_t device,
                           uint32_t flags, int32_t *src_out, int32_t *dst_out, float *w_out,
                           sssp_graph **out);
/* Device ingest of a binary edge-list file (the reference's format: 1-byte
 * directed flag, int32 vertex count, int64 edge count, then packed records of
 * int32 src, int32 dst and a 4-byte weight). The records are read in chunks
 * into two pinned host buffers while the previous chunk is copied and split
 * into the edge arrays on the device; the layouts are then built there. No
 * host-side conversion or full host copy of the graph is made.
 * `directed` must match the file's flag (as the reference throws on a
 * mismatch); weight_type is SSSP_W_F32 or SSSP_W_U32 (files with 8-byte
 * weights are rejected: device graphs hold 4-byte weights).
 * Replaces GraphStorageAPI::load_from_edges_list_bin_file
 * (include/storage/graph_storage_API.hpp:113-157) followed by the upload of
 * gpu_bellman_ford (algorithm/sssp/sssp.hpp:245-253). */
int sssp_graph_create_from_bin(const char *path, int32_t directed, int32_t weight_type,
                               int32_t device, uint32_t flags, sssp_graph **out);
int sssp_graph_destroy(sssp_graph *g);
int sssp_graph_info(const sssp_graph *g, int32_t *vertices, int64_t *edges,
                    int32_t *weight_type, int64_t *device_bytes);
/* The order the device layouts use: slot_of_vertex[v] (host, V elements) =
 * the slot of vertex v. On a graph of at least 2^23 vertices whose largest
 * in- plus out-degree is at least 16 times the mean (R-MAT-like), or with
 * SSSP_RELABEL, slots rank the vertices by descending in- plus out-degree
 * (ties by descending out-degree, then id), so that the distances most edges
 * read share cache lines; otherwise (or with SSSP_NO_RELABEL) slot = vertex.
 * Every argument and result of this API stays in the caller's vertex ids. No
 * counterpart in the reference (its layouts keep the ids). */
int sssp_graph_vertex_order(const sssp_graph *g, int32_t *slot_of_vertex);
/* Copies the device layouts back (tests): row_ptr[V+1], sorted src/dst/w[E];
 * col_ptr[V+1], csc_src[E], csc_w[E] (csc arrays may be NULL). The layouts
 * index vertices by slot (see sssp_graph_vertex_order). */
int sssp_graph_download_layout(const sssp_graph *g, int64_t *row_ptr, int32_t *el_src,
                               int32_t *el_dst, void *el_w, int64_t *col_ptr,
                               int32_t *csc_src, void *csc_w);

/* ---- single-source solve ------------------------------------------------- */
/* Solves from `source` with the given schedule. dist_out (host, V elements of
 * the graph's weight type) may be NULL to keep the result on the device only.
 * Drop-in for gpu_bellman_ford's device part (sssp.hpp:256-262): the timed
 * window (stats->t_solve_ms) covers distance init up to the fixpoint and
 * excludes uploads and downloads, as the reference's t2..t3. */
int sssp_solve(sssp_graph *g, int32_t source, int32_t schedule, void *dist_out,
               sssp_stats *stats);
/* Downloads the distances of the last solve. */
int sssp_get_distances(sssp_graph *g, void *dist_out);
/* Per-round trace of the last solve: mode, frontier size, frontier out-edges,
 * device time (ms) of each round. Arrays of length `cap`; *rounds = count. */
int sssp_get_trace(sssp_graph *g, int32_t cap, int32_t *modes, int64_t *frontier,
                   int64_t *frontier_edges, float *round_ms, int32_t *rounds);
/* Certificate of the last solve, computed on the device: counts[0] violated
 * edges, [1] d[source] != 0, [2] reachable vertices without a tight parent
 * path, [5] reachable vertex count, [7] tight-tree size. parent_out (host, V,
 * may be NULL) receives a tight parent tree rooted at source (-1 for the
 * source and for unreachable vertices). Returns SSSP_OK and sets *valid. */
int sssp_certify(sssp_graph *g, int32_t source, int32_t *parent_out, int64_t *counts,
                 int32_t *valid);
/* Tuning knobs by name ("push_chunk", "all_chunk", "pull_chunk",
 * "advance_blocks", "verify", "max_rounds", "push_edge_frac", "pull_frac",
 * "delta", "hub_degree", "plain_below_edges", "rollout_policy", "collect_kind",
 * "collect_budget", "dense_all_frac", "dense_unreached", "dense_kernel",
 * "batch_capture", "decide_spin", "blk_tile_cap", "blk_tile_order",
 * "blk_range_bits", "plain_kernel", "tail_edges", "tail_group", "tail_blocks",
 * "push_bits_edges", "trace_events"; "cu_count",
 * "collect_rounds", "multi_passes", "blk_tiles", "blk_wide_chunks" and
 * "plain_blocked" read-only).
 * dense_all_frac (default 0.15, or 0.4 for a graph without hubs: no block of
 * 2^14 destinations with twice the mean in-edges) and dense_unreached
 * (default 0.5): an all-edge round whose frontier has at least
 * dense_all_frac * E out-edges, or that starts with at least dense_unreached
 * of a fixed vertex sample unreached, runs dense instead of appending per
 * edge; the fixpoint check always does. dense_kernel 0 (default) runs dense
 * rounds over the destination-blocked edge list (LDS minima per block, the
 * lowered vertices marked by the same kernel and queued when a push round
 * needs them); 1 runs the plain pass, then
 * compacts the lowered vertices against a snapshot into the next frontier.
 * rollout_policy and collect_kind steer sssp_collect_timings: the rounds
 * after the tried one follow the rule (0) or the attached network (1); the
 * collected solve is an SSSP (0) or a BFS (1). collect_budget (default 2^22)
 * bounds the rounds one collection runs in total (its cost grows with the
 * square of the solve's round count); past it the call returns
 * SSSP_ERR_NOT_CONVERGED. "collect_rounds" (read-only) = rounds the last
 * collection ran.
 * plain_below_edges (default 2^26): a single-source heuristic or learned
 * solve (or BFS) of a graph with fewer edges runs the plain pass loop, whose
 * passes then cost less than the rounds with their fixed per-round cost.
 * plain_kernel picks the plain pass: 0 (default) the destination-blocked pass
 * (block minima in LDS, as the dense rounds) once the blocked list has 64
 * tiles or more, else the XCD-partitioned edge-list pass; 1 always the
 * edge-list pass, 2 always the blocked pass. "plain_blocked" (read-only) = 1
 * when plain solves use the blocked pass.
 * tail_edges (default 2^23): once a heuristic or learned solve's frontier has
 * had more than tail_edges out-edges, a round decided as a push round over a
 * frontier with at most that many (far pile off) starts the tail rounds: push
 * rounds queued tail_group (default 4) at a time with no decision or host
 * round trip between them (three launches each: entry list, relaxation over
 * tail_blocks workgroups -- 0 = 2 per CU --, bookkeeping), until the frontier
 * is empty or has more out-edges than tail_edges again; 0 turns them off.
 * Traces show them as SSSP_MODE_TAIL, with sizes and times only under
 * trace_events (which queues them one at a time).
 * push_bits_edges (default 0 = off): a decided push round over a frontier
 * with at least this many out-edges (far pile off) marks the vertices it
 * lowers in a bitmap (one non-returning OR each) and builds the next queue
 * from it in one pass, instead of stamping and appending per lowered vertex.
 * delta > 0 turns on the far pile of push rounds: a frontier vertex with
 * d >= (min distance over the frontier and pile) + delta and out-degree >=
 * hub_degree waits in the pile instead of being relaxed; 0 (default) = off.
 * max_rounds (default 100000) is bounded to [1, 2^28]; a solve that needs more
 * rounds returns SSSP_ERR_NOT_CONVERGED.
 * The plain schedule queues its passes four at a time (the change flags are
 * kept and tested on the device; the host reads them after each four); with
 * trace_events = 1 each pass is launched alone between HIP events.
 * batch_capture (default 1): a batch of policy 3 runs each chunk of up to 256
 * sources as captured hipGraph launches -- init and a run of 32 passes (one
 * launch each, the pass number read from device memory, passes after the
 * fixpoint returning at once) and the transposed output -- replayed per call,
 * one host check per run of 32 passes; 0 queues the passes from the host, four
 * per check. "multi_passes" (read-only) = passes the last policy-3 batch ran.
 * blk_tile_cap / blk_tile_order rebuild the tile table of the dense rounds'
 * destination-blocked list: at most blk_tile_cap edge slots per tile (a
 * multiple of 256; default 2^17, doubled while more than 4096 tiles of the
 * cap fit in E) and the dispatch order 0 = smallest tiles first (default), 1 = by block,
 * 2 = largest first, 3 = by descending block. blk_range_bits > 0 (18-31)
 * also cuts every block's tiles where their sources cross multiples of
 * 2^blk_range_bits and dispatches the tiles range by range, so that the
 * source gathers of the workgroups in flight stay in one range; the default
 * is 0, or, for a graph of 2^22 vertices or more without hub blocks, the size
 * that makes 8 ranges. "blk_wide_chunks" (read-only) = chunks of 256 edges
 * whose sources are stored whole (a gap of 2^18 or more between two of them). decide_spin (default 1): the host polls the
 * mapped control block for each round's decision instead of synchronizing
 * the stream.
 * Distances do not depend on any knob. */
int sssp_set_param(sssp_graph *g, const char *name, double value);
int sssp_get_param(const sssp_graph *g, const char *name, double *value);

/* ---- BFS on the same engine ----------------------------------------------- */
/* Breadth-first search from `source` over the graph's edges (weights
 * ignored), run by the same round engine as the SSSP with every edge counting
 * one hop: the fixpoint of lev[v] <= lev[u] + 1 is the hop distance. Any
 * schedule may be used (push rounds are top-down, pull rounds bottom-up, the
 * rule or the network picks per round). visited_out (host, V bytes, may be
 * NULL) receives the reference's result, 1 for every vertex reachable from
 * source, else 0; levels_out (host, V, may be NULL) the hop levels
 * (0xFFFFFFFF unreached). sssp_certify then checks the levels (every edge
 * satisfies lev[v] <= lev[u] + 1, a tight parent tree reaches every visited
 * vertex). Replaces BFS::gpu_bfs (algorithm/bfs/bfs.hpp:213-253). */
int sssp_bfs(sssp_graph *g, int32_t source, int32_t schedule, uint8_t *visited_out,
             uint32_t *levels_out, sssp_stats *stats);
/* bfs_wrapper (algorithm/bfs/gpu/bfs.cu:44-85) over caller-owned device
 * arrays: d_visited (device, V bytes) = 1 for every vertex reachable from
 * source over the edges (d_src[i] -> d_dst[i]), after all-edge passes until
 * one marks nothing; *passes (may be NULL) = passes run. Ids outside
 * [0, vertices) return SSSP_ERR_ARG before any pass. */
int sssp_bfs_edges(const int32_t *d_src, const int32_t *d_dst, int32_t vertices, int64_t edges,
                   uint8_t *d_visited, int32_t source, int32_t *passes);

/* ---- batched solves (one persistent launch, replayed through a hipGraph) -- */
/* Many sources on one graph. dist_out: n*V elements (host) or NULL.
 * policy 0 = autotuned (untuned: 3 for n >= 32, else 2), 1 = one workgroup
 * per source, 2 = the sources solved one after another on the full chip,
 * 3 = vertex-major all-edge passes: the distances of up to 256 sources are
 * held per vertex as one row, each pass reads every edge once for all of
 * them; the layout it needs (16 bytes per edge) is built on its first use,
 * and 256 * V distances are kept on the device. stats->iterations = the
 * policy run. */
int sssp_solve_batch(sssp_graph *g, const int32_t *sources, int32_t n, int32_t policy,
                     void *dist_out, sssp_stats *stats);
/* Many independent graphs: graph k has vertex_counts[k] vertices and edges
 * [edge_offsets[k], edge_offsets[k+1]) of the concatenated arrays, with ids
 * local to the graph. f32 weights. */
int sssp_pack_create(const int32_t *src, const int32_t *dst, const float *weights,
                     const int64_t *edge_offsets, const int32_t *vertex_counts,
                     int32_t n_graphs, int32_t device, sssp_pack **out);
int sssp_pack_destroy(sssp_pack *p);
/* sources[k] is local to graph k. dist_out: sum(vertex_counts) floats or NULL. */
int sssp_pack_solve(sssp_pack *p, const int32_t *sources, int32_t policy, float *dist_out,
                    sssp_stats *stats);
/* Autotuner: times candidate (policy, block size, chunk width, frontier
 * threshold) settings on the given batch and keeps the fastest; writes the
 * table to `tuning_file` (text, may be NULL). */
int sssp_autotune_batch(sssp_graph *g, const int32_t *sources, int32_t n,
                        const char *tuning_file, double *best_ms);
int sssp_autotune_pack(sssp_pack *p, const int32_t *sources, const char *tuning_file,
                       double *best_ms);

/* ---- reference-shaped wrappers (caller-owned device pointers) ------------ */
/* init_distances_wrapper<W> (bellman_ford.cu:38-56): d[i] = max, d[source] = 0 */
int sssp_init_distances_f32(float *d_dist, int32_t vertices, int32_t source);
int sssp_init_distances_f64(double *d_dist, int32_t vertices, int32_t source);
int sssp_init_distances_u32(uint32_t *d_dist, int32_t vertices, int32_t source);
/* bellman_ford_wrapper<W> (bellman_ford.cu:103-144): all-edge passes over the
 * caller's edge list, in its order, until a pass changes nothing (at most
 * `vertices` passes). Convergence is tested on the device; *passes (may be
 * NULL) receives the pass count. The edge arrays are checked on the device
 * first: a vertex id outside [0, vertices) or a negative or NaN weight returns
 * SSSP_ERR_ARG before any pass (the distances are compared as unsigned bit
 * patterns, which order non-negative values only). */
int sssp_bellman_ford_f32(const int32_t *d_src, const int32_t *d_dst, const float *d_w,
                          int32_t vertices, int64_t edges, int32_t source, float *d_dist,
                          int32_t *passes);
int sssp_bellman_ford_f64(const int32_t *d_src, const int32_t *d_dst, const double *d_w,
                          int32_t vertices, int64_t edges, int32_t source, double *d_dist,
                          int32_t *passes);
int sssp_bellman_ford_u32(const int32_t *d_src, const int32_t *d_dst, const uint32_t *d_w,
                          int32_t vertices, int64_t edges, int32_t source, uint32_t *d_dist,
                          int32_t *passes);
/* Allocation-free sssp_bellman_ford_f32: no counterpart in the reference. For
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc). Same arguments,
 * checks, passes and results as sssp_bellman_ford_f32, on the current device;
 * it calls no hipMalloc/hipFree. Its scratch (64 change flags and 2 validation
 * counters, 272 of 512 bytes) is a file-scope device array that each device
 * keeps until the process ends; the scratch does not grow with the graph, so
 * every size sssp_bellman_ford_f32 accepts is accepted. Calls on one device
 * share that array: calls that overlap on one device race. */
int sssp_bellman_ford_noalloc_f32(const int32_t *d_src, const int32_t *d_dst, const float *d_w,
                                  int32_t vertices, int64_t edges, int32_t source, float *d_dist,
                                  int32_t *passes);

/* ---- host services ------------------------------------------------------- */
/* GraphGenerationAPI::R_MAT_parallel (include/test/generation/
 * graph_generation_API.hpp:404-484), directed, with the per-thread seed
 * time(NULL)*tid (:430) replaced by base_seed*tid and the OpenMP thread count
 * replaced by a fixed `virtual_threads` (contiguous static chunks), so the
 * edge sequence does not depend on the host's core count. virtual_threads=1
 * reproduces the reference run with one thread exactly (seed 0). Arrays have
 * 2^scale * edge_factor elements. */
int sssp_rmat_generate(int32_t scale, int32_t edge_factor, int32_t a, int32_t b, int32_t c,
                       int32_t d, uint32_t base_seed, int32_t virtual_threads,
                       int32_t *src, int32_t *dst, float *weights);
/* GraphGenerationAPI::random_uniform (graph_generation_API.hpp:4-71), directed,
 * not optimized, with srand(seed): V vertex-value draws, then per edge
 * from = rand()%V, to = rand()%V and one weight draw.
 * weight_mode 0: f32 (rand()%10000)/10000 (reference rand_uniform_val<float>);
 * 1: u32 1 + rand()%255 (integer-weight configuration, a documented deviation);
 * 2: u32 rand()%10000 (reference rand_uniform_val<unsigned>). */
int sssp_er_generate(int32_t vertices, int32_t avg_degree, uint32_t seed, int32_t weight_mode,
                     int32_t *src, int32_t *dst, void *weights);
/* Multi-graph RMAT set: graph k uses base seed (seed0 + k) with one virtual
 * thread; its scale is scales[k]. Arrays are concatenated; edge_offsets has
 * n+1 entries (computed by the caller from the scales). */
int sssp_rmat_generate_many(int32_t n, const int32_t *scales, int32_t edge_factor, int32_t a,
                            int32_t b, int32_t c, int32_t d, uint32_t seed0,
                            const int64_t *edge_offsets, int32_t *src, int32_t *dst,
                            float *weights);
/* Binary edge-list file (GraphStorageAPI, include/storage/graph_storage_API.hpp
 * :40-65 save, :115-160 load): bool directed (1 byte), int32 V, int64 E, then
 * E x {int32 src, int32 dst, weight (weight_bytes)}, little-endian, packed. */
int sssp_save_bin(const char *path, int32_t directed, const int32_t *src, const int32_t *dst,
                  const void *weights, int32_t weight_bytes, int32_t vertices, int64_t edges);
int sssp_load_bin_header(const char *path, int32_t *directed, int32_t *vertices, int64_t *edges);
int sssp_load_bin(const char *path, int32_t weight_bytes, int32_t *src, int32_t *dst,
                  void *weights, int64_t edges_capacity);
/* `n` distinct sources drawn uniformly among vertices with out-degree > 0,
 * std::mt19937_64(seed), v = rng() % V, redrawn until acceptable. */
int sssp_pick_sources(const int32_t *src, int64_t edges, int32_t vertices, uint64_t seed,
                      int32_t n, int32_t *out);

/* ---- learned schedule: MLP 5 -> 64 -> 64 -> (3 modes + chunk buckets) ---- */
#define SSSP_MLP_IN      5
#define SSSP_MLP_HID     64
#define SSSP_MLP_MODES   3
#define SSSP_MLP_CHUNKS  4
#define SSSP_MLP_OUT     (SSSP_MLP_MODES + SSSP_MLP_CHUNKS)
/* parameter count: 5*64+64 + 64*64+64 + 64*7+7 */
#define SSSP_MLP_NPARAM  (SSSP_MLP_IN * SSSP_MLP_HID + SSSP_MLP_HID + SSSP_MLP_HID * SSSP_MLP_HID + SSSP_MLP_HID + SSSP_MLP_HID * SSSP_MLP_OUT + SSSP_MLP_OUT)

int sssp_mlp_create(int32_t device, uint64_t seed, sssp_mlp **out);
int sssp_mlp_destroy(sssp_mlp *m);
int sssp_mlp_get_params(const sssp_mlp *m, float *params);   /* SSSP_MLP_NPARAM floats */
int sssp_mlp_set_params(sssp_mlp *m, const float *params);
/* Forward pass on the device for n rows of features (host in, host out). */
int sssp_mlp_forward(sssp_mlp *m, const float *x, int32_t n, float *logits);
/* Loss and gradient on the device for one batch: loss = mean over rows of
 * CE(mode logits, y_mode) + CE(chunk logits, y_chunk). */
int sssp_mlp_loss_grad(sssp_mlp *m, const float *x, const int32_t *y_mode,
                       const int32_t *y_chunk, int32_t n, float *loss, float *grad);
/* Adam training on the device; losses (epochs entries, may be NULL). */
int sssp_mlp_train(sssp_mlp *m, const float *x, const int32_t *y_mode, const int32_t *y_chunk,
                   int32_t n, int32_t epochs, int32_t batch, float lr, uint64_t seed,
                   float *losses);
/* Attach the network to a graph for SSSP_SCHED_LEARNED SSSP solves (NULL
 * detaches). */
int sssp_graph_set_mlp(sssp_graph *g, sssp_mlp *m);
/* The same for sssp_bfs with SSSP_SCHED_LEARNED; without one, a learned BFS
 * runs the hand-written rule. */
int sssp_graph_set_bfs_mlp(sssp_graph *g, sssp_mlp *m);
/* Timing collection for training (rollouts): solves from `source` (a BFS when
 * the graph's "collect_kind" is 1); before every round the state is saved and
 * each of the 3 modes x SSSP_MLP_CHUNKS chunk widths is run for that round,
 * the rest of the solve following the rule ("rollout_policy" 0) or the
 * attached network (1); times_ms is the best of two device timings of each
 * remainder. The solve then continues with the best choice. Writes up to
 * `cap` rows: features[5], times_ms[3*SSSP_MLP_CHUNKS]. */
int sssp_collect_timings(sssp_graph *g, int32_t source, int32_t cap, float *features,
                         float *times_ms, int32_t *rows);
/* Feature vector definition used by both training and inference. */
int sssp_features(double frontier_vertices, double frontier_edges, double unsettled_frac,
                  double round_index, double mean_degree, double vertices, double edges,
                  float *out5);

/* ---- test hooks ---------------------------------------------------------- */
/* Makes the next device allocation fail with SSSP_ERR_OOM. */
int sssp_debug_fail_next_alloc(void);

#ifdef __cplusplus
}
#endif

#endif /* SSSP_MI355X_H */
