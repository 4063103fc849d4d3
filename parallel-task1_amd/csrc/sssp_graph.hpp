// sssp_graph.hpp -- the device-resident graph object and the kernels' shared
// launch helpers. Internal to the library.
#pragma once

#include "sssp_internal.hpp"

namespace sssp {

constexpr int kBlock = 256;          // threads per workgroup (4 waves)
constexpr int kMaxRounds = 1 << 16;  // hard cap on relaxation rounds per solve
constexpr int kChunkCount = SSSP_MLP_CHUNKS;
// chunk width bucket -> edges per lane per tile (tile = 256 * K edges)
constexpr int kChunkK[kChunkCount] = {1, 2, 4, 8};
// Frontier stamp of a vertex without out-edges: above every round tick (ticks
// restart below 2^30), so such a vertex is never appended to a frontier.
constexpr int32_t kSinkStamp = 0x7FFFFFFF;
// Pile state of a vertex: kPileIn while it waits in the far pile, else the
// tick of the round that last moved it out (0 = never).
constexpr int32_t kPileIn = -1;

// Per-round counters written by the relaxation kernels. Two slots indexed by
// round parity: round r writes slot (r+1)&1 (the frontier it builds) while
// the host reads slot r&1 (the frontier it consumes).
struct alignas(64) RoundSlot {
    int32_t qlen;          // vertices appended (out-degree > 0)
    uint32_t minw;         // min distance written this round (bit pattern)
    int64_t qedges;        // sum of out-degrees of appended vertices
    int64_t relaxed;       // edge relaxations evaluated
    int32_t lowered;       // relaxations whose candidate beat the distance read
    // filled by the round's sample (k_round_open) for the consuming round:
    int32_t sample_unreached;   // sampled vertices with d = INF
    int32_t sample_unsettled;   // sampled vertices with d > minF
    int32_t sample_n;           // sample size
    int64_t sample_unsettled_in;  // sum of in-degrees of the unsettled sample
};

struct Ctrl {
    RoundSlot slot[2];
    int32_t changed[64];   // plain passes: changed[p & 63] = 1 when pass p lowered a distance
    int64_t n_compact;     // lb scan: entries
    int64_t lb_total;      // lb scan: total edges
    int64_t cert[8];       // certificate counters
    int64_t n_reached;     // vertices with a finite distance after a solve
    int32_t err;           // device-side error flag
    // per-round decision written by the decision kernel
    int32_t dec_mode;
    int32_t dec_chunk;
    int32_t dec_mlp;       // 1 when the network made the choice
    int32_t dec_heur_mode;
    float feat[SSSP_MLP_IN];
    float logits[SSSP_MLP_OUT];
    // far pile of push rounds (double-buffered): frontier vertices whose
    // relaxation is deferred until their distance is near the frontier minimum
    int32_t pile_n[2];
    uint32_t pile_min[2];     // min distance over the pile when it was built
    int64_t pile_edges[2];    // out-edges of the pile's vertices
    int32_t lb_n;             // queue entries plus pile entries moved onto the queue
    int32_t dec_seq;          // host mirror: written last by the decision kernel (round
                              // sequence number), after the fields above
    int32_t open_done;        // workgroups of k_round_open done sampling (the last one decides)
    int32_t tail_stop;        // first round of the current run of tail rounds that did not
                              // run (-1: all so far ran)
    int32_t tail_rounds;      // tail rounds of the current run that ran
    int64_t tail_relaxed;     // edge relaxations evaluated by them
    unsigned long long tail_list;  // entries << 32 | edges listed by the tail round in flight
    int32_t mat_n;            // vertices written by k_frontier_from_bits
};

// Destination-blocked edge list: edges sorted by (dst >> kBlkBits, src), each
// block's run padded to whole chunks of kBlkChunk slots with copies of its
// last edge (relaxing an edge twice changes nothing). An edge is 8 bytes: a
// 32-bit word holding its destination's offset in the block (low kBlkBits
// bits) and its source as the delta to the previous edge of the chunk, and
// its weight; each chunk keeps its first source whole. A tile is a run of
// at most kBlkTileEdges slots of one block, whole chunks (the cap doubles
// while the graph holds more than kBlkTileCount tiles' worth of edges);
// `split` = the block has more than one tile.
constexpr int kBlkBits = 14;
constexpr int kBlkThreads = 1024;  // threads of a blocked-pass workgroup
static_assert(kBlkBits <= 16, "block offsets fit the low half of a packed edge");
constexpr int kBlkDeltaBits = 32 - kBlkBits;  // source delta bits of a packed edge
constexpr int64_t kBlkChunk = 256;  // edge slots per chunk: 4 per lane of one wave
constexpr int64_t kBlkTileEdges = (int64_t)1 << 17;
constexpr int64_t kBlkTileCount = 4096;
struct BlkTile {
    int32_t blk;
    int32_t split;
    int64_t e0, e1;
};
// Source-range cuts of the blocked list's block runs (host copy): off[b *
// (n_fine + 1) + k] = first slot of block b whose source is at least
// k << fine_bits, aligned down to its chunk (k = n_fine: the run's end).
// Tiles cut at every stride-th cut (stride 0: no cuts).
constexpr int kBlkRangeFineBits = 18;
constexpr int64_t kBlkRangeMaxEntries = (int64_t)1 << 22;
constexpr int kBlkRangeMinVertexBits = 22;  // source ranges by default from 2^22 vertices
struct BlkRanges {
    std::vector<int64_t> off;
    int32_t fine_bits = 0;
    int32_t n_fine = 0;
    int32_t stride = 0;
};

// Multi-source batches on one graph: edges as 16-byte records sorted by
// (dst >> kMBits, src) -- 16-vertex destination blocks, so a 256-source block
// of distances is 16 KB of LDS; tiles of at most kMTileEdges edges of one
// block.
constexpr int kMBits = 4;
constexpr int kMultiSlots = 4;  // captured chunk launches kept per graph
constexpr int kMultiKey = 10;
constexpr int64_t kMTileEdges = 4096;
struct MRec {
    int32_t src;
    uint32_t w;
    int32_t loc;  // dst - block base
    int32_t pad;
};

struct TraceRow {
    int32_t mode;
    int32_t chunk;
    int64_t frontier;
    int64_t frontier_edges;
    float ms;
    float unsettled;
};

// mlp.hip
const float *mlp_device_params(sssp_mlp *m);

}  // namespace sssp

struct sssp_graph {
    int device = 0;
    int wtype = SSSP_W_F32;
    int32_t V = 0;
    int64_t E = 0;
    hipStream_t stream = nullptr;
    int cu_count = 256;

    // src-sorted edge list == CSR (row_ptr, el_dst, el_w) plus the explicit src
    int64_t *row_ptr = nullptr;
    int32_t *el_src = nullptr;
    int32_t *el_dst = nullptr;
    uint32_t *el_w = nullptr;
    // XCD-partitioned edge list: 8 groups by hashed destination block, each
    // src-sorted and padded to a multiple of 4 edges; group x = edges
    // [xgoff[x], xgoff[x+1])
    int32_t *x_src = nullptr;
    int32_t *x_dst = nullptr;
    uint32_t *x_w = nullptr;
    int64_t *xgoff = nullptr;
    std::vector<int64_t> xgoff_host;
    // destination-blocked edge list (dense all-edge rounds and the blocked
    // plain pass), in chunks of kBlkChunk edge slots: b_pk[e] = source delta
    // (to the previous edge of the chunk) << kBlkBits | (destination - block
    // base); b_cbase[c] = source of chunk c's first edge, or -(1 + i) for a
    // chunk whose sources are kept whole in b_wide[i * kBlkChunk ...] (a gap
    // of 2^kBlkDeltaBits or more between two of its sources)
    uint32_t *b_pk = nullptr;
    int32_t *b_cbase = nullptr;
    int32_t *b_wide = nullptr;
    int64_t b_nwide = 0;  // chunks with whole sources
    uint32_t *b_w = nullptr;
    sssp::BlkTile *b_tile = nullptr;
    int32_t n_btile = 0;
    uint16_t *b_app = nullptr;  // dense rounds' append masks: kBlkThreads per tile
    bool frontier_bits = false;  // the coming round's frontier is in b_app, not yet queued
    std::vector<int64_t> b_off_host;  // padded block offsets (tiles are rebuilt from them)
    int64_t blk_tile_cap = 0;         // edges per tile at most
    int blk_tile_order = 0;           // dispatch order of the tiles (make_blk_tiles)
    sssp::BlkRanges b_ranges;         // source-range cuts of the block runs (may be empty)
    int blk_range_bits = 0;           // tiles cut by source ranges of 2^bits vertices (0: no)
    int64_t b_edges = 0;  // edges in the blocked list, padding included
    // multi-source batches (built on first use): records, tiles, and the
    // vertex-major distances of up to 256 sources (D[v * n + s])
    sssp::MRec *m_rec = nullptr;
    sssp::BlkTile *m_tile = nullptr;
    int32_t n_mtile = 0;
    uint32_t *mdist = nullptr;
    int64_t mdist_elems = 0;
    // the batch's sources on the device, and the captured launches of its
    // chunks (hipGraphExec_t), keyed on chunk shape and buffers
    int32_t *msrc = nullptr;
    int32_t msrc_cap = 0;
    int32_t *mflags = nullptr;  // 64 change flags of the passes, then the pass base
    void *mexec[sssp::kMultiSlots][2] = {};  // first run, next runs
    int64_t mkey[sssp::kMultiSlots][sssp::kMultiKey] = {};
    int mnext = 0;
    int batch_capture = 1;     // policy 3: 1 = captured launches (hipGraph) per chunk
    int32_t multi_passes = 0;  // passes the last policy-3 batch ran (over its chunks)
    // CSC: in-edges grouped by dst
    int64_t *col_ptr = nullptr;
    int32_t *csc_src = nullptr;
    uint32_t *csc_w = nullptr;
    bool has_csc = false;
    // static entry lists of vertices with out- (in-) degree > 0; *_scan has
    // n+1 entries = their row starts followed by E
    int32_t *nzo_v = nullptr;
    int64_t *nzo_scan = nullptr;
    int32_t n_nzo = 0;
    int32_t *nzi_v = nullptr;
    int64_t *nzi_scan = nullptr;
    int32_t n_nzi = 0;
    // tile -> first entry tables of the static lists, one per chunk bucket
    int32_t *nzo_tile[sssp::kChunkCount] = {};
    int32_t *nzi_tile[sssp::kChunkCount] = {};
    // vertices sampled for the unsettled-fraction estimate
    int32_t *sample_v = nullptr;
    int32_t *sample_in = nullptr;  // in-degree of each sample vertex (with the CSC)
    int32_t n_sample = 0;

    // vertex order: the layouts and every per-vertex array index vertices by
    // slot, slots in descending order of in- plus out-degree (ties by vertex
    // id), so that the distances of the vertices most edges read share cache
    // lines. vperm[v] = slot of caller vertex v, vinv[slot] = caller vertex
    // (both null, and slot = vertex, when the graph is not relabelled)
    int relabel_mode = 0;  // 0: by size and skew, 1: always, -1: never
    bool relabel = false;  // the layouts use degree-ordered slots
    int32_t *vperm = nullptr;
    int32_t *vinv = nullptr;
    std::vector<int32_t> vperm_host;
    // distances (or BFS levels) of the last solve in caller order: a buffer
    // of its own when relabelled, else the same memory as dist
    uint32_t *dout = nullptr;

    // solve state
    uint32_t *dist = nullptr;
    int32_t *stamp = nullptr;
    int32_t *queue[2] = {nullptr, nullptr};
    int32_t *lb_v = nullptr;      // compacted frontier (sparse push)
    int64_t *lb_start = nullptr;  // row start per entry
    int64_t *lb_scan = nullptr;   // exclusive degree scan, n+1
    int32_t *lb_tile = nullptr;   // tile -> first entry
    int64_t *blk_a = nullptr;     // block partials for scans
    int64_t *blk_b = nullptr;
    int64_t n_blk = 0;
    int32_t *parent = nullptr;    // certificate tight tree
    int32_t *pile[2] = {nullptr, nullptr};  // far pile (push rounds), double-buffered
    int32_t *farst = nullptr;     // per-vertex pile state (kPileIn / tick moved out)
    uint32_t *snap = nullptr;     // distances before a dense all-edge round
    int32_t snap_round = -1;      // snap equals dist as dense round snap_round left it (-1: stale)
    int pile_cur = 0;             // pile buffer the coming round reads
    bool pile_dirty = false;      // a solve stopped with vertices still piled
    sssp::Ctrl *ctrl = nullptr;
    sssp::Ctrl *ctrl_host = nullptr;  // pinned mirror (coherent, mapped)
    sssp::Ctrl *ctrl_host_dev = nullptr;  // the mirror's device address
    int32_t tick = 1;
    int32_t dec_seq = 0;  // sequence number of the last decision enqueued
    int decide_spin = 1;  // 1: poll the mirror for the decision; 0: stream synchronize
    hipEvent_t ev_a = nullptr, ev_b = nullptr, ev_r0 = nullptr, ev_r1 = nullptr;

    // knobs
    int push_chunk = 2, all_chunk = 2, pull_chunk = 2;
    int advance_blocks = 0;  // 0 = derive from the CU count
    int verify = 1;
    int max_rounds = 100000;
    double push_edge_frac = 0.05;  // heuristic: sparse push below this frontier-edge fraction
    double pull_frac = 0.5;        // heuristic: pull when unsettled in-edges < pull_frac * frontier edges
    double delta = 0.0;            // push rounds defer vertices with d >= min + delta (0 = off)
    int hub_degree = 0;            // ... unless their out-degree is below this
    int64_t plain_below_edges = (int64_t)1 << 26;  // heuristic/learned: plain passes below this E
    double dense_all_frac = 0.15;  // all-edge rounds with frontier edges >= this * E run dense
                                   // (0.4 when the layout build finds no hub blocks),
    double dense_unreached = 0.5;  // ... and so do those with this sampled unreached fraction
    int dense_kernel = 0;  // dense rounds: 0 = destination-blocked pass, 1 = pass + compaction
    int plain_kernel = 0;  // plain passes: 0 = by tile count, 1 = edge-list pass, 2 = blocked pass
    int64_t tail_edges = (int64_t)1 << 23;  // heuristic/learned: a push round whose frontier has at
                                            // most this many out-edges starts the tail rounds (0 = off)
    int tail_group = 4;    // tail rounds queued per host check
    int tail_blocks = 0;   // workgroups of a tail round's relaxation (0 = 2 per CU)
    bool tail_armed = false;  // this solve's frontier has had more than tail_edges out-edges
    uint32_t *fbits = nullptr;  // push rounds' lowered-vertex bitmap (2 * ceil(V / 64) words, zero
                                // between rounds)
    int64_t push_bits_edges = 0;  // decided push rounds over at least this many frontier
                                  // out-edges mark lowered vertices in fbits (0 = never)
    bool last_round_dense = false;  // the last round ran dense (traces)
    int rollout_policy = 0;        // timing collection: rounds after the tried one follow 0 rule, 1 network
    int collect_kind = 0;          // timing collection of 0 SSSP, 1 BFS rounds
    int64_t collect_budget = (int64_t)1 << 22;  // rounds one timing collection may run in total
    int64_t collect_rounds = 0;    // rounds the last timing collection ran
    int trace_events = 0;

    // last solve
    int32_t last_source = -1;
    int last_sched = -1;
    bool have_result = false;     // g->dist holds SSSP distances of last_source
    bool bfs_result = false;      // g->dist holds BFS levels of last_source
    std::vector<sssp::TraceRow> trace;

    sssp_mlp *mlp = nullptr;      // scheduler network of SSSP solves
    sssp_mlp *bfs_mlp = nullptr;  // scheduler network of BFS solves
    bool solving_bfs = false;     // the rounds in flight are a BFS

    // batch state (allocated on demand)
    uint32_t *bdist = nullptr;
    int64_t bdist_elems = 0;
    int32_t *bscratch = nullptr;
    int64_t bscratch_elems = 0;
    int64_t bscratch_stride = 0;

    int64_t device_bytes = 0;

    // multi-source solves: device copy of the source list
    int32_t *src_list = nullptr;
    int32_t src_cap = 0;
    // batch policy chosen by the autotuner (0 = not tuned)
    int batch_policy = 0, batch_bs = 256, batch_wgpcu = 2;
    void *batch_exec = nullptr;  // hipGraphExec_t of the captured persistent launch
    int64_t batch_key[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    void *btasks = nullptr;      // device BatchTask array
    int32_t btask_cap = 0;
    int32_t *bcounter = nullptr;  // task counter + error flag
    int32_t *bround = nullptr;    // per-workgroup round counters
    int32_t bround_cap = 0;
};

namespace sssp {
// The plain pass kernel a graph uses: the destination-blocked pass once the
// blocked list has at least kPlainBlkMinTiles tiles (one 1024-thread
// workgroup each), else the XCD-partitioned edge-list pass, whose grid fills
// the chip whatever the graph's block structure. Measured crossover (plain
// solve, R-MAT edge factor 16): 44 tiles (scale 18) 1.08 against 0.96 ms for
// the edge-list pass, 83 tiles (scale 19) 1.27 against 1.57 ms; R-MAT
// 22 x 32 (1196 tiles) 7.1 against 15.6 ms.
constexpr int64_t kPlainBlkMinTiles = 64;

inline bool plain_uses_blocked(const sssp_graph *g)
{
    if (g->plain_kernel == 1 || g->n_btile == 0) return false;
    if (g->plain_kernel == 2) return true;
    return g->n_btile >= kPlainBlkMinTiles;
}
}  // namespace sssp

namespace sssp {

// Slot of caller vertex v (the identity on a graph that is not relabelled).
inline int32_t slot_of(const sssp_graph *g, int32_t v)
{
    return g->vperm_host.empty() ? v : g->vperm_host[(size_t)v];
}

// Block-wide exclusive scan of one int64 per thread (kBlock threads);
// s_warp must hold kBlock/64 entries. Contains barriers: call uniformly.
__device__ inline int64_t block_excl_scan_i64(int64_t x, int64_t *total, int64_t *s_warp)
{
    const int lane = __lane_id(), w = threadIdx.x >> 6;
    int64_t inc = x;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        int64_t y = __shfl_up(inc, o);
        if (lane >= o) inc += y;
    }
    if (lane == 63) s_warp[w] = inc;
    __syncthreads();
    int64_t wpre = 0, tot = 0;
#pragma unroll
    for (int i = 0; i < kBlock / 64; i++) {
        if (i < w) wpre += s_warp[i];
        tot += s_warp[i];
    }
    __syncthreads();
    *total = tot;
    return wpre + inc - x;
}

// gen.hip
void rmat_device(int scale, int ef, int a, int b, int c, int d, uint32_t base_seed, int T,
                 int32_t *src, int32_t *dst, float *w, hipStream_t s);

// layout.hip
// ids in [0, V), and for wtype SSSP_W_F32/F64/U32 weights >= 0 and not NaN
// (any other wtype: ids only); throws SSSP_ERR_ARG. bad_buf: 2 device
// counters to use, or null (allocated for the call)
void validate_edges(const int32_t *src, const int32_t *dst, const void *w, int wtype, int64_t E,
                    int32_t V, hipStream_t s, int64_t *bad_buf = nullptr);
// in-place exclusive scan of two arrays of nb + 1 values (the last one
// receives the total), one workgroup each, in one launch
void scan_blocks_pair_inplace(int64_t *a, int64_t *b, int64_t nb, hipStream_t s);
// packed 12-byte binary edge-list records (device) -> src, dst, w (device)
void split_records(const uint32_t *rec, int64_t n, int32_t *src, int32_t *dst, uint32_t *w,
                   hipStream_t s);
void build_layouts(sssp_graph *g, const int32_t *d_src, const int32_t *d_dst, const uint32_t *d_w,
                   bool build_csc);
// the multi-source records and tiles (m_rec, m_tile), built once per graph
void build_multi_layout(sssp_graph *g);
// rebuilds the destination-blocked list's tile table (cap: edge slots per
// tile, a multiple of kBlkChunk; order: 0 smallest first, 1 by block, 2 largest first, 3 by
// descending block)
void retile_blocked(sssp_graph *g, int64_t cap, int order, int range_bits = 0);
void device_exclusive_scan_i64(const int64_t *in, int64_t *out, int64_t n, int64_t *blk_a,
                               int64_t *blk_b, int64_t n_blk_cap, hipStream_t s);
int64_t scan_blocks_needed(int64_t n);
void build_tile_table(const int64_t *scan, int32_t n, int64_t total, int32_t tile_edges,
                      int32_t *tile_out, hipStream_t s);

// engine.hip
// in-degrees of the sample vertices (after the layouts and the sample)
void sample_in_degrees(sssp_graph *g);
// out[r * V + v] = in[r * V + vperm[v]] for rows r < rows: caller order from
// slot order (in and out must not overlap)
void unpermute_rows(const uint32_t *in, const int32_t *vperm, int64_t V, int64_t rows,
                    uint32_t *out, hipStream_t s);
// the certificate's parent tree of the last certify call, in caller ids, to host
void parents_out(sssp_graph *g, int32_t *parent_host);
template <class WT, class DT>
void ref_init(DT *d, int32_t V, int32_t source);
// changed_buf: the 64 change flags (device), or null (allocated for the call)
template <class WT, class DT>
void ref_bf(const int32_t *src, const int32_t *dst, const DT *w, int32_t V, int64_t E, DT *d,
            int32_t *passes_out, int32_t *changed_buf = nullptr);
// bellman_ford_wrapper<float> with its scratch in a file-scope device array:
// no device allocation (sssp_bellman_ford_noalloc_f32)
void ref_bf_noalloc_f32(const int32_t *src, const int32_t *dst, const uint32_t *w, int32_t V,
                        int64_t E, uint32_t *d, int32_t *passes_out);
// BFS on the graph's edges (weights ignored): hop levels in g->dist, any schedule
void bfs(sssp_graph *g, int32_t source, int sched, sssp_stats *st);
void certify_levels(sssp_graph *g, int32_t source, int64_t *counts);
void levels_to_visited(const uint32_t *lev, int32_t V, uint8_t *vis, hipStream_t s);
void ref_bfs(const int32_t *src, const int32_t *dst, int32_t V, int64_t E, uint8_t *visited,
             int32_t source, int32_t *passes_out);
void reset_stamps(sssp_graph *g);  // stamps: 0 (out-edges) / kSinkStamp; pile states: 0
void solve(sssp_graph *g, const int32_t *sources, int32_t n_src, int sched, sssp_stats *st);
void certify(sssp_graph *g, const int32_t *sources, int32_t n_src, int64_t *counts);
void collect_timings(sssp_graph *g, int32_t source, int32_t cap, float *feats, float *times,
                     int32_t *rows);

// mlp.hip
const float *mlp_device_params(sssp_mlp *m);

}  // namespace sssp
