// batch.hip -- many SSSP solves in one launch.
//
// Policy 1 (workgroup per task): one persistent kernel; each workgroup takes
// (graph range, source) tasks from an atomic counter -- largest tasks first --
// and runs a complete frontier Bellman-Ford for its task with workgroup
// barriers only (no grid barrier, so any occupancy is deadlock-free):
//   frontier chunk of BS vertices -> degree scan in LDS -> the chunk's edges
//   dealt over the BS lanes (binary search in LDS) -> check-then-atomicMin ->
//   appends to the next frontier through an LDS counter, deduplicated by a
//   per-workgroup stamp array. The launch (counter reset + persistent kernel)
//   is captured once in a hipGraph and replayed.
// Policy 2 (merged frontier): many disjoint graphs are one union graph, so all
//   their sources share one full-chip frontier solve; many sources on one
//   graph are solved one after another on the full chip.
// Policy 3 (many sources on one graph, vertex-major): the distances of up to
//   256 sources are one array D[v * n + s], so a vertex's distances are one
//   contiguous row. All-edge passes over the edges sorted by 16-vertex
//   destination block: a 256-thread workgroup (thread s = source s) holds its
//   block's 16 x 256 distances in LDS, reads each edge record once for all its
//   sources (a wave-uniform load), gathers the source vertex's row (coalesced,
//   eight rows in flight)
//   and lowers its own LDS column -- no atomics inside a block; the block is
//   written back at the end (atomicMin for blocks split over several tiles).
//   Passes run until one changes nothing, the plain Bellman-Ford fixpoint of
//   every source; the rows are then transposed into the per-source output.
// The autotuner times block size x workgroups per CU for policy 1, frontier
// threshold x push chunk width for policy 2, and policy 3, and keeps the
// fastest.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <string>

#include "sssp_graph.hpp"

namespace sssp {

struct BatchTask {
    int64_t vlo;     // first vertex of the task's graph (global id)
    int64_t doff;    // offset of the task's distance array in the batch output
    int32_t nv;      // vertices of the task's graph
    int32_t source;  // global id
};

namespace {

// Block-wide exclusive scan, BS threads.
template <int BS>
__device__ int64_t bscan(int64_t x, int64_t *total, int64_t *s_warp)
{
    const int lane = __lane_id(), w = threadIdx.x >> 6;
    int64_t inc = x;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int64_t y = __shfl_up(inc, o);
        if (lane >= o) inc += y;
    }
    if (lane == 63) s_warp[w] = inc;
    __syncthreads();
    int64_t pre = 0, tot = 0;
#pragma unroll
    for (int i = 0; i < BS / 64; i++) {
        if (i < w) pre += s_warp[i];
        tot += s_warp[i];
    }
    __syncthreads();
    *total = tot;
    return pre + inc - x;
}

__global__ void k_batch_reset(int32_t *counter)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        counter[0] = 0;  // next task
        counter[1] = 0;  // error flag
    }
}

template <class WT, int BS>
__global__ __launch_bounds__(BS) void k_wg_solve(const BatchTask *__restrict__ tasks, int32_t ntasks,
                                                 int32_t *counter, const int64_t *__restrict__ row_ptr,
                                                 const int32_t *__restrict__ adj,
                                                 const uint32_t *__restrict__ wts, uint32_t *bdist,
                                                 int32_t *scratch, int64_t maxv, int32_t *round_ctr)
{
    constexpr uint32_t INF = WT::INF;
    __shared__ int s_task, s_n;
    __shared__ int32_t s_v[BS];
    __shared__ int64_t s_start[BS];
    __shared__ int64_t s_off[BS + 1];
    __shared__ uint32_t s_du[BS];
    __shared__ int64_t s_warp[BS / 64];
    const int tid = threadIdx.x;
    int32_t *qa = scratch + (int64_t)blockIdx.x * 3 * maxv;
    int32_t *qb = qa + maxv;
    int32_t *stamp = qb + maxv;  // local index; rounds only ever increase
    int32_t rt = round_ctr[blockIdx.x];
    for (;;) {
        if (tid == 0) s_task = atomicAdd(&counter[0], 1);
        __syncthreads();
        const int t = s_task;
        __syncthreads();
        if (t >= ntasks) break;
        const BatchTask T = tasks[t];
        uint32_t *dist = bdist + T.doff - T.vlo;  // indexed by global vertex id
        for (int64_t v = tid; v < T.nv; v += BS) dist[T.vlo + v] = INF;
        __syncthreads();
        if (tid == 0) {
            dist[T.source] = 0u;
            qa[0] = T.source;
        }
        int32_t n = 1;
        int32_t rounds = 0;
        int32_t *cur = qa, *nxt = qb;
        __syncthreads();
        while (n > 0) {
            if (++rounds > T.nv + 2) {  // cannot happen for w >= 0; bounded anyway
                if (tid == 0) counter[1] = 1;
                break;
            }
            rt++;
            if (tid == 0) s_n = 0;
            __syncthreads();
            for (int32_t base = 0; base < n; base += BS) {
                const int32_t i = base + tid;
                int32_t u = 0;
                int64_t a = 0, deg = 0;
                uint32_t du = INF;
                if (i < n) {
                    u = cur[i];
                    a = row_ptr[u];
                    deg = row_ptr[u + 1] - a;
                    du = dist[u];
                }
                s_v[tid] = u;
                s_start[tid] = a;
                s_du[tid] = du;
                int64_t tot;
                const int64_t off = bscan<BS>(deg, &tot, s_warp);
                s_off[tid] = off;
                if (tid == BS - 1) s_off[BS] = tot;
                __syncthreads();
                for (int64_t e = tid; e < tot; e += BS) {
                    int lo = 0, hi = BS - 1;  // largest j with s_off[j] <= e
                    while (lo < hi) {
                        const int mid = (lo + hi + 1) >> 1;
                        if (s_off[mid] <= e) lo = mid;
                        else hi = mid - 1;
                    }
                    const int64_t edge = s_start[lo] + (e - s_off[lo]);
                    const int32_t v = adj[edge];
                    const uint32_t cand = WT::add(s_du[lo], wts[edge]);
                    // rounds share one launch, so a plain read may be from an
                    // earlier round: the returned old value decides "lowered";
                    // a stale stamp is an older round and falls through to the
                    // exchange
                    if (cand < dist[v]) {
                        const uint32_t old = atomicMin(&dist[v], cand);
                        int32_t *sp = &stamp[v - T.vlo];
                        if (cand < old && *sp != rt && atomicExch(sp, rt) != rt)
                            nxt[atomicAdd(&s_n, 1)] = v;
                    }
                }
                __syncthreads();
            }
            n = s_n;
            int32_t *tmp = cur;
            cur = nxt;
            nxt = tmp;
            __syncthreads();
        }
    }
    if (tid == 0) round_ctr[blockIdx.x] = rt;
}

template <int BS>
void launch_wg(int wtype, unsigned grid, hipStream_t s, const BatchTask *tasks, int32_t ntasks,
               int32_t *counter, sssp_graph *g, uint32_t *bdist, int32_t *scratch, int64_t maxv)
{
    if (wtype == SSSP_W_F32)
        k_wg_solve<WF32, BS><<<grid, BS, 0, s>>>(tasks, ntasks, counter, g->row_ptr, g->el_dst,
                                                 g->el_w, bdist, scratch, maxv, g->bround);
    else
        k_wg_solve<WU32, BS><<<grid, BS, 0, s>>>(tasks, ntasks, counter, g->row_ptr, g->el_dst,
                                                 g->el_w, bdist, scratch, maxv, g->bround);
    SSSP_LAUNCH_CHECK("k_wg_solve");
}

struct Config {
    int policy;      // 1 workgroup per task, 2 merged frontier / sequential
    int bs;          // policy 1 block size
    int wgpcu;       // policy 1 workgroups per CU
    double frac;     // policy 2 push/all frontier threshold
    int push_chunk;  // policy 2 sparse push chunk bucket
};

void ensure_batch_buffers(sssp_graph *g, int64_t dist_elems, int32_t ntasks, int32_t nwg,
                          int64_t maxv)
{
    if (dist_elems > g->bdist_elems) {
        dfree(g->bdist);
        g->bdist = nullptr;
        g->bdist = dmalloc<uint32_t>((size_t)dist_elems);
        g->bdist_elems = dist_elems;
    }
    if (ntasks > g->btask_cap) {
        dfree(g->btasks);
        g->btasks = nullptr;
        g->btasks = dmalloc<BatchTask>((size_t)ntasks);
        g->btask_cap = ntasks;
    }
    if (!g->bcounter) g->bcounter = dmalloc<int32_t>(2);
    // per-workgroup scratch = [queue a | queue b | stamps], each `stride`
    // entries; the stride only grows, so stamp regions never move onto old
    // queue contents
    if (nwg > 0 && (maxv > g->bscratch_stride || nwg > g->bround_cap)) {
        const int64_t stride = std::max<int64_t>(maxv, g->bscratch_stride);
        const int32_t cap = std::max<int32_t>(nwg, g->bround_cap);
        const int64_t need = (int64_t)cap * 3 * stride;
        dfree(g->bscratch);
        dfree(g->bround);
        g->bscratch = nullptr;
        g->bround = nullptr;
        g->bscratch = dmalloc<int32_t>((size_t)need);
        g->bscratch_elems = need;
        g->bscratch_stride = stride;
        g->bround = dmalloc<int32_t>((size_t)cap);
        g->bround_cap = cap;
        // stamps start below every round number the kernel will use
        SSSP_HIP(hipMemsetAsync(g->bscratch, 0, 4 * (size_t)need, g->stream));
        SSSP_HIP(hipMemsetAsync(g->bround, 0, 4 * (size_t)cap, g->stream));
    }
}

// Policy 1: tasks already uploaded; captures (counter reset, persistent
// kernel) in a hipGraph keyed on the launch shape and replays it.
double run_wg(sssp_graph *g, int32_t ntasks, int64_t maxv, int bs, int wgpcu)
{
    hipStream_t s = g->stream;
    if (maxv > g->bscratch_stride) throw Error{SSSP_ERR_INTERNAL, "batch scratch too small"};
    maxv = g->bscratch_stride;
    int32_t nwg = g->cu_count * wgpcu;
    if (nwg > ntasks) nwg = ntasks;
    if (nwg < 1) nwg = 1;
    const int64_t key[8] = {ntasks, maxv, bs, nwg, (int64_t)(uintptr_t)g->bdist,
                            (int64_t)(uintptr_t)g->bscratch, (int64_t)(uintptr_t)g->btasks,
                            (int64_t)(uintptr_t)g->bround};
    if (!g->batch_exec || std::memcmp(key, g->batch_key, sizeof(key)) != 0) {
        if (g->batch_exec) {
            (void)hipGraphExecDestroy((hipGraphExec_t)g->batch_exec);
            g->batch_exec = nullptr;
        }
        SSSP_HIP(hipStreamSynchronize(s));
        hipGraph_t graph;
        SSSP_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        k_batch_reset<<<1, 64, 0, s>>>(g->bcounter);
        const BatchTask *tasks = static_cast<const BatchTask *>(g->btasks);
        switch (bs) {
        case 256: launch_wg<256>(g->wtype, nwg, s, tasks, ntasks, g->bcounter, g, g->bdist, g->bscratch, maxv); break;
        case 512: launch_wg<512>(g->wtype, nwg, s, tasks, ntasks, g->bcounter, g, g->bdist, g->bscratch, maxv); break;
        default: launch_wg<1024>(g->wtype, nwg, s, tasks, ntasks, g->bcounter, g, g->bdist, g->bscratch, maxv); break;
        }
        SSSP_HIP(hipStreamEndCapture(s, &graph));
        hipGraphExec_t exec;
        const hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        SSSP_HIP(e);
        g->batch_exec = exec;
        std::memcpy(g->batch_key, key, sizeof(key));
    }
    SSSP_HIP(hipEventRecord(g->ev_a, s));
    SSSP_HIP(hipGraphLaunch((hipGraphExec_t)g->batch_exec, s));
    SSSP_HIP(hipEventRecord(g->ev_b, s));
    SSSP_HIP(hipEventSynchronize(g->ev_b));
    float ms = 0.f;
    SSSP_HIP(hipEventElapsedTime(&ms, g->ev_a, g->ev_b));
    int32_t flags[2] = {0, 0};
    SSSP_HIP(hipMemcpy(flags, g->bcounter, sizeof(flags), hipMemcpyDeviceToHost));
    if (flags[1]) throw Error{SSSP_ERR_INTERNAL, "batched solve exceeded its round bound"};
    return ms;
}

void upload_tasks(sssp_graph *g, std::vector<BatchTask> &tasks, const std::vector<int64_t> &work)
{
    // largest tasks first (longest-processing-time order)
    std::vector<size_t> order(tasks.size());
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) { return work[a] > work[b]; });
    std::vector<BatchTask> sorted(tasks.size());
    for (size_t i = 0; i < order.size(); i++) sorted[i] = tasks[order[i]];
    SSSP_HIP(hipMemcpyAsync(g->btasks, sorted.data(), sizeof(BatchTask) * sorted.size(),
                            hipMemcpyHostToDevice, g->stream));
}

double run_multi(sssp_graph *g, const int32_t *sources, int32_t n);

// ---- many sources on one graph ----------------------------------------------
double batch_sources(sssp_graph *g, const int32_t *sources, int32_t n, const Config &c)
{
    const int64_t V = g->V;
    if (c.policy == 3) {
        ensure_batch_buffers(g, (int64_t)n * V, 0, 0, 0);
        return run_multi(g, sources, n);
    }
    if (c.policy == 1) {
        int32_t nwg = std::min<int32_t>(n, g->cu_count * c.wgpcu);
        ensure_batch_buffers(g, (int64_t)n * V, n, nwg, V);
        std::vector<BatchTask> tasks((size_t)n);
        std::vector<int64_t> work((size_t)n, 1);
        for (int32_t i = 0; i < n; i++) tasks[(size_t)i] = BatchTask{0, (int64_t)i * V, (int32_t)V, sources[i]};
        upload_tasks(g, tasks, work);
        // sources sorted by work are all equal here: order is preserved
        double ms = run_wg(g, n, V, c.bs, c.wgpcu);
        if (g->vperm) {
            // rows to the caller's vertex order, one row at a time through
            // dout (timed with the solve)
            hipStream_t s = g->stream;
            SSSP_HIP(hipEventRecord(g->ev_a, s));
            for (int32_t i = 0; i < n; i++) {
                uint32_t *row = g->bdist + (int64_t)i * V;
                unpermute_rows(row, g->vperm, V, 1, g->dout, s);
                SSSP_HIP(hipMemcpyAsync(row, g->dout, 4 * (size_t)V, hipMemcpyDeviceToDevice, s));
            }
            SSSP_HIP(hipEventRecord(g->ev_b, s));
            SSSP_HIP(hipEventSynchronize(g->ev_b));
            float extra = 0.f;
            SSSP_HIP(hipEventElapsedTime(&extra, g->ev_a, g->ev_b));
            ms += extra;
        }
        return ms;
    }
    ensure_batch_buffers(g, (int64_t)n * V, 0, 0, 0);
    const double saved = g->push_edge_frac;
    const int saved_chunk = g->push_chunk;
    g->push_edge_frac = c.frac;
    g->push_chunk = c.push_chunk;
    double total = 0;
    for (int32_t i = 0; i < n; i++) {
        sssp_stats st{};
        solve(g, &sources[i], 1, SSSP_SCHED_HEURISTIC, &st);
        SSSP_HIP(hipMemcpyAsync(g->bdist + (int64_t)i * V, g->dout, 4 * (size_t)V,
                                hipMemcpyDeviceToDevice, g->stream));
        total += st.t_solve_ms;
    }
    SSSP_HIP(hipStreamSynchronize(g->stream));
    g->push_edge_frac = saved;
    g->push_chunk = saved_chunk;
    return total;
}

// ---- policy 3: vertex-major all-edge passes ---------------------------------
constexpr int kMThreads = 256;  // sources per chunk: one thread per source
constexpr int kMBlock = 1 << kMBits;

// One tile of a vertex-major pass: thread s holds source s's distances of
// the tile's 16-vertex destination block in its LDS column, lowers them with
// the tile's records (a record's source row gathered for all sources at
// once), and writes back what it lowered. Returns whether this lane lowered a
// distance.
template <class WT>
__device__ __forceinline__ bool multi_tile(const BlkTile &t, const MRec *__restrict__ rec,
                                           uint32_t *D, int32_t n, int64_t V,
                                           uint32_t (&ld)[kMBlock][kMThreads])
{
    constexpr uint32_t INF = WT::INF;
    const int s = threadIdx.x;
    const bool act = s < n;
    const int64_t base = (int64_t)t.blk << kMBits;
    const int nb = (int)min((int64_t)kMBlock, V - base);
    // each thread touches only its own column: no barrier is needed
    for (int v = 0; v < nb; v++) ld[v][s] = act ? D[(base + v) * n + s] : INF;
    auto relax = [&](const MRec &r, uint32_t du) {
        if (du == INF) return;
        const uint32_t cand = WT::add(du, r.w);
        if (cand < ld[r.loc][s]) ld[r.loc][s] = cand;
    };
    int64_t e = t.e0;
    // eight records at a time: their row gathers are in flight together.
    // Measured at R-MAT scale 20 x 256 sources (GTEPS): 64-vertex blocks (64
    // KB of LDS, 2 waves per SIMD) 23.0 with 8 in flight, 25.8-27.4 with 16;
    // 32-vertex blocks 44.0 (16); 16-vertex blocks (16 KB, 8 workgroups per
    // CU) 51.4 (16), 54.7 (8); 8-vertex blocks 51.1 (16)
    constexpr int R = 8;
    MRec r[R];
    if (e + R <= t.e1) {
#pragma unroll
        for (int k = 0; k < R; k++) r[k] = rec[e + k];
    }
    for (; e + R <= t.e1; e += R) {
        uint32_t du[R];
#pragma unroll
        for (int k = 0; k < R; k++) du[k] = act ? D[(int64_t)r[k].src * n + s] : INF;
        // the next step's records load under this step's gathers
        MRec rn[R];
        if (e + 2 * R <= t.e1) {
#pragma unroll
            for (int k = 0; k < R; k++) rn[k] = rec[e + R + k];
        }
#pragma unroll
        for (int k = 0; k < R; k++) relax(r[k], du[k]);
#pragma unroll
        for (int k = 0; k < R; k++) r[k] = rn[k];
    }
    for (; e < t.e1; e++) {
        const MRec r = rec[e];
        relax(r, act ? D[(int64_t)r.src * n + s] : INF);
    }
    bool low = false;
    if (act) {
        for (int v = 0; v < nb; v++) {
            const uint32_t nv = ld[v][s];
            const int64_t idx = (base + v) * n + s;
            if (nv < D[idx]) {
                if (t.split)
                    __hip_atomic_fetch_min(&D[idx], nv, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                else
                    D[idx] = nv;  // only this workgroup writes the block
                low = true;
            }
        }
    }
    return low;
}

// One pass per launch, one workgroup per tile. The pass number is base[0] +
// k (base in device memory, so a captured run of passes is replayed as the
// next run); pass p zeroes the flag pass p+1 sets and returns at once when
// pass p-1 changed nothing.
template <class WT>
__global__ __launch_bounds__(kMThreads) void k_multi_pass(const BlkTile *__restrict__ tiles,
                                                          const MRec *__restrict__ rec,
                                                          uint32_t *D, int32_t n, int64_t V,
                                                          int32_t *changed, const int32_t *base,
                                                          int32_t k)
{
    const int32_t pass = *base + k;
    if (blockIdx.x == 0 && threadIdx.x == 0) changed[(pass + 1) & 63] = 0;
    if (pass > 0 && __hip_atomic_load(&changed[(pass - 1) & 63], __ATOMIC_RELAXED,
                                      __HIP_MEMORY_SCOPE_AGENT) == 0)
        return;
    __shared__ uint32_t ld[kMBlock][kMThreads];
    const bool low = multi_tile<WT>(tiles[blockIdx.x], rec, D, n, V, ld);
    if (__ballot(low) != 0 && __lane_id() == 0) changed[pass & 63] = 1;
}

// D[v * n + s] = INF, D[sources[s] * n + s] = 0
template <class WT>
__global__ __launch_bounds__(kBlock) void k_multi_init(uint32_t *D, int32_t n, int64_t V,
                                                       const int32_t *sources)
{
    const int64_t total = V * n;
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < total;
         i += (int64_t)gridDim.x * kBlock) {
        const int64_t v = i / n;
        const int32_t s = (int32_t)(i - v * n);
        D[i] = sources[s] == v ? 0u : WT::INF;
    }
}

// out[s * V + v] = D[slot(v) * n + s] through 64 x 64 LDS tiles (slot(v) =
// vperm[v] on a relabelled graph, else v): per source, caller vertex order
__global__ __launch_bounds__(kBlock) void k_multi_out(const uint32_t *D, int32_t n, int64_t V,
                                                      const int32_t *vperm, uint32_t *out)
{
    __shared__ uint32_t tl[64][65];
    const int64_t v0 = (int64_t)blockIdx.x * 64;
    const int s0 = blockIdx.y * 64;
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;  // 64 x 4
    for (int r = ty; r < 64; r += 4) {
        const int64_t v = v0 + r;
        const int s = s0 + tx;
        if (v < V && s < n) tl[r][tx] = D[(vperm ? (int64_t)vperm[v] : v) * n + s];
    }
    __syncthreads();
    for (int r = ty; r < 64; r += 4) {
        const int s = s0 + r;
        const int64_t v = v0 + tx;
        if (v < V && s < n) out[(int64_t)s * V + v] = tl[tx][r];
    }
}

__global__ void k_zero_words(int32_t *p, int32_t n)
{
    for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0;
}

// base[0] += k (the next run of captured passes)
__global__ void k_advance_base(int32_t *base, int32_t k)
{
    if (threadIdx.x == 0) base[0] += k;
}

constexpr int32_t kMultiRun = 32;  // passes per captured run (the flag ring has 64)

template <class WT>
void launch_multi_passes(sssp_graph *g, int32_t m, hipStream_t s)
{
    for (int32_t k = 0; k < kMultiRun; k++)
        k_multi_pass<WT><<<(unsigned)g->n_mtile, kMThreads, 0, s>>>(
            g->m_tile, g->m_rec, g->mdist, m, g->V, g->mflags, g->mflags + 64, k);
}

// The captured launches of one chunk of at most 256 sources, keyed on chunk
// shape and buffers: `first` clears the change flags and the pass base with a
// kernel, initialises the vertex-major distances, runs kMultiRun passes and
// transposes the result into the per-source output; `more` advances the pass
// base by kMultiRun and runs the next kMultiRun passes and the transpose.
// Passes after the fixpoint return at once.
struct MultiGraphs {
    hipGraphExec_t first, more;
};

MultiGraphs multi_graphs(sssp_graph *g, int32_t c0, int32_t m)
{
    const int64_t V = g->V;
    const int64_t key[kMultiKey] = {m,
                                    c0,
                                    V,
                                    g->n_mtile,
                                    (int64_t)(uintptr_t)g->mflags,
                                    (int64_t)(uintptr_t)g->mdist,
                                    (int64_t)(uintptr_t)g->bdist,
                                    (int64_t)(uintptr_t)g->msrc,
                                    g->wtype,
                                    (int64_t)(uintptr_t)g->m_tile};
    for (int i = 0; i < kMultiSlots; i++)
        if (g->mexec[i][0] && std::memcmp(key, g->mkey[i], sizeof(key)) == 0)
            return MultiGraphs{(hipGraphExec_t)g->mexec[i][0], (hipGraphExec_t)g->mexec[i][1]};
    const int slot = g->mnext;
    g->mnext = (g->mnext + 1) % kMultiSlots;
    for (int j = 0; j < 2; j++)
        if (g->mexec[slot][j]) {
            (void)hipGraphExecDestroy((hipGraphExec_t)g->mexec[slot][j]);
            g->mexec[slot][j] = nullptr;
        }
    hipStream_t s = g->stream;
    SSSP_HIP(hipStreamSynchronize(s));
    const bool u32 = g->wtype == SSSP_W_U32;
    const unsigned gi = (unsigned)std::min<int64_t>((V * m + kBlock - 1) / kBlock, 65536);
    for (int j = 0; j < 2; j++) {
        hipGraph_t graph;
        SSSP_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        if (j == 0) {
            k_zero_words<<<1, 128, 0, s>>>(g->mflags, 65);
            if (u32) k_multi_init<WU32><<<gi, kBlock, 0, s>>>(g->mdist, m, V, g->msrc + c0);
            else k_multi_init<WF32><<<gi, kBlock, 0, s>>>(g->mdist, m, V, g->msrc + c0);
        } else {
            k_advance_base<<<1, 64, 0, s>>>(g->mflags + 64, kMultiRun);
        }
        if (g->n_mtile > 0) {
            if (u32) launch_multi_passes<WU32>(g, m, s);
            else launch_multi_passes<WF32>(g, m, s);
        }
        k_multi_out<<<dim3((unsigned)((V + 63) / 64), (unsigned)((m + 63) / 64)), kBlock, 0, s>>>(
            g->mdist, m, V, g->vperm, g->bdist + (int64_t)c0 * V);
        const hipError_t le = hipGetLastError();
        const hipError_t ce = hipStreamEndCapture(s, &graph);
        if (le != hipSuccess || ce != hipSuccess) {
            if (ce == hipSuccess) (void)hipGraphDestroy(graph);
            throw Error{SSSP_ERR_HIP, std::string("capturing the batch launch failed: ") +
                                          hipGetErrorString(le != hipSuccess ? le : ce)};
        }
        hipGraphExec_t exec;
        const hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        SSSP_HIP(ie);
        g->mexec[slot][j] = exec;
    }
    std::memcpy(g->mkey[slot], key, sizeof(key));
    return MultiGraphs{(hipGraphExec_t)g->mexec[slot][0], (hipGraphExec_t)g->mexec[slot][1]};
}

// Passes of one chunk as queued launches (batch_capture = 0): four at a time,
// the host reading the change flags after each four.
int32_t multi_passes_launched(sssp_graph *g, int32_t m)
{
    hipStream_t s = g->stream;
    const int64_t V = g->V;
    int32_t *changed = g->mflags;
    k_zero_words<<<1, 128, 0, s>>>(changed, 65);
    int32_t pass = 0, done_at = -1;
    while (done_at < 0 && g->n_mtile > 0) {
        for (int i = 0; i < 4; i++) {
            if (g->wtype == SSSP_W_U32)
                k_multi_pass<WU32><<<(unsigned)g->n_mtile, kMThreads, 0, s>>>(
                    g->m_tile, g->m_rec, g->mdist, m, V, changed, changed + 64, pass + i);
            else
                k_multi_pass<WF32><<<(unsigned)g->n_mtile, kMThreads, 0, s>>>(
                    g->m_tile, g->m_rec, g->mdist, m, V, changed, changed + 64, pass + i);
            SSSP_LAUNCH_CHECK("k_multi_pass");
        }
        SSSP_HIP(hipMemcpyAsync(g->ctrl_host->changed, changed, 4 * 64, hipMemcpyDeviceToHost, s));
        SSSP_HIP(hipStreamSynchronize(s));
        for (int i = 0; i < 4 && done_at < 0; i++)
            if (g->ctrl_host->changed[(pass + i) & 63] == 0) done_at = pass + i;
        pass += 4;
        if (done_at < 0 && pass > g->V + 4)
            throw Error{SSSP_ERR_NOT_CONVERGED, "multi-source passes exceeded V passes"};
    }
    return done_at + 1;
}

double run_multi(sssp_graph *g, const int32_t *sources, int32_t n)
{
    hipStream_t s = g->stream;
    build_multi_layout(g);
    const int64_t V = g->V;
    const int32_t chunk = std::min<int32_t>(n, kMThreads);
    if ((int64_t)chunk * V > g->mdist_elems) {
        dfree(g->mdist);
        g->mdist = nullptr;
        g->mdist = dmalloc<uint32_t>((size_t)chunk * (size_t)V);
        g->mdist_elems = (int64_t)chunk * V;
    }
    if (n > g->msrc_cap) {
        dfree(g->msrc);
        g->msrc = nullptr;
        g->msrc = dmalloc<int32_t>((size_t)n);
        g->msrc_cap = n;
    }
    if (!g->mflags) g->mflags = dmalloc<int32_t>(65);  // 64 change flags, the pass base
    SSSP_HIP(hipMemcpyAsync(g->msrc, sources, 4 * (size_t)n, hipMemcpyHostToDevice, s));
    float total = 0.f;
    g->multi_passes = 0;
    for (int32_t c0 = 0; c0 < n; c0 += kMThreads) {
        const int32_t m = std::min<int32_t>(kMThreads, n - c0);
        if (g->batch_capture) {
            const MultiGraphs mg = multi_graphs(g, c0, m);
            SSSP_HIP(hipEventRecord(g->ev_a, s));
            int32_t done_at = -1;
            // replay until a run holds a pass that changed nothing (at most
            // V + 1 passes, as the pass loop)
            for (int32_t run = 0; done_at < 0; run++) {
                if (g->n_mtile == 0) {
                    SSSP_HIP(hipGraphLaunch(mg.first, s));
                    done_at = 0;
                    break;
                }
                if ((int64_t)run * kMultiRun > V + 1)
                    throw Error{SSSP_ERR_NOT_CONVERGED, "multi-source passes exceeded V passes"};
                SSSP_HIP(hipGraphLaunch(run == 0 ? mg.first : mg.more, s));
                SSSP_HIP(hipMemcpyAsync(g->ctrl_host->changed, g->mflags, 4 * 64,
                                        hipMemcpyDeviceToHost, s));
                SSSP_HIP(hipStreamSynchronize(s));
                for (int32_t i = 0; i < kMultiRun && done_at < 0; i++) {
                    const int32_t p = run * kMultiRun + i;
                    if (g->ctrl_host->changed[p & 63] == 0) done_at = p;
                }
            }
            SSSP_HIP(hipEventRecord(g->ev_b, s));
            SSSP_HIP(hipEventSynchronize(g->ev_b));
            g->multi_passes += done_at + 1;
        } else {
            SSSP_HIP(hipEventRecord(g->ev_a, s));
            const unsigned gi = (unsigned)std::min<int64_t>((V * m + kBlock - 1) / kBlock, 65536);
            if (g->wtype == SSSP_W_U32)
                k_multi_init<WU32><<<gi, kBlock, 0, s>>>(g->mdist, m, V, g->msrc + c0);
            else
                k_multi_init<WF32><<<gi, kBlock, 0, s>>>(g->mdist, m, V, g->msrc + c0);
            SSSP_LAUNCH_CHECK("k_multi_init");
            g->multi_passes += multi_passes_launched(g, m);
            k_multi_out<<<dim3((unsigned)((V + 63) / 64), (unsigned)((m + 63) / 64)), kBlock, 0,
                          s>>>(g->mdist, m, V, g->vperm, g->bdist + (int64_t)c0 * V);
            SSSP_LAUNCH_CHECK("k_multi_out");
            SSSP_HIP(hipEventRecord(g->ev_b, s));
            SSSP_HIP(hipEventSynchronize(g->ev_b));
        }
        float ms = 0.f;
        SSSP_HIP(hipEventElapsedTime(&ms, g->ev_a, g->ev_b));
        total += ms;
    }
    return total;
}

// one_graph: many sources on one graph (policy 3 applies); else a pack
std::vector<Config> candidates(bool one_graph)
{
    std::vector<Config> c;
    if (one_graph) c.push_back(Config{3, kMThreads, 1, 0.0, 0});
    for (int bs : {256, 512, 1024})
        for (int w : {1, 2, 4}) c.push_back(Config{1, bs, w, 0.05, 1});
    for (double f : {0.02, 0.05, 0.15})
        for (int pc : {1, 2}) c.push_back(Config{2, 256, 2, f, pc});
    return c;
}

std::string describe(const Config &c)
{
    char buf[160];
    if (c.policy == 1)
        std::snprintf(buf, sizeof(buf), "policy=workgroup_per_task block=%d wg_per_cu=%d", c.bs, c.wgpcu);
    else if (c.policy == 3)
        std::snprintf(buf, sizeof(buf), "policy=vertex_major_all_edge sources_per_chunk=%d", c.bs);
    else
        std::snprintf(buf, sizeof(buf), "policy=merged_frontier push_edge_frac=%.3f push_chunk=%d", c.frac,
                      c.push_chunk);
    return buf;
}

template <class Run>
Config autotune(Run &&run, const char *file, double *best_ms, const char *what, bool one_graph)
{
    std::vector<Config> cs = candidates(one_graph);
    std::vector<double> t(cs.size(), 1e300);
    for (size_t i = 0; i < cs.size(); i++) {
        (void)run(cs[i]);  // warm-up (and graph capture)
        t[i] = std::min(run(cs[i]), run(cs[i]));
    }
    const size_t b = std::min_element(t.begin(), t.end()) - t.begin();
    if (file) {
        FILE *f = std::fopen(file, "w");
        if (!f) throw Error{SSSP_ERR_IO, std::string("cannot open tuning file ") + file};
        std::fprintf(f, "# autotune %s: min of 2 timed runs after 1 warm-up, device ms\n", what);
        for (size_t i = 0; i < cs.size(); i++)
            std::fprintf(f, "%s ms=%.4f%s\n", describe(cs[i]).c_str(), t[i], i == b ? " best" : "");
        std::fclose(f);
    }
    if (best_ms) *best_ms = t[b];
    return cs[b];
}

Config default_config(sssp_graph *g, int32_t n)
{
    if (g->batch_policy)
        return Config{g->batch_policy, g->batch_bs, g->batch_wgpcu, g->push_edge_frac, g->push_chunk};
    // untuned: policy 3 for 32 or more sources on one graph (R-MAT scale 20
    // x 256 sources: 54.7 GTEPS against 10.5 for policy 1), policy 2 below
    if (n >= 32) return Config{3, kMThreads, 1, 0.0, 0};
    return Config{2, 256, 2, g->push_edge_frac, g->push_chunk};
}

void fill_stats(sssp_stats *st, double ms, double edges)
{
    if (!st) return;
    std::memset(st, 0, sizeof(*st));
    st->t_solve_ms = ms;
    st->gteps = ms > 0 ? edges / (ms * 1e-3) / 1e9 : 0.0;
}

}  // namespace
}  // namespace sssp

using namespace sssp;

struct sssp_pack {
    sssp_graph *g = nullptr;  // the disjoint union of all graphs
    int32_t n = 0;
    std::vector<int64_t> voff;  // n+1 vertex offsets
    std::vector<int64_t> eoff;  // n+1 edge offsets
    int policy = 0, bs = 256, wgpcu = 2;
    double frac = 0.05;
    int push_chunk = 1;
};

namespace sssp {
namespace {

double pack_run(sssp_pack *p, const int32_t *local_sources, const Config &c)
{
    sssp_graph *g = p->g;
    const int32_t n = p->n;
    if (c.policy == 1) {
        int64_t maxv = 0;
        for (int32_t k = 0; k < n; k++) maxv = std::max<int64_t>(maxv, p->voff[k + 1] - p->voff[k]);
        const int32_t nwg = std::min<int32_t>(n, g->cu_count * c.wgpcu);
        ensure_batch_buffers(g, g->V, n, nwg, maxv);
        std::vector<BatchTask> tasks((size_t)n);
        std::vector<int64_t> work((size_t)n);
        for (int32_t k = 0; k < n; k++) {
            tasks[(size_t)k] = BatchTask{p->voff[k], p->voff[k], (int32_t)(p->voff[k + 1] - p->voff[k]),
                                         (int32_t)(p->voff[k] + local_sources[k])};
            work[(size_t)k] = p->eoff[k + 1] - p->eoff[k];
        }
        upload_tasks(g, tasks, work);
        return run_wg(g, n, maxv, c.bs, c.wgpcu);
    }
    std::vector<int32_t> gs((size_t)n);
    for (int32_t k = 0; k < n; k++) gs[(size_t)k] = (int32_t)(p->voff[k] + local_sources[k]);
    const double saved = g->push_edge_frac;
    const int saved_chunk = g->push_chunk;
    g->push_edge_frac = c.frac;
    g->push_chunk = c.push_chunk;
    sssp_stats st{};
    solve(g, gs.data(), n, SSSP_SCHED_HEURISTIC, &st);
    g->push_edge_frac = saved;
    g->push_chunk = saved_chunk;
    ensure_batch_buffers(g, g->V, 0, 0, 0);
    SSSP_HIP(hipMemcpyAsync(g->bdist, g->dout, 4 * (size_t)g->V, hipMemcpyDeviceToDevice, g->stream));
    SSSP_HIP(hipStreamSynchronize(g->stream));
    return st.t_solve_ms;
}

std::vector<int32_t> source_slots(const sssp_graph *g, const int32_t *sources, int32_t n)
{
    std::vector<int32_t> out((size_t)n);
    for (int32_t i = 0; i < n; i++) out[(size_t)i] = slot_of(g, sources[i]);
    return out;
}

void check_pack_sources(const sssp_pack *p, const int32_t *sources)
{
    SSSP_REQUIRE(sources, "null sources");
    for (int32_t k = 0; k < p->n; k++)
        SSSP_REQUIRE(sources[k] >= 0 && sources[k] < p->voff[k + 1] - p->voff[k],
                     "pack source out of range for graph " + std::to_string(k));
}

}  // namespace
}  // namespace sssp

extern "C" {

int sssp_solve_batch(sssp_graph *g, const int32_t *sources, int32_t n, int32_t policy,
                     void *dist_out, sssp_stats *stats)
{
    return guard([&] {
        SSSP_REQUIRE(g && sources && n >= 1, "bad arguments");
        SSSP_REQUIRE(policy >= 0 && policy <= 3, "policy must be 0, 1, 2 or 3");
        for (int32_t i = 0; i < n; i++)
            SSSP_REQUIRE(sources[i] >= 0 && sources[i] < g->V, "source vertex out of range");
        SSSP_HIP(hipSetDevice(g->device));
        // the merged/sequential policy solves through g->dist: no single-solve
        // result (distances or BFS levels) survives a batch
        g->have_result = false;
        g->bfs_result = false;
        Config c = default_config(g, n);
        if (policy != 0 && policy != c.policy)
            c = (policy == 1)   ? Config{1, 1024, 1, 0.05, 1}
                : (policy == 3) ? Config{3, kMThreads, 1, 0.0, 0}
                                : Config{2, 256, 2, g->push_edge_frac, g->push_chunk};
        const std::vector<int32_t> slots = source_slots(g, sources, n);
        const double ms = batch_sources(g, slots.data(), n, c);
        if (dist_out)
            SSSP_HIP(hipMemcpy(dist_out, g->bdist, 4 * (size_t)n * (size_t)g->V, hipMemcpyDeviceToHost));
        fill_stats(stats, ms, (double)g->E * n);
        if (stats) stats->iterations = c.policy;
    });
}

int sssp_autotune_batch(sssp_graph *g, const int32_t *sources, int32_t n, const char *tuning_file,
                        double *best_ms)
{
    return guard([&] {
        SSSP_REQUIRE(g && sources && n >= 1, "bad arguments");
        for (int32_t i = 0; i < n; i++)
            SSSP_REQUIRE(sources[i] >= 0 && sources[i] < g->V, "source vertex out of range");
        SSSP_HIP(hipSetDevice(g->device));
        const std::vector<int32_t> slots = source_slots(g, sources, n);
        const Config b = autotune([&](const Config &c) { return batch_sources(g, slots.data(), n, c); },
                                  tuning_file, best_ms, "batch of sources", true);
        g->batch_policy = b.policy;
        g->batch_bs = b.bs;
        g->batch_wgpcu = b.wgpcu;
        if (b.policy == 2) {
            g->push_edge_frac = b.frac;
            g->push_chunk = b.push_chunk;
        }
    });
}

int sssp_pack_create(const int32_t *src, const int32_t *dst, const float *weights,
                     const int64_t *edge_offsets, const int32_t *vertex_counts, int32_t n_graphs,
                     int32_t device, sssp_pack **out)
{
    sssp_pack *p = nullptr;
    int rc = guard([&] {
        SSSP_REQUIRE(out && edge_offsets && vertex_counts && n_graphs >= 1, "bad arguments");
        *out = nullptr;
        p = new sssp_pack();
        p->n = n_graphs;
        p->voff.assign((size_t)n_graphs + 1, 0);
        p->eoff.assign(edge_offsets, edge_offsets + n_graphs + 1);
        SSSP_REQUIRE(p->eoff[0] == 0, "edge_offsets must start at 0");
        for (int32_t k = 0; k < n_graphs; k++) {
            SSSP_REQUIRE(vertex_counts[k] >= 1, "every graph needs a vertex");
            SSSP_REQUIRE(p->eoff[k + 1] >= p->eoff[k], "edge_offsets must not decrease");
            p->voff[k + 1] = p->voff[k] + vertex_counts[k];
        }
        const int64_t V = p->voff[n_graphs], E = p->eoff[n_graphs];
        SSSP_REQUIRE(V < ((int64_t)1 << 31) - 1, "packed vertex count must fit int32");
        SSSP_REQUIRE(E == 0 || (src && dst && weights), "null edge array");
        std::vector<int32_t> us((size_t)E), ud((size_t)E);
        int64_t bad = 0;
#pragma omp parallel for schedule(dynamic, 16) reduction(+ : bad)
        for (int32_t k = 0; k < n_graphs; k++) {
            const int32_t nv = vertex_counts[k];
            const int32_t o = (int32_t)p->voff[k];
            for (int64_t e = p->eoff[k]; e < p->eoff[k + 1]; e++) {
                if (src[e] < 0 || src[e] >= nv || dst[e] < 0 || dst[e] >= nv) bad++;
                us[(size_t)e] = src[e] + o;
                ud[(size_t)e] = dst[e] + o;
            }
        }
        SSSP_REQUIRE(bad == 0, "pack: edge endpoint outside its graph");
        // the union keeps its ids: a graph's vertices stay one contiguous
        // range, which the workgroup-per-graph policy relies on
        const int rc2 = sssp_graph_create(SSSP_W_F32, us.data(), ud.data(), weights, (int32_t)V, E,
                                          device, SSSP_NO_RELABEL, &p->g);
        if (rc2 != SSSP_OK) throw Error{rc2, sssp_last_error()};
        *out = p;
    });
    if (rc != SSSP_OK && p) {
        if (p->g) sssp_graph_destroy(p->g);
        delete p;
    }
    return rc;
}

int sssp_pack_destroy(sssp_pack *p)
{
    return guard([&] {
        SSSP_REQUIRE(p, "null pack");
        if (p->g) sssp_graph_destroy(p->g);
        delete p;
    });
}

int sssp_pack_solve(sssp_pack *p, const int32_t *sources, int32_t policy, float *dist_out,
                    sssp_stats *stats)
{
    return guard([&] {
        SSSP_REQUIRE(p && p->g, "null pack");
        SSSP_REQUIRE(policy >= 0 && policy <= 2, "policy must be 0, 1 or 2");
        check_pack_sources(p, sources);
        SSSP_HIP(hipSetDevice(p->g->device));
        Config c{p->policy ? p->policy : 1, p->bs, p->wgpcu, p->frac, p->push_chunk};
        if (policy == 1) c = Config{1, p->policy == 1 ? p->bs : 1024, p->policy == 1 ? p->wgpcu : 1, 0.05, 1};
        if (policy == 2) c = Config{2, 256, 2, p->policy == 2 ? p->frac : 0.05, p->policy == 2 ? p->push_chunk : 1};
        const double ms = pack_run(p, sources, c);
        if (dist_out)
            SSSP_HIP(hipMemcpy(dist_out, p->g->bdist, 4 * (size_t)p->g->V, hipMemcpyDeviceToHost));
        fill_stats(stats, ms, (double)p->g->E);
        if (stats) stats->iterations = c.policy;
    });
}

int sssp_autotune_pack(sssp_pack *p, const int32_t *sources, const char *tuning_file,
                       double *best_ms)
{
    return guard([&] {
        SSSP_REQUIRE(p && p->g, "null pack");
        check_pack_sources(p, sources);
        SSSP_HIP(hipSetDevice(p->g->device));
        const Config b = autotune([&](const Config &c) { return pack_run(p, sources, c); },
                                  tuning_file, best_ms, "pack of graphs", false);
        p->policy = b.policy;
        p->bs = b.bs;
        p->wgpcu = b.wgpcu;
        p->frac = b.frac;
        p->push_chunk = b.push_chunk;
    });
}

}  // extern "C"
