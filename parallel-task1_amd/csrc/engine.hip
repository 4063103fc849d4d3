// engine.hip -- the relaxation kernels and the round loop of the MI355X
// Bellman-Ford SSSP.
//
// One solve = distance init + relaxation rounds until the frontier is empty
// and (by default) one full all-edge round confirms the fixpoint. Each round
// runs in one of three modes, chosen per round on the device:
//   PUSH  sparse frontier push: the frontier queue is compacted into entries
//         (vertex, row start, exclusive degree scan) and edges are dealt out
//         in fixed tiles of 256*K edges (merge-path style load balancing, so a
//         hub of degree 10^5 is spread over many workgroups);
//   ALL   all-edge sweep over the CSR of every vertex with out-degree > 0,
//         relaxing the out-edges of active vertices (stamped in the previous
//         or the current round, so updates propagate within the sweep);
//   PULL  for every unsettled vertex (d > min over the frontier, which no
//         future update can undercut) the min over its in-edges (CSC), with a
//         segmented wave reduction before one atomicMin per vertex.
// Relaxations use check-then-atomicMin on the 32-bit distance pattern. The
// frontier of the next round is built with wave-aggregated appends and an
// iteration stamp per vertex that deduplicates.
//
// Reference semantics followed (algorithm/sssp/gpu/bellman_ford.cu):
//   init:      d[i] = max, d[source] = 0                      (:10-20, :38-56)
//   relax:     if d[dst] > d[src] + w then d[dst] = d[src] + w  (:94-112)
//   converge:  stop after a pass that changes nothing          (:126-141)
// The plain schedule (SSSP_SCHED_PLAIN) is exactly that loop over the
// src-sorted edge list, with the convergence flag tested on the device.
#include <algorithm>
#include <atomic>
#include <cstring>
#include <mutex>

#include "mlp_dev.hpp"
#include "sssp_graph.hpp"

namespace sssp {

namespace {

enum { DIR_PUSH = 0, DIR_PULL = 1, DIR_BFS = 2 };
enum { SRC_SPARSE = 0, SRC_DENSE = 1 };
enum { PRED_NONE = 0, PRED_STAMP = 1, PRED_REACHED = 2, PRED_UNSETTLED = 3 };

constexpr int kLbItems = 8;
constexpr int kLbTile = kBlock * kLbItems;  // queue entries per prep block
constexpr int kSampleMax = 1 << 16;

struct AdvArgs {
    const int32_t *ent_v;      // entry vertex
    const int64_t *ent_start;  // entry row start (sparse entries)
    const int64_t *ent_scan;   // exclusive edge scan over entries, n+1 values
    const int32_t *tile;       // tile -> first entry, ntiles+1 values
    int32_t n;
    int64_t T;                 // total edges over entries
    int64_t ntiles;
    const int64_t *row_ptr;    // out-degree lookup at append time
    const int32_t *adj;        // el_dst (push) or csc_src (pull)
    const uint32_t *wts;       // el_w or csc_w
    uint32_t *dist;
    int32_t *stamp;
    int32_t *parent;           // certificate BFS
    int32_t tick_prev;
    int32_t tick_cur;
    int32_t *qnext;
    RoundSlot *slot;
    uint32_t minF;
    const int64_t *dev_n;      // when set, entry count and edge total are read
    const int64_t *dev_T;      // from the device (near-far compaction)
    const unsigned long long *dev_nt;  // when set, both from one word: count << 32 | total
                                       // (tail rounds)
    uint32_t *fbits;           // push: when set, lowered vertices are marked here (one bit
                               // each) instead of appended; k_bits_frontier queues them
};

constexpr int kStageCap = 512;  // per-wave LDS staging slots for frontier appends

// Per-block accumulators of one round, reduced once per block at the end.
struct RoundAcc {
    int64_t relaxed = 0, qedges = 0;
    int lowered = 0;
    int appended = 0;  // frontier entries not counted in the slot otherwise (dense rounds)
    uint32_t minw = 0xFFFFFFFFu;
};

template <int NW>
__device__ __forceinline__ void flush_acc_n(RoundAcc &acc, RoundSlot *s)
{
    __shared__ int64_t r_rel[NW], r_qe[NW];
    __shared__ int r_low[NW], r_app[NW];
    __shared__ uint32_t r_min[NW];
    const int lane = __lane_id(), w = threadIdx.x >> 6;
    const int64_t rel = wave_sum64(acc.relaxed), qe = wave_sum64(acc.qedges);
    const int low = wave_sum32(acc.lowered), app = wave_sum32(acc.appended);
    const uint32_t mn = wave_min_u32(acc.minw);
    if (lane == 0) {
        r_rel[w] = rel;
        r_qe[w] = qe;
        r_low[w] = low;
        r_app[w] = app;
        r_min[w] = mn;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int64_t a = 0, b = 0;
        int c = 0, q = 0;
        uint32_t m = 0xFFFFFFFFu;
        for (int i = 0; i < NW; i++) {
            a += r_rel[i];
            b += r_qe[i];
            c += r_low[i];
            q += r_app[i];
            m = r_min[i] < m ? r_min[i] : m;
        }
        if (a) atomicAdd((unsigned long long *)&s->relaxed, (unsigned long long)a);
        if (b) atomicAdd((unsigned long long *)&s->qedges, (unsigned long long)b);
        if (c) atomicAdd(&s->lowered, c);
        if (q) atomicAdd(&s->qlen, q);
        if (m != 0xFFFFFFFFu) atomicMin(&s->minw, m);
    }
}

__device__ __forceinline__ void flush_acc(RoundAcc &acc, RoundSlot *s)
{
    flush_acc_n<kBlock / 64>(acc, s);
}

// End-of-kernel flush of every wave's staged appends with one queue
// reservation per workgroup: a returning atomic on the one queue counter per
// wave, from every wave of a large grid at about the same time, queued at the
// counter's memory channel for tens of microseconds. Every thread of the
// workgroup calls it.
template <int NW>
__device__ __forceinline__ void stage_flush_block(int32_t (*stage)[kStageCap], int wcnt, int32_t *q,
                                                  int32_t *len)
{
    __shared__ int s_off[NW];
    __shared__ int s_base;
    const int w = threadIdx.x >> 6, lane = __lane_id();
    if (lane == 0) s_off[w] = wcnt;
    __syncthreads();
    if (threadIdx.x == 0) {
        int tot = 0;
        for (int i = 0; i < NW; i++) {
            const int c = s_off[i];
            s_off[i] = tot;
            tot += c;
        }
        s_base = tot > 0 ? atomicAdd(len, tot) : 0;
    }
    __syncthreads();
    const int base = s_base + s_off[w];
    for (int i = lane; i < wcnt; i += 64) q[base + i] = stage[w][i];
}

// A vertex lowered in this round joins the next frontier when it has
// out-edges and no other lane appended it in this round (stamp). A plain read
// of the stamp filters sinks (kSinkStamp) and vertices already appended, so
// the exchange is paid about once per appended vertex. This handles the N
// edges a lane relaxed in one step of its loop: the stamp and degree loads of
// all N are issued together, then the exchanges, so the protocol's latency is
// paid once per step rather than once per edge (the stamps still land within
// the step that lowered the vertex). Flushes the wave's staging buffer
// whenever fewer than 64 slots are left. (The certificate BFS, DIR_BFS,
// appends every vertex with out-edges it claimed, without stamps.)
template <int DIR, int N>
__device__ __forceinline__ void append_frontier_n(const bool (&low)[N], const int32_t (&v)[N],
                                                  const int64_t *row_ptr, int32_t *stamp,
                                                  int32_t tick_cur, int64_t &qedges,
                                                  int32_t *wbuf, int &wcnt, int32_t *q,
                                                  int32_t *qlen)
{
    int32_t sv[N];
    int64_t r0[N], r1[N];
#pragma unroll
    for (int k = 0; k < N; k++) {
        sv[k] = kSinkStamp;
        r0[k] = r1[k] = 0;
        if (low[k]) {
            if constexpr (DIR != DIR_BFS) sv[k] = stamp[v[k]];
            r0[k] = row_ptr[v[k]];
            r1[k] = row_ptr[v[k] + 1];
        }
    }
    // the exchanges are issued together and their results used after all
    // of them (a use between two made the compiler wait for each in turn)
    bool app[N];
    int32_t old[N];
#pragma unroll
    for (int k = 0; k < N; k++) {
        old[k] = tick_cur;
        if constexpr (DIR != DIR_BFS)
            if (low[k] && sv[k] != tick_cur && sv[k] != kSinkStamp)
                old[k] = atomicExch(&stamp[v[k]], tick_cur);
    }
    // pins each result in a register after the last exchange, so that the
    // compares below cannot be moved between the exchanges
#pragma unroll
    for (int k = 0; k < N; k++) asm volatile("" : "+v"(old[k]));
#pragma unroll
    for (int k = 0; k < N; k++) {
        if constexpr (DIR == DIR_BFS) app[k] = low[k] && r1[k] > r0[k];
        else app[k] = low[k] && old[k] != tick_cur;
    }
#pragma unroll
    for (int k = 0; k < N; k++) {
        if (app[k]) qedges += r1[k] - r0[k];
        stage_push(app[k], v[k], wbuf, wcnt);
        if (wcnt > kStageCap - 64) stage_flush(wbuf, wcnt, q, qlen);
    }
}

// Lowers dist[v] to cand without returning the old value. The caller has read
// dist[v] > cand in this kernel, and distances only decrease, so dist[v] was
// lowered during the kernel whoever wins: v belongs in the next frontier.
__device__ __forceinline__ void lower_dist(uint32_t *dist, int32_t v, uint32_t cand)
{
    __hip_atomic_fetch_min(&dist[v], cand, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Per-workgroup filter of lowering atomics. Device-scope atomics on the
// distance array execute past the L2, whose copy of the line keeps the value
// read before them, so every edge of a round into a much-reached vertex would
// send its own atomic. The filter keeps, per hashed vertex slot, the smallest
// value this workgroup has sent in this kernel: a candidate not below it
// cannot lower the distance (already, or once the earlier atomic lands, at
// most that value) and is dropped; the lane that sent the earlier value
// appended the vertex. Entries are (vertex << 32 | value), one 64-bit LDS word.
// Sizes (measured at R-MAT scale 25): 4096 entries (32 KB) in the plain pass,
// 2.98 -> 2.86 ms; 1024 in the round kernels, whose grids of up to 16
// workgroups per CU the larger table would no longer keep resident.
constexpr int kFiltBitsPass = 12;
constexpr int kFiltBitsRound = 10;

template <int BITS>
__device__ __forceinline__ void filt_init(uint64_t *f)
{
    for (int i = threadIdx.x; i < (1 << BITS); i += blockDim.x) f[i] = ~0ull;
}

// true: send the atomic (the slot now records cand for v)
template <int BITS>
__device__ __forceinline__ bool filt_send(uint64_t *f, int32_t v, uint32_t cand)
{
    const uint32_t h = ((uint32_t)v * 0x9E3779B1u) >> (32 - BITS);
    const uint64_t e = f[h];
    if ((uint32_t)(e >> 32) == (uint32_t)v && (uint32_t)e <= cand) return false;
    f[h] = ((uint64_t)(uint32_t)v << 32) | cand;
    return true;
}

template <class WT, int DIR, int SRC, int PRED, int K>
__global__ __launch_bounds__(kBlock) void k_advance(AdvArgs a)
{
    constexpr int TE = kBlock * K;
    constexpr uint32_t INF = WT::INF;
    __shared__ int32_t s_off[TE + 1];
    // entry vertex: pull and BFS rounds need it after the edge loads (push
    // rounds take the edge's own target)
    __shared__ int32_t s_v[DIR == DIR_PUSH ? 1 : TE + 1];
    __shared__ uint32_t s_d[TE + 1];
    __shared__ int64_t s_start[SRC == SRC_SPARSE ? TE + 1 : 1];
    __shared__ int32_t s_stage[kBlock / 64][kStageCap];
    __shared__ uint64_t s_filt[DIR == DIR_PUSH ? (1 << kFiltBitsRound) : 1];

    const int tid = threadIdx.x, lane = __lane_id();
    int32_t *wbuf = s_stage[tid >> 6];
    int wcnt = 0;
    RoundAcc acc;
    if constexpr (DIR == DIR_PUSH) filt_init<kFiltBitsRound>(s_filt);  // first use is after a barrier
    if (a.dev_n) {
        a.n = (int32_t)*a.dev_n;
        a.T = *a.dev_T;
        a.ntiles = (a.T + TE - 1) / TE;
    } else if (a.dev_nt) {
        const unsigned long long nt = *a.dev_nt;
        a.n = (int32_t)(nt >> 32);
        a.T = (int64_t)(nt & 0xFFFFFFFFull);
        a.ntiles = (a.T + TE - 1) / TE;
    }

    // XCD-aware tile order: groups of 16 consecutive tiles stay on one XCD
    // (blocks are dispatched round-robin over the 8 XCDs); groups interleave
    // across XCDs so skewed active regions still spread over the chip.
    const int64_t x = blockIdx.x & 7;
    const int64_t nbx = gridDim.x >> 3;
    constexpr int64_t G = 16;
    for (int64_t idx = blockIdx.x >> 3;; idx += nbx) {
        const int64_t t = ((idx / G) * 8 + x) * G + (idx % G);
        if (t >= a.ntiles) break;
        const int64_t e0 = t * TE;
        const int64_t e1 = min(e0 + (int64_t)TE, a.T);
        const int32_t j0 = a.tile[t];
        const int32_t jl = t + 1 < a.ntiles ? min(a.tile[t + 1], a.n - 1) : a.n - 1;
        const int m = jl - j0 + 1;
        int any = 0;
        for (int jj = tid; jj < m; jj += kBlock) {
            const int32_t j = j0 + jj;
            s_off[jj] = (int32_t)(a.ent_scan[j] - e0);
            const int32_t v = a.ent_v[j];
            if constexpr (DIR != DIR_PUSH) s_v[jj] = v;
            if constexpr (SRC == SRC_SPARSE) s_start[jj] = a.ent_start[j];
            const uint32_t dv = a.dist[v];
            bool act;
            if constexpr (DIR == DIR_PULL) {
                act = dv > a.minF;
                s_d[jj] = act ? dv : 0u;  // 0 marks "skip": no candidate is below 0
            } else {
                act = dv != INF;
                if constexpr (PRED == PRED_STAMP) act = act && (a.stamp[v] >= a.tick_prev);
                s_d[jj] = act ? dv : INF;
            }
            any |= act ? 1 : 0;
        }
        if (!__syncthreads_or(any)) continue;
        // Three stages so that the K edges of a thread have their loads in
        // flight together: (1) the entry of each edge (binary search in LDS,
        // non-decreasing in k), (2) the loads, (3) compare / lower / append.
        int jk[K];
        {
            int jj = 0;
#pragma unroll
            for (int k = 0; k < K; k++) {
                const int32_t rel = k * kBlock + tid;
                int lo = jj, hi = m - 1;  // largest jj' with s_off[jj'] <= rel
                while (lo < hi) {
                    const int mid = (lo + hi + 1) >> 1;
                    if (s_off[mid] <= rel) lo = mid;
                    else hi = mid - 1;
                }
                jj = lo;
                jk[k] = lo;
            }
        }
        uint32_t sk[K];  // push/bfs: d of the edge's source; pull: d of its target (0 = skip)
        int32_t ok[K];   // the edge's other endpoint
        uint32_t wk[K];
        uint32_t dk[K];  // d of the other endpoint
#pragma unroll
        for (int k = 0; k < K; k++) {
            const int32_t rel = k * kBlock + tid;
            const bool valid = e0 + rel < e1;
            if constexpr (DIR == DIR_PULL) sk[k] = valid ? s_d[jk[k]] : 0u;
            else sk[k] = valid ? s_d[jk[k]] : INF;
            const bool act = (DIR == DIR_PULL) ? sk[k] != 0u : sk[k] != INF;
            ok[k] = 0;
            wk[k] = 0;
            if (act) {
                const int64_t edge =
                    (SRC == SRC_DENSE) ? e0 + rel : s_start[jk[k]] + (rel - s_off[jk[k]]);
                ok[k] = a.adj[edge];
                wk[k] = a.wts[edge];
            }
        }
#pragma unroll
        for (int k = 0; k < K; k++) {
            const bool act = (DIR == DIR_PULL) ? sk[k] != 0u : sk[k] != INF;
            dk[k] = act ? a.dist[ok[k]] : INF;
        }
        // Every compare of the K edges is done before the first lowering
        // atomic is issued (they go out together below): an atomic issued
        // between two compares made the compiler wait for it -- a trip to the
        // memory side -- before the next compare, K times per tile.
        bool lowk[K];
        int32_t vk[K];
        uint32_t ck[K];  // the candidate to send when lowk (push, pull)
#pragma unroll
        for (int k = 0; k < K; k++) {
            const int32_t rel = k * kBlock + tid;
            const bool valid = e0 + rel < e1;
            const int jj = jk[k];
            bool low = false;
            int32_t v = 0;
            ck[k] = 0u;
            if constexpr (DIR == DIR_PULL) {
                const uint32_t dv = sk[k];
                uint32_t cand = INF;
                if (dv != 0u) {
                    cand = (dk[k] == INF) ? INF : WT::add(dk[k], wk[k]);
                    acc.relaxed++;
                }
                // segmented min over lanes sharing an entry (jj is non-decreasing in lane)
                const int seg = valid ? jj : (0x40000000 | lane);
#pragma unroll
                for (int o = 1; o < 64; o <<= 1) {
                    const uint32_t oc = __shfl_down(cand, o);
                    const int os = __shfl_down(seg, o);
                    if (lane + o < 64 && os == seg) cand = min(cand, oc);
                }
                const int pseg = __shfl_up(seg, 1);
                const bool head = (lane == 0) || (pseg != seg);
                if (head && dv != 0u && cand < dv) {
                    v = s_v[jj];
                    ck[k] = cand;
                    low = true;
                    acc.minw = min(acc.minw, cand);
                    acc.lowered++;
                }
            } else {
                const uint32_t du = sk[k];
                if (du != INF) {
                    v = ok[k];
                    const uint32_t cand = WT::add(du, wk[k]);
                    acc.relaxed++;
                    const uint32_t dv = dk[k];
                    if constexpr (DIR == DIR_BFS) {
                        // certificate: tight edge into a finite vertex, first claimer wins
                        if (dv != INF && cand == dv && atomicCAS(&a.parent[v], -1, s_v[jj]) == -1)
                            low = true;
                    } else if (cand < dv && filt_send<kFiltBitsRound>(s_filt, v, cand)) {
                        ck[k] = cand;
                        low = true;
                        acc.minw = min(acc.minw, cand);
                        acc.lowered++;
                    }
                }
            }
            lowk[k] = low;
            vk[k] = v;
        }
        if constexpr (DIR != DIR_BFS) {
#pragma unroll
            for (int k = 0; k < K; k++)
                if (lowk[k]) lower_dist(a.dist, vk[k], ck[k]);
        }
        if (DIR == DIR_PUSH && a.fbits) {
            // one non-returning OR per lowered vertex: no stamp read, no
            // degree loads, no returning exchange here (k_bits_frontier)
#pragma unroll
            for (int k = 0; k < K; k++)
                if (lowk[k])
                    __hip_atomic_fetch_or(&a.fbits[vk[k] >> 5], 1u << (vk[k] & 31),
                                          __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            append_frontier_n<DIR, K>(lowk, vk, a.row_ptr, a.stamp, a.tick_cur, acc.qedges,
                                      wbuf, wcnt, a.qnext, &a.slot->qlen);
        }
        __syncthreads();
    }
    stage_flush_block<kBlock / 64>(s_stage, wcnt, a.qnext, &a.slot->qlen);
    flush_acc(acc, a.slot);
}

// The next frontier of a push round that marked its lowered vertices in a
// bitmap (AdvArgs::fbits): every vertex with out-edges whose bit is set joins
// the queue (in vertex order) with the round's stamp, and its bit is
// cleared. A wave takes 8 chunks of 64 vertices per step (two bitmap words
// each): the 8 words, then the row offsets of the set bits, are loaded
// together; chunks without a set bit cost only their word load.
__global__ __launch_bounds__(kBlock) void k_bits_frontier(uint32_t *fbits, int64_t V,
                                                          const int64_t *row_ptr, int32_t *stamp,
                                                          int32_t tick, int32_t *q,
                                                          RoundSlot *slot)
{
    constexpr int J = 8;
    __shared__ int32_t s_stage[kBlock / 64][kStageCap];
    const int tid = threadIdx.x, lane = __lane_id();
    int32_t *wbuf = s_stage[tid >> 6];
    int wcnt = 0;
    RoundAcc acc;
    const int64_t nch = (V + 63) >> 6;  // the bitmap holds 2 * nch words
    const int64_t ngroups = (nch + J - 1) / J;
    const int64_t nw = (int64_t)gridDim.x * (kBlock / 64);
    for (int64_t gi = (int64_t)blockIdx.x * (kBlock / 64) + (tid >> 6); gi < ngroups; gi += nw) {
        uint32_t wd[J];
#pragma unroll
        for (int j = 0; j < J; j++) {
            const int64_t c = gi * J + j;
            wd[j] = c < nch ? fbits[c * 2 + (lane >> 5)] : 0u;
        }
        bool set[J];
        int64_t r0[J], r1[J];
#pragma unroll
        for (int j = 0; j < J; j++) {
            const int64_t v = (gi * J + j) * 64 + lane;
            set[j] = v < V && ((wd[j] >> (lane & 31)) & 1u);
            r0[j] = r1[j] = 0;
            if (set[j]) {
                r0[j] = row_ptr[v];
                r1[j] = row_ptr[v + 1];
            }
        }
#pragma unroll
        for (int j = 0; j < J; j++) {
            const int64_t c = gi * J + j;
            const int32_t v = (int32_t)(c * 64 + lane);
            const int64_t deg = r1[j] - r0[j];
            const bool app = set[j] && deg > 0;
            if (app) {
                stamp[v] = tick;
                acc.qedges += deg;
            }
            stage_push(app, v, wbuf, wcnt);
            if (wcnt > kStageCap - 64) stage_flush(wbuf, wcnt, q, &slot->qlen);
            if ((lane & 31) == 0 && wd[j] != 0u) fbits[c * 2 + (lane >> 5)] = 0u;
        }
    }
    stage_flush_block<kBlock / 64>(s_stage, wcnt, q, &slot->qlen);
    flush_acc(acc, slot);
}

// ---------------------------------------------------------------------------
// All-edge round over the src-sorted edge list: the plain pass's streaming
// access (16-byte loads, 4 edges per lane) restricted to active sources
// (PRED_STAMP: stamped in the previous or the current round; PRED_REACHED:
// any finite distance -- the fixpoint check), with frontier appends.
// ---------------------------------------------------------------------------
template <class WT, int PRED>
__global__ __launch_bounds__(kBlock) void k_el_active(const int32_t *__restrict__ src,
                                                      const int32_t *__restrict__ dst,
                                                      const uint32_t *__restrict__ w, int64_t E,
                                                      uint32_t *dist, int32_t *stamp,
                                                      int32_t tick_prev, int32_t tick_cur,
                                                      const int64_t *row_ptr, int32_t *qnext,
                                                      RoundSlot *slot, const int64_t *goff)
{
    typedef int32_t v4i __attribute__((ext_vector_type(4)));
    typedef uint32_t v4u __attribute__((ext_vector_type(4)));
    constexpr uint32_t INF = WT::INF;
    __shared__ int32_t s_stage[kBlock / 64][kStageCap];
    __shared__ uint64_t s_filt[1 << kFiltBitsRound];
    int32_t *wbuf = s_stage[threadIdx.x >> 6];
    int wcnt = 0;
    RoundAcc acc;
    filt_init<kFiltBitsRound>(s_filt);
    __syncthreads();
    // XCD-partitioned list (see k_el_pass): group blockIdx % 8, padded to
    // whole 16-byte vectors
    const int64_t x = blockIdx.x & 7, nbx = gridDim.x >> 3, bi = blockIdx.x >> 3;
    const int64_t ga = goff[x] >> 2, gb = goff[x + 1] >> 2;
    const int64_t per = (gb - ga + nbx - 1) / nbx;
    const int64_t g0 = ga + bi * per, g1 = min(g0 + per, gb);
    const int64_t iters = g1 > g0 ? (g1 - g0 + kBlock - 1) / kBlock : 0;
    const v4i *s4 = reinterpret_cast<const v4i *>(src);
    const v4i *t4 = reinterpret_cast<const v4i *>(dst);
    const v4u *w4 = reinterpret_cast<const v4u *>(w);
    for (int64_t it = 0; it < iters; it++) {
        int32_t es[4], et[4];
        uint32_t ew[4];
        bool val[4] = {false, false, false, false};
        {
            const int64_t g = g0 + it * kBlock + threadIdx.x;
            if (g < g1) {
                const v4i sv = __builtin_nontemporal_load(&s4[g]);
                es[0] = sv.x; es[1] = sv.y; es[2] = sv.z; es[3] = sv.w;
                val[0] = val[1] = val[2] = val[3] = true;
            }
        }
        uint32_t du[4];
        bool any = false;
        // the distances and stamps of the four sources are loaded together
        // (equal sources of the src-sorted runs hit the same line): a reuse
        // test between them serialised the loads (rounds 20-30 % slower)
        uint32_t dd[4];
        int32_t sg[4];
#pragma unroll
        for (int i = 0; i < 4; i++) {
            dd[i] = val[i] ? dist[es[i]] : INF;
            if constexpr (PRED == PRED_STAMP) sg[i] = val[i] ? stamp[es[i]] : 0;
        }
#pragma unroll
        for (int i = 0; i < 4; i++) {
            bool act = dd[i] != INF;
            if constexpr (PRED == PRED_STAMP) act = act && sg[i] >= tick_prev;
            du[i] = act ? dd[i] : INF;
            any |= du[i] != INF;
        }
        uint32_t dv[4];
        if (any) {
            const int64_t g = g0 + it * kBlock + threadIdx.x;
            const v4i tv = __builtin_nontemporal_load(&t4[g]);
            const v4u wv = __builtin_nontemporal_load(&w4[g]);
            et[0] = tv.x; et[1] = tv.y; et[2] = tv.z; et[3] = tv.w;
            ew[0] = wv.x; ew[1] = wv.y; ew[2] = wv.z; ew[3] = wv.w;
#pragma unroll
            for (int i = 0; i < 4; i++) dv[i] = du[i] != INF ? dist[et[i]] : 0u;
        }
        bool low[4];
        int32_t vv[4];
        uint32_t ck[4];
#pragma unroll
        for (int i = 0; i < 4; i++) {
            low[i] = false;
            vv[i] = 0;
            ck[i] = 0u;
            if (du[i] != INF) {
                vv[i] = et[i];
                const uint32_t cand = WT::add(du[i], ew[i]);
                acc.relaxed++;
                if (cand < dv[i] && filt_send<kFiltBitsRound>(s_filt, vv[i], cand)) {
                    ck[i] = cand;
                    low[i] = true;
                    acc.minw = min(acc.minw, cand);
                    acc.lowered++;
                }
            }
        }
        // the lowering atomics after every compare (see k_advance)
#pragma unroll
        for (int i = 0; i < 4; i++)
            if (low[i]) lower_dist(dist, vv[i], ck[i]);
        append_frontier_n<DIR_PUSH, 4>(low, vv, row_ptr, stamp, tick_cur, acc.qedges, wbuf, wcnt,
                                       qnext, &slot->qlen);
    }
    stage_flush_block<kBlock / 64>(s_stage, wcnt, qnext, &slot->qlen);
    flush_acc(acc, slot);
}

// ---------------------------------------------------------------------------
// Sparse-frontier preparation with the far pile. A queue entry is "near"
// (relaxed in this round) when its out-degree is below hub_degree or its
// distance is below thr; otherwise it moves to the next far pile and waits
// until its distance comes within delta of the frontier minimum. Near
// entries become load-balancing entries (vertex, row start, exclusive degree
// scan, tile table). Entries still in the pile are skipped (the pile holds
// them); entries the pile filter moved out this round are near.
// thr = 0xFFFFFFFF with no pile (farst == nullptr): every entry is near.
// ---------------------------------------------------------------------------
struct NearRule {
    uint32_t thr;
    int32_t hub_degree;
    const int32_t *farst;  // nullptr: no pile
    int32_t tick_cur;
};

enum { kSkip = 0, kNear = 1, kFar = 2 };

__device__ __forceinline__ int classify(const NearRule &r, int32_t v, uint32_t d, int64_t deg)
{
    if (!r.farst) return kNear;
    const int32_t f = r.farst[v];
    if (f == kPileIn) return kSkip;
    if (f == r.tick_cur) return kNear;
    return (d < r.thr || deg < r.hub_degree) ? kNear : kFar;
}

__global__ __launch_bounds__(kBlock) void k_lb_reduce(const int32_t *q, int32_t n,
                                                      const int32_t *dev_n, const int64_t *row_ptr,
                                                      const uint32_t *dist, NearRule rule,
                                                      int64_t *blk_deg, int64_t *blk_cnt)
{
    __shared__ int64_t s_warp[kBlock / 64];
    if (dev_n) n = *dev_n;
    const int64_t base = (int64_t)blockIdx.x * kLbTile + (int64_t)threadIdx.x * kLbItems;
    int64_t s = 0, c = 0;
#pragma unroll
    for (int i = 0; i < kLbItems; i++) {
        const int64_t j = base + i;
        if (j < n) {
            const int32_t v = q[j];
            const int64_t deg = row_ptr[v + 1] - row_ptr[v];
            if (classify(rule, v, dist[v], deg) == kNear) {
                s += deg;
                c++;
            }
        }
    }
    int64_t tot_s, tot_c;
    (void)block_excl_scan_i64(s, &tot_s, s_warp);
    (void)block_excl_scan_i64(c, &tot_c, s_warp);
    if (threadIdx.x == 0) {
        blk_deg[blockIdx.x] = tot_s;
        blk_cnt[blockIdx.x] = tot_c;
    }
}

struct LbWriteArgs {
    const int32_t *q;
    int32_t n;
    const int32_t *dev_n;    // when set, the entry count is read from the device
    const int64_t *row_ptr;
    const uint32_t *dist;
    NearRule rule;
    const int64_t *blk_deg;  // scanned, [nb] = total edges
    const int64_t *blk_cnt;  // scanned, [nb] = near entries
    int64_t nb;
    int32_t *lv;
    int64_t *ls;
    int64_t *lscan;
    int32_t *ltile;
    int64_t TE;
    Ctrl *ctrl;              // receives n_compact / lb_total and the next pile's counters
    int32_t *farst;          // far entries are marked kPileIn
    int32_t *pile_next;
    int nxt;                 // index of the next pile's counters
};

__global__ __launch_bounds__(kBlock) void k_lb_write(LbWriteArgs a)
{
    __shared__ int64_t s_warp[kBlock / 64];
    const int32_t n = a.dev_n ? *a.dev_n : a.n;
    const int64_t base = (int64_t)blockIdx.x * kLbTile + (int64_t)threadIdx.x * kLbItems;
    int32_t vv[kLbItems];
    uint32_t dd[kLbItems];
    int64_t st[kLbItems], dg[kLbItems];
    int cls[kLbItems];
    int64_t s = 0, c = 0;
#pragma unroll
    for (int i = 0; i < kLbItems; i++) {
        const int64_t j = base + i;
        vv[i] = 0;
        dd[i] = 0xFFFFFFFFu;
        st[i] = 0;
        dg[i] = 0;
        cls[i] = kSkip;
        if (j < n) {
            vv[i] = a.q[j];
            dd[i] = a.dist[vv[i]];
            st[i] = a.row_ptr[vv[i]];
            dg[i] = a.row_ptr[vv[i] + 1] - st[i];
            cls[i] = classify(a.rule, vv[i], dd[i], dg[i]);
        }
        if (cls[i] == kNear) {
            s += dg[i];
            c++;
        }
    }
    const int64_t total_edges = a.blk_deg[a.nb], total_near = a.blk_cnt[a.nb];
    int64_t tot;
    int64_t off = block_excl_scan_i64(s, &tot, s_warp) + a.blk_deg[blockIdx.x];
    int64_t idx = block_excl_scan_i64(c, &tot, s_warp) + a.blk_cnt[blockIdx.x];
    const int64_t ntiles = (total_edges + a.TE - 1) / a.TE;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        a.ctrl->n_compact = total_near;
        a.ctrl->lb_total = total_edges;
        if (total_near == 0) a.ltile[0] = 0;
    }
    int64_t far_edges = 0;
    uint32_t far_min = 0xFFFFFFFFu;
#pragma unroll
    for (int i = 0; i < kLbItems; i++) {
        if (cls[i] == kNear) {
            const int64_t j = idx;
            a.lv[j] = vv[i];
            a.ls[j] = st[i];
            a.lscan[j] = off;
            for (int64_t t = (off + a.TE - 1) / a.TE; t * a.TE < off + dg[i]; t++)
                a.ltile[t] = (int32_t)j;
            off += dg[i];
            idx++;
            if (j == total_near - 1) {
                a.lscan[total_near] = off;
                a.ltile[ntiles] = (int32_t)total_near;
            }
        }
        const bool far = cls[i] == kFar;
        if (far) {
            a.farst[vv[i]] = kPileIn;
            far_edges += dg[i];
            far_min = min(far_min, dd[i]);
        }
        wave_append(far, vv[i], a.pile_next, &a.ctrl->pile_n[a.nxt]);
    }
    far_edges = wave_sum64(far_edges);
    far_min = wave_min_u32(far_min);
    if (__lane_id() == 0) {
        if (far_edges)
            atomicAdd((unsigned long long *)&a.ctrl->pile_edges[a.nxt], (unsigned long long)far_edges);
        if (far_min != 0xFFFFFFFFu) atomicMin(&a.ctrl->pile_min[a.nxt], far_min);
    }
}

// Pile filter of a push round: pile entries whose distance is below thr move
// out (state = this round's tick) and join the queue unless they are on it
// already (stamped by the previous round); the others go to the next pile.
__global__ __launch_bounds__(kBlock) void k_pile_filter(const int32_t *pile, int32_t n,
                                                        const uint32_t *dist, uint32_t thr,
                                                        int32_t *farst, const int32_t *stamp,
                                                        int32_t tick_prev, int32_t tick_cur,
                                                        int32_t *q, const int64_t *row_ptr,
                                                        int32_t *pile_next, Ctrl *c, int nxt)
{
    int64_t kept_edges = 0;
    uint32_t kept_min = 0xFFFFFFFFu;
    for (int64_t base = (int64_t)blockIdx.x * kBlock; base < n;
         base += (int64_t)gridDim.x * kBlock) {
        const int64_t i = base + threadIdx.x;
        bool mv = false, keep = false;
        int32_t v = 0;
        if (i < n) {
            v = pile[i];
            const uint32_t d = dist[v];
            if (d < thr) {
                farst[v] = tick_cur;
                mv = stamp[v] != tick_prev;
            } else {
                keep = true;
                kept_edges += row_ptr[v + 1] - row_ptr[v];
                kept_min = min(kept_min, d);
            }
        }
        wave_append(mv, v, q, &c->lb_n);
        wave_append(keep, v, pile_next, &c->pile_n[nxt]);
    }
    kept_edges = wave_sum64(kept_edges);
    kept_min = wave_min_u32(kept_min);
    if (__lane_id() == 0) {
        if (kept_edges)
            atomicAdd((unsigned long long *)&c->pile_edges[nxt], (unsigned long long)kept_edges);
        if (kept_min != 0xFFFFFFFFu) atomicMin(&c->pile_min[nxt], kept_min);
    }
}

// Before an all-edge or pull round: every piled vertex becomes active for
// this round (stamp of the consumed frontier) and leaves the pile.
__global__ __launch_bounds__(kBlock) void k_pile_flush(const int32_t *pile, int32_t n,
                                                       int32_t *stamp, int32_t tick_prev,
                                                       int32_t *farst, Ctrl *c, int cur)
{
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * kBlock) {
        const int32_t v = pile[i];
        stamp[v] = tick_prev;
        farst[v] = 0;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        c->pile_n[cur] = 0;
        c->pile_min[cur] = 0xFFFFFFFFu;
        c->pile_edges[cur] = 0;
    }
}

// ---------------------------------------------------------------------------
// Init, round bookkeeping and the per-round decision.
// ---------------------------------------------------------------------------
template <class T>
__global__ __launch_bounds__(kBlock) void k_fill(T *p, int64_t n, T val)
{
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * kBlock)
        p[i] = val;
}

__global__ __launch_bounds__(kBlock) void k_stamp_init(int32_t *stamp, const int64_t *row_ptr,
                                                       int64_t V)
{
    for (int64_t v = (int64_t)blockIdx.x * kBlock + threadIdx.x; v < V;
         v += (int64_t)gridDim.x * kBlock)
        stamp[v] = row_ptr[v + 1] > row_ptr[v] ? 0 : kSinkStamp;
}

template <class WT>
__global__ __launch_bounds__(kBlock) void k_count_reached(const uint32_t *dist, int64_t V,
                                                          int64_t *out)
{
    int64_t n = 0;
    for (int64_t v = (int64_t)blockIdx.x * kBlock + threadIdx.x; v < V;
         v += (int64_t)gridDim.x * kBlock)
        n += dist[v] != WT::INF;
    n = wave_sum64(n);
    if (__lane_id() == 0 && n) atomicAdd((unsigned long long *)out, (unsigned long long)n);
}

__device__ void zero_slot(RoundSlot *s, uint32_t minw)
{
    s->qlen = 0;
    s->minw = minw;
    s->qedges = 0;
    s->relaxed = 0;
    s->lowered = 0;
    s->sample_unreached = 0;
    s->sample_unsettled = 0;
    s->sample_n = 0;
    s->sample_unsettled_in = 0;
}

// source < 0: only the control block is reset (k_init_sources seeds the queue)
__global__ void k_init_solve(Ctrl *c, uint32_t *dist, int32_t *stamp, int32_t *q0,
                             const int64_t *row_ptr, int32_t source, int32_t tick0)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    zero_slot(&c->slot[0], 0u);
    zero_slot(&c->slot[1], 0xFFFFFFFFu);
    for (int i = 0; i < 64; i++) c->changed[i] = 0;
    for (int i = 0; i < 8; i++) c->cert[i] = 0;
    c->n_reached = 0;
    c->err = 0;
    for (int i = 0; i < 2; i++) {
        c->pile_n[i] = 0;
        c->pile_min[i] = 0xFFFFFFFFu;
        c->pile_edges[i] = 0;
    }
    c->lb_n = 0;
    c->tail_stop = -1;
    c->tail_rounds = 0;
    c->tail_relaxed = 0;
    c->tail_list = 0;
    if (source < 0) return;
    dist[source] = 0u;
    const int64_t deg = row_ptr[source + 1] - row_ptr[source];
    c->slot[0].qlen = deg > 0 ? 1 : 0;
    c->slot[0].qedges = deg;
    if (deg > 0) {
        q0[0] = source;
        stamp[source] = tick0;
    }
}

// Several sources at distance 0 (a batch of disjoint graphs solved as one
// union graph): every source with out-edges starts in the first frontier.
__global__ __launch_bounds__(kBlock) void k_init_sources(Ctrl *c, uint32_t *dist, int32_t *stamp,
                                                         int32_t *q0, const int64_t *row_ptr,
                                                         const int32_t *srcs, int32_t n,
                                                         int32_t tick0)
{
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const int32_t s = srcs[i];
    dist[s] = 0u;
    const int64_t deg = row_ptr[s + 1] - row_ptr[s];
    if (deg > 0 && atomicExch(&stamp[s], tick0) != tick0) {
        q0[atomicAdd(&c->slot[0].qlen, 1)] = s;
        atomicAdd((unsigned long long *)&c->slot[0].qedges, (unsigned long long)deg);
    }
}

// Zeroes the slot the coming round fills and samples the unsettled fraction
// of the frontier it consumes (d > minF, minF = min distance over the frontier
// and, pile >= 0, over the far pile).
// Round bookkeeping: resets the next round's slot and samples the unsettled
// (distance above the frontier minimum) and unreached fractions of a fixed
// vertex sample, with the in-degree of the unsettled ones (pull estimate).
__device__ __forceinline__ void round_sample(Ctrl *c, int32_t r, int pile,
                                             const int32_t *sample_v, const int32_t *sample_in,
                                             int32_t n_sample, const uint32_t *dist, uint32_t inf)
{
    __shared__ int s_cnt[kBlock / 64], s_unr[kBlock / 64];
    __shared__ int64_t s_ind[kBlock / 64];
    RoundSlot *cur = &c->slot[r & 1];
    RoundSlot *nxt = &c->slot[(r + 1) & 1];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        zero_slot(nxt, 0xFFFFFFFFu);
        cur->sample_n = n_sample;
    }
    const uint32_t minF = pile >= 0 ? min(cur->minw, c->pile_min[pile]) : cur->minw;
    int cnt = 0, unr = 0;
    int64_t ind = 0;
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n_sample;
         i += (int64_t)gridDim.x * kBlock) {
        const int32_t v = sample_v[i];
        const int32_t din = sample_in ? sample_in[i] : 0;  // loaded beside v
        const uint32_t d = dist[v];
        if (d > minF) {
            cnt++;
            ind += din;
        }
        unr += d == inf ? 1 : 0;
    }
    // one set of counter atomics per workgroup
    cnt = wave_sum32(cnt);
    unr = wave_sum32(unr);
    ind = wave_sum64(ind);
    const int w = threadIdx.x >> 6;
    if (__lane_id() == 0) {
        s_cnt[w] = cnt;
        s_unr[w] = unr;
        s_ind[w] = ind;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        cnt = unr = 0;
        ind = 0;
        for (int i = 0; i < kBlock / 64; i++) {
            cnt += s_cnt[i];
            unr += s_unr[i];
            ind += s_ind[i];
        }
        if (cnt) atomicAdd(&cur->sample_unsettled, cnt);
        if (unr) atomicAdd(&cur->sample_unreached, unr);
        if (ind) atomicAdd((unsigned long long *)&cur->sample_unsettled_in, (unsigned long long)ind);
    }
}

__global__ __launch_bounds__(kBlock) void k_round_begin(Ctrl *c, int32_t r, int pile,
                                                        const int32_t *sample_v,
                                                        const int32_t *sample_in, int32_t n_sample,
                                                        const uint32_t *dist, uint32_t inf)
{
    round_sample(c, r, pile, sample_v, sample_in, n_sample, dist, inf);
}

// In-degree of each sample vertex (the pull estimate's weight), once per graph.
__global__ __launch_bounds__(kBlock) void k_sample_in(const int32_t *sample_v, int32_t n,
                                                      const int64_t *col_ptr, int32_t *out)
{
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const int32_t v = sample_v[i];
    const int64_t d = col_ptr[v + 1] - col_ptr[v];
    out[i] = (int32_t)min(d, (int64_t)0x7FFFFFFF);
}

struct DecideArgs {
    int32_t r;
    int32_t sched;
    double V, E;
    int has_csc;
    float push_edge_frac;
    float pull_frac;
    int push_chunk, all_chunk, pull_chunk;
    const float *mlp;  // nullable
    int pile;          // pile buffer the round reads
    Ctrl *mirror;      // device address of the host mirror (coherent host memory)
    int32_t seq;       // written to mirror->dec_seq once the mirror holds this round
};

// Features, the hand-written rule and (learned schedule) the MLP, computed
// by threads 0-63 of one workgroup (every thread of it calls this: it holds
// workgroup barriers); writes the decision and the host mirror.
__device__ __forceinline__ void decide_round(Ctrl *c, const DecideArgs &d)
{
    static_assert(kHid == 64, "one hidden unit per thread of the first wave");
    __shared__ float s_f[kIn];
    __shared__ float s_h1[kHid];
    __shared__ float s_h2[kHid];
    __shared__ float s_o[kOut];
    const int j = threadIdx.x;
    const RoundSlot *cur = &c->slot[d.r & 1];
    // the frontier is the queue plus the far pile
    const double fv = (double)cur->qlen + c->pile_n[d.pile];
    const double fe = (double)cur->qedges + (double)c->pile_edges[d.pile];
    const double sn = cur->sample_n > 0 ? cur->sample_n : 1;
    const double uns = cur->sample_unsettled / sn;
    const double uns_in = (double)cur->sample_unsettled_in * (d.V / sn);
    if (j == 0) {
        float f[kIn];
        make_features(fv, fe, uns, d.r, d.V > 0 ? d.E / d.V : 0.0, d.V, d.E, f);
        for (int i = 0; i < kIn; i++) s_f[i] = f[i];
    }
    __syncthreads();
    // hand-written rule: sparse push while the frontier's out-edges are a
    // small fraction of E; pull when the unsettled in-edges are fewer than
    // the frontier's out-edges; otherwise the all-edge sweep
    int hmode, hchunk;
    if (fe <= d.push_edge_frac * d.E) {
        hmode = SSSP_MODE_PUSH;
        hchunk = d.push_chunk;
    } else if (d.has_csc && uns_in < d.pull_frac * fe) {
        hmode = SSSP_MODE_PULL;
        hchunk = d.pull_chunk;
    } else {
        hmode = SSSP_MODE_ALL;
        hchunk = d.all_chunk;
    }
    int mode = hmode, chunk = hchunk, by_mlp = 0;
    if (d.mlp && d.sched == SSSP_SCHED_LEARNED) {
        const float *P = d.mlp;
        if (j < kHid) {
            float acc = P[kOffB1 + j];
            for (int i = 0; i < kIn; i++) acc += P[kOffW1 + j * kIn + i] * s_f[i];
            s_h1[j] = acc > 0.f ? acc : 0.f;
        }
        __syncthreads();
        if (j < kHid) {
            float acc = P[kOffB2 + j];
            for (int i = 0; i < kHid; i++) acc += P[kOffW2 + j * kHid + i] * s_h1[i];
            s_h2[j] = acc > 0.f ? acc : 0.f;
        }
        __syncthreads();
        if (j < kOut) {
            float acc = P[kOffB3 + j];
            for (int i = 0; i < kHid; i++) acc += P[kOffW3 + j * kHid + i] * s_h2[i];
            s_o[j] = acc;
        }
        __syncthreads();
        bool ok = true;
        for (int i = 0; i < kOut; i++) ok = ok && isfinite(s_o[i]);
        if (ok) {
            int bm = 0;
            for (int i = 1; i < kModes; i++) {
                if (i == SSSP_MODE_PULL && !d.has_csc) continue;
                if (s_o[i] > s_o[bm]) bm = i;
            }
            int bc = 0;
            for (int i = 1; i < SSSP_MLP_CHUNKS; i++)
                if (s_o[kModes + i] > s_o[kModes + bc]) bc = i;
            mode = bm;
            chunk = bc;
            by_mlp = 1;
        }
    }
    if (j == 0) {
        // counters a push round fills: the next pile and the queue length
        // after the pile filter
        const int nx = d.pile ^ 1;
        c->pile_n[nx] = 0;
        c->pile_min[nx] = 0xFFFFFFFFu;
        c->pile_edges[nx] = 0;
        c->lb_n = cur->qlen;
        c->dec_mode = mode;
        c->dec_chunk = chunk;
        c->dec_mlp = by_mlp;
        c->dec_heur_mode = hmode;
        for (int i = 0; i < kIn; i++) c->feat[i] = s_f[i];
        for (int i = 0; i < kOut; i++) c->logits[i] = by_mlp ? s_o[i] : 0.f;
        // what the host reads to launch the round, written straight into
        // host memory (visible once the kernel has completed)
        Ctrl *m = d.mirror;
        m->slot[d.r & 1] = *cur;
        m->pile_n[d.pile] = c->pile_n[d.pile];
        m->pile_min[d.pile] = c->pile_min[d.pile];
        m->pile_edges[d.pile] = c->pile_edges[d.pile];
        m->lb_n = cur->qlen;
        m->dec_mode = mode;
        m->dec_chunk = chunk;
        m->dec_mlp = by_mlp;
        m->dec_heur_mode = hmode;
        for (int i = 0; i < kIn; i++) m->feat[i] = s_f[i];
        for (int i = 0; i < kOut; i++) m->logits[i] = by_mlp ? s_o[i] : 0.f;
        // the host polls dec_seq instead of waiting for the stream: every
        // field above is visible to it before the sequence number
        __threadfence_system();
        __hip_atomic_store(&m->dec_seq, d.seq, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    }
}

// Round bookkeeping and decision in one launch: every workgroup samples
// (round_sample), then counts itself done; the last one to finish -- which
// sees every other workgroup's sample counts through the fences around the
// counter -- resets the counter and decides.
__global__ __launch_bounds__(kBlock) void k_round_open(Ctrl *c, int32_t r, int pile,
                                                       const int32_t *sample_v,
                                                       const int32_t *sample_in, int32_t n_sample,
                                                       const uint32_t *dist, uint32_t inf,
                                                       DecideArgs d)
{
    __shared__ int s_last;
    round_sample(c, r, pile, sample_v, sample_in, n_sample, dist, inf);
    __threadfence();
    __syncthreads();
    if (threadIdx.x == 0) s_last = atomicAdd(&c->open_done, 1) == (int)gridDim.x - 1;
    __syncthreads();
    if (!s_last) return;  // workgroup-uniform
    __threadfence();
    if (threadIdx.x == 0) c->open_done = 0;
    decide_round(c, d);
}

// ---------------------------------------------------------------------------
// Tail rounds: push rounds over a small frontier, queued several at a time
// with no decision between them. Once the frontier has shrunk to few
// out-edges, the rounds left are short (tens of microseconds of work), and
// each decided round would also pay its decision launch, the host's round
// trip and three load-balancing launches. A tail round is three launches that
// read their frontier from the device: k_tail_list builds the load-balancing
// entries in one pass, k_advance (the push rounds' kernel) relaxes, and
// k_tail_close (one thread) does the bookkeeping. A tail round runs only while
// the rounds before it ran and its frontier is non-empty with at most
// tail_edges out-edges; the first one that does not is where the host's
// decided rounds resume (tail_stop).
// ---------------------------------------------------------------------------
constexpr int kTailChunk = 2;  // chunk bucket of the tail rounds' relaxation (4 edges per lane)
// ... and of a run of tail rounds whose first frontier has more out-edges
// than kTailBigEdges (8 edges per lane, the decided push rounds' grid)
constexpr int kTailBigChunk = 3;
constexpr int64_t kTailBigEdges = (int64_t)1 << 22;

struct TailArgs {
    Ctrl *ctrl;
    Ctrl *mirror;          // device address of the host mirror (coherent host memory)
    const int64_t *row_ptr;
    const int32_t *adj;    // CSR targets (src-sorted edge list)
    const uint32_t *wts;
    uint32_t *dist;
    int32_t *stamp;
    const int32_t *q;      // frontier this round consumes
    int32_t *qnext;        // frontier it builds
    int64_t tail_edges;    // out-edge bound of a frontier a tail round takes
    int32_t r;             // round index: consumes slot r & 1, fills slot (r + 1) & 1
    int32_t tick_cur;
    int32_t seq;           // written to mirror->dec_seq once the mirror holds the round
    int32_t first;         // 1: the first round of a run of tail rounds
};

// Whether tail round a.r runs; every workgroup of its kernels reads the same
// values (written by earlier kernels).
__device__ __forceinline__ bool tail_runs(const TailArgs &a)
{
    const Ctrl *c = a.ctrl;
    const RoundSlot *in = &c->slot[a.r & 1];
    return (a.first || c->tail_stop < 0) && in->qlen > 0 && in->qedges <= a.tail_edges;
}

// The entry list of a tail round, built without a separate scan: each wave
// takes 64 frontier vertices, scans their out-degrees across its lanes and
// reserves its entries and their edges with one 64-bit atomic (count << 32 |
// edges), so entries stay in edge order across waves; neither half carries
// into the other, as the frontier holds distinct vertices (count < V < 2^31)
// and their out-edges number at most E < 2^31. Writes the entries
// (vertex, row start, exclusive edge offset) and, for each tile of TE edges,
// the entry holding its first edge -- the load-balancing lists k_advance
// reads (k_lb_reduce / k_lb_write build the same for decided rounds).
template <class WT>
__global__ __launch_bounds__(kBlock) void k_tail_list(TailArgs a, int32_t *ent_v,
                                                      int64_t *ent_start, int64_t *ent_scan,
                                                      int32_t *tile, int64_t TE)
{
    if (!tail_runs(a)) return;  // grid-uniform
    const int32_t n = a.ctrl->slot[a.r & 1].qlen;
    const int lane = __lane_id();
    const int64_t nw = (int64_t)gridDim.x * (kBlock / 64);
    for (int64_t b = ((int64_t)blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6)) * 64; b < n;
         b += nw * 64) {
        const int64_t j = b + lane;
        int32_t v = 0;
        int64_t r0 = 0, deg = 0;
        if (j < n) {
            v = a.q[j];
            const uint32_t d = a.dist[v];
            const int64_t p0 = a.row_ptr[v], p1 = a.row_ptr[v + 1];
            if (d != WT::INF) {
                r0 = p0;
                deg = p1 - p0;
            }
        }
        const bool has = deg > 0;
        const uint64_t m = __ballot(has);
        if (m == 0) continue;  // wave-uniform
        int64_t inc = deg;  // inclusive scan of the degrees over the wave
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const int64_t y = __shfl_up(inc, o);
            if (lane >= o) inc += y;
        }
        const int64_t tot = __shfl(inc, 63);
        unsigned long long old = 0;
        if (lane == 0)
            old = atomicAdd(&a.ctrl->tail_list,
                            ((unsigned long long)__popcll(m) << 32) | (unsigned long long)tot);
        old = __shfl(old, 0);
        if (has) {
            const int64_t slot = (int64_t)(old >> 32) + __popcll(m & lanemask_lt());
            const int64_t off = (int64_t)(old & 0xFFFFFFFFull) + inc - deg;
            ent_v[slot] = v;
            ent_start[slot] = r0;
            ent_scan[slot] = off;
            for (int64_t t = (off + TE - 1) / TE; t * TE < off + deg; t++) tile[t] = (int32_t)slot;
        }
    }
}

// Bookkeeping of tail round a.r (one thread): a round that ran resets its
// input slot (the one round a.r + 1 fills) and adds to the run's totals; the
// first round that did not run records itself as tail_stop. Reports to the
// host mirror, the sequence number last.
__global__ void k_tail_close(TailArgs a)
{
    if (threadIdx.x != 0) return;
    Ctrl *c = a.ctrl;
    Ctrl *m = a.mirror;
    const bool ran = tail_runs(a);
    if (a.first) {
        c->tail_stop = -1;
        c->tail_rounds = 0;
        c->tail_relaxed = 0;
    }
    if (ran) {
        const RoundSlot o = c->slot[(a.r + 1) & 1];
        zero_slot(&c->slot[a.r & 1], 0xFFFFFFFFu);
        c->tail_list = 0;
        c->tail_rounds += 1;
        c->tail_relaxed += o.relaxed;
        m->slot[(a.r + 1) & 1] = o;
    } else if (c->tail_stop < 0) {
        c->tail_stop = a.r;
    }
    m->tail_stop = c->tail_stop;
    m->tail_rounds = c->tail_rounds;
    m->tail_relaxed = c->tail_relaxed;
    __threadfence_system();
    __hip_atomic_store(&m->dec_seq, a.seq, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

// ---------------------------------------------------------------------------
// Plain Bellman-Ford pass over an edge list (bellman_ford.cu:62-99), with
// check-then-atomicMin and the change flag kept on the device: pass p zeroes
// the flag pass p+1 will set, and returns at once when pass p-1 changed
// nothing, so several passes can be queued without a host round trip.
// ---------------------------------------------------------------------------
template <class WT, class DT, int VEC, bool GROUPED>
__global__ __launch_bounds__(kBlock) void k_el_pass(const int32_t *__restrict__ src,
                                                    const int32_t *__restrict__ dst,
                                                    const DT *__restrict__ w, int64_t E,
                                                    DT *dist, int32_t *changed, int32_t pass,
                                                    int64_t per_block, const int64_t *goff)
{
    if (blockIdx.x == 0 && threadIdx.x == 0) changed[(pass + 1) & 63] = 0;
    if (pass > 0 && __hip_atomic_load(&changed[(pass - 1) & 63], __ATOMIC_RELAXED,
                                      __HIP_MEMORY_SCOPE_AGENT) == 0)
        return;
    constexpr bool kUseFilt = sizeof(DT) == 4 && VEC == 4;
    __shared__ uint64_t s_filt[kUseFilt ? (1 << kFiltBitsPass) : 1];
    if constexpr (kUseFilt) {
        filt_init<kFiltBitsPass>(s_filt);
        __syncthreads();
    }
    bool low = false;
    // both distances of an edge are loaded before either is used, so the
    // random d[dst] gather does not wait for the d[src] test
    auto relax = [&](DT du, DT dv, int32_t t, DT ww) {
        if (du == WT::INF) return;
        const DT cand = WT::add(du, ww);
        if (cand < dv) {
            if constexpr (kUseFilt)
                if (!filt_send<kFiltBitsPass>(s_filt, t, (uint32_t)cand)) return;
            // dv was read in this pass: the distance drops in this pass
            // whoever wins, so the result of the min is not needed
            __hip_atomic_fetch_min(&dist[t], cand, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            low = true;
        }
    };
    if constexpr (VEC == 4) {
        int64_t g0, g1;
        if constexpr (GROUPED) {
            // XCD-partitioned list: the workgroups of XCD x (blockIdx % 8 == x,
            // round-robin dispatch) share group x, whose destinations map to
            // a disjoint eighth of the distance array
            const int64_t x = blockIdx.x & 7, nbx = gridDim.x >> 3, bi = blockIdx.x >> 3;
            const int64_t ga = goff[x] >> 2, gb = goff[x + 1] >> 2;
            const int64_t per = (gb - ga + nbx - 1) / nbx;
            g0 = ga + bi * per;
            g1 = min(g0 + per, gb);
        } else {
            const int64_t b = xcd_tile(blockIdx.x, gridDim.x);
            g0 = b * per_block;
            g1 = min(g0 + per_block, E >> 2);
        }
        typedef int32_t v4i __attribute__((ext_vector_type(4)));
        typedef uint32_t v4u __attribute__((ext_vector_type(4)));
        const v4i *s4 = reinterpret_cast<const v4i *>(src);
        const v4i *t4 = reinterpret_cast<const v4i *>(dst);
        const v4u *w4 = reinterpret_cast<const v4u *>(w);
        for (int64_t g = g0 + threadIdx.x; g < g1; g += kBlock) {
            const v4i s = __builtin_nontemporal_load(&s4[g]);
            const v4i t = __builtin_nontemporal_load(&t4[g]);
            const v4u ww = __builtin_nontemporal_load(&w4[g]);
            // the list is src-sorted: a lane's next edge usually has the
            // same source, whose distance is then not loaded again
            const DT u0 = dist[s.x];
            const DT u1 = s.y == s.x ? u0 : dist[s.y];
            const DT u2 = s.z == s.y ? u1 : dist[s.z];
            const DT u3 = s.w == s.z ? u2 : dist[s.w];
            const DT v0 = dist[t.x], v1 = dist[t.y], v2 = dist[t.z], v3 = dist[t.w];
            relax(u0, v0, t.x, (DT)ww.x);
            relax(u1, v1, t.y, (DT)ww.y);
            relax(u2, v2, t.z, (DT)ww.z);
            relax(u3, v3, t.w, (DT)ww.w);
        }
        if constexpr (!GROUPED) {
            if (xcd_tile(blockIdx.x, gridDim.x) == 0 && threadIdx.x < (E & 3)) {
                const int64_t e = ((E >> 2) << 2) + threadIdx.x;
                relax(dist[src[e]], dist[dst[e]], dst[e], w[e]);
            }
        }
    } else {
        const int64_t b = xcd_tile(blockIdx.x, gridDim.x);
        const int64_t e0 = b * per_block, e1 = min(e0 + per_block, E);
        for (int64_t e = e0 + threadIdx.x; e < e1; e += kBlock)
            relax(dist[src[e]], dist[dst[e]], dst[e], w[e]);
    }
    if (__ballot(low) != 0 && __lane_id() == 0) changed[pass & 63] = 1;
}

// Frontier of a dense all-edge round (the plain pass over every reached
// source, without appends): the vertices the round lowered -- distance below
// the snapshot taken before it -- that have out-edges join the next queue with
// the round's stamp, the queue and stamps the edge loop's appends would have
// produced (nothing else stamps in such a round, and out-degree 0 is what the
// sink stamp marks). Also fills the round's counters (lowered vertices, their
// minimum) and brings the snapshot up to date, so that a dense round right
// after this one needs no copy. One vertex per lane: the row offsets and
// stamps of a wave's lanes share lines (four vertices per lane measured
// 0.40 ms against 0.31 ms at R-MAT scale 25).
__global__ __launch_bounds__(kBlock) void k_lowered_compact(const uint32_t *dist, uint32_t *snap,
                                                            int32_t *stamp, int32_t tick_cur,
                                                            const int64_t *row_ptr, int64_t V,
                                                            int64_t relaxed, int32_t *qnext,
                                                            RoundSlot *slot)
{
    __shared__ int32_t s_stage[kBlock / 64][kStageCap];
    int32_t *wbuf = s_stage[threadIdx.x >> 6];
    int wcnt = 0;
    RoundAcc acc;
    if (blockIdx.x == 0 && threadIdx.x == 0) acc.relaxed = relaxed;
    const int64_t stride = (int64_t)gridDim.x * kBlock;
    // block-uniform trip count: every lane of a wave reaches stage_push
    for (int64_t base = (int64_t)blockIdx.x * kBlock; base < V; base += stride) {
        const int64_t v = base + threadIdx.x;
        bool app = false;
        if (v < V) {
            const uint32_t d = dist[v];
            if (d < snap[v]) {
                snap[v] = d;
                acc.lowered++;
                acc.minw = min(acc.minw, d);
                const int64_t deg = row_ptr[v + 1] - row_ptr[v];
                if (deg > 0) {
                    stamp[v] = tick_cur;  // only this lane handles v
                    app = true;
                    acc.qedges += deg;
                }
            }
        }
        stage_push(app, (int32_t)v, wbuf, wcnt);
        if (wcnt > kStageCap - 64) stage_flush(wbuf, wcnt, qnext, &slot->qlen);
    }
    stage_flush_block<kBlock / 64>(s_stage, wcnt, qnext, &slot->qlen);
    flush_acc(acc, slot);
}

constexpr int kBlkPer = (1 << kBlkBits) / kBlkThreads;  // block vertices per thread
static_assert(kBlkPer <= 16, "a thread's appends of one tile fit a 16-bit mask");

// One tile of the destination-blocked edge list, shared by the dense round
// and the plain blocked pass: loads the tile's block of 2^kBlkBits distances
// into LDS, then relaxes the tile's edges into it -- d[src] gathered from
// global memory, the minimum taken with LDS atomics. The callers compare the
// LDS minima with the block read again from global memory (64 KB per tile,
// 2 % of a round's bytes at R-MAT scale 25): keeping the loaded values in
// registers instead took 16 VGPRs per lane, which left room for one
// 1024-thread workgroup per CU; with two per CU (8 waves per SIMD, 64 VGPRs
// at most) a dense round at R-MAT scale 25 takes 1.04-1.24 ms instead of
// 1.29-1.35 and a plain blocked pass 0.98-1.09 instead of 1.10-1.19. A wave takes one chunk of 256
// edge slots per step, 4 per lane: the sources are the chunk's first source
// plus a wave prefix sum of the packed deltas (or, for a chunk with a large
// gap, read whole). Returns the number of relaxed edges with a reached source
// (counted only when COUNT).
struct BlkList {
    const uint32_t *pk;     // packed edges: source delta << kBlkBits | destination offset
    const int32_t *cbase;   // per chunk: first source, or -(1 + index) of whole sources
    const int32_t *wide;    // whole sources of the chunks that need them
    const uint32_t *w;      // weights
};

BlkList blk_list(const sssp_graph *g) { return BlkList{g->b_pk, g->b_cbase, g->b_wide, g->b_w}; }

template <class WT, bool COUNT>
__device__ __forceinline__ int64_t blk_relax_tile(const BlkTile &t, const BlkList L,
                                                  const uint32_t *dist, int64_t base, int nb,
                                                  uint32_t *s_d)
{
    typedef int32_t v4i __attribute__((ext_vector_type(4)));
    typedef uint32_t v4u __attribute__((ext_vector_type(4)));
    constexpr uint32_t INF = WT::INF;
    constexpr uint32_t kOffMask = (1u << kBlkBits) - 1;
    static_assert(kBlkChunk == 4 * 64, "a chunk is 4 slots per lane of one wave");
    // the block's distances: all rows' loads in flight before the first LDS
    // store (a load per row under its bounds test made the compiler wait for
    // each row in turn: 16 memory round trips per tile)
    uint32_t row[kBlkPer];
#pragma unroll
    for (int k = 0; k < kBlkPer; k++) {
        const int i = k * kBlkThreads + threadIdx.x;
        row[k] = dist[base + min(i, nb - 1)];  // nb >= 1: an in-bounds read
    }
#pragma unroll
    for (int k = 0; k < kBlkPer; k++) {
        const int i = k * kBlkThreads + threadIdx.x;
        s_d[i] = i < nb ? row[k] : INF;
    }
    __syncthreads();
    const v4u *p4 = reinterpret_cast<const v4u *>(L.pk);
    const v4u *w4 = reinterpret_cast<const v4u *>(L.w);
    const v4i *x4 = reinterpret_cast<const v4i *>(L.wide);
    const int lane = __lane_id();
    int64_t relaxed = 0;
    // branch-free: an unreached source gives INF, which lowers nothing; the
    // LDS minimum returns nothing, so no lane waits on it
    auto relax = [&](uint32_t du, uint32_t loc, uint32_t ww) {
        const bool fin = du != INF;
        if constexpr (COUNT) relaxed += fin ? 1 : 0;
        atomicMin(&s_d[loc], fin ? WT::add(du, ww) : INF);
    };
    // tiles hold whole chunks and the wave's 64 lanes cover one chunk per
    // step, so the loop bound and the chunk index are wave-uniform. The next
    // step's stream loads are issued before this step's gathers.
    const int64_t gend = t.e1 >> 2;
    int64_t g = (t.e0 >> 2) + threadIdx.x;
    int32_t cb = 0;
    v4u pv = {0u, 0u, 0u, 0u}, wv = {0u, 0u, 0u, 0u};
    if (g < gend) {
        cb = L.cbase[g >> 6];
        pv = __builtin_nontemporal_load(&p4[g]);
        wv = __builtin_nontemporal_load(&w4[g]);
    }
    for (; g < gend; g += kBlkThreads) {
        int32_t s0, s1, s2, s3;
        if (cb >= 0) {
            const int32_t a0 = (int32_t)(pv.x >> kBlkBits);
            const int32_t a1 = a0 + (int32_t)(pv.y >> kBlkBits);
            const int32_t a2 = a1 + (int32_t)(pv.z >> kBlkBits);
            const int32_t a3 = a2 + (int32_t)(pv.w >> kBlkBits);
            int32_t incl = a3;
#pragma unroll
            for (int o = 1; o < 64; o <<= 1) {
                const int32_t y = __shfl_up(incl, o);
                if (lane >= o) incl += y;
            }
            const int32_t b0 = cb + (incl - a3);
            s0 = b0 + a0;
            s1 = b0 + a1;
            s2 = b0 + a2;
            s3 = b0 + a3;
        } else {
            const v4i sv = x4[(int64_t)(-1 - cb) * (kBlkChunk / 4) + lane];
            s0 = sv.x;
            s1 = sv.y;
            s2 = sv.z;
            s3 = sv.w;
        }
        // the four gathers are issued together (equal sources, frequent in
        // the src-sorted runs, hit the same line); a reuse test between them
        // would serialise them behind one another's latency
        const uint32_t u0 = dist[s0], u1 = dist[s1], u2 = dist[s2], u3 = dist[s3];
        const int64_t gn = g + kBlkThreads;
        int32_t cbn = 0;
        v4u pvn = {0u, 0u, 0u, 0u}, wvn = {0u, 0u, 0u, 0u};
        if (gn < gend) {
            cbn = L.cbase[gn >> 6];
            pvn = __builtin_nontemporal_load(&p4[gn]);
            wvn = __builtin_nontemporal_load(&w4[gn]);
        }
        relax(u0, pv.x & kOffMask, wv.x);
        relax(u1, pv.y & kOffMask, wv.y);
        relax(u2, pv.z & kOffMask, wv.z);
        relax(u3, pv.w & kOffMask, wv.w);
        cb = cbn;
        pv = pvn;
        wv = wvn;
    }
    __syncthreads();
    return relaxed;
}

// Round counters of a kBlkThreads workgroup through LDS (buf: 5 * kBlkThreads
// words no longer in use): every thread stores its partials, then the first
// wave sums them. The shuffle reductions of flush_acc_n, run by every wave
// with the write-back's values live, pushed k_blk_pass past its 64 registers
// (20 spilled); this keeps it at 57, none spilled.
__device__ __forceinline__ void flush_acc_lds(const RoundAcc &acc, RoundSlot *s, uint32_t *buf)
{
    constexpr int N = kBlkThreads;
    const int tid = threadIdx.x;
    __syncthreads();  // buf's last readers are done
    buf[0 * N + tid] = (uint32_t)acc.relaxed;  // per thread < 2^32: a tile has fewer edges
    buf[1 * N + tid] = (uint32_t)acc.qedges;   // at most 16 degrees, E < 2^31
    buf[2 * N + tid] = (uint32_t)acc.lowered;
    buf[3 * N + tid] = (uint32_t)acc.appended;
    buf[4 * N + tid] = acc.minw;
    __syncthreads();
    if (tid < 64) {
        unsigned long long a = 0, b = 0;
        uint32_t c = 0, q = 0, m = 0xFFFFFFFFu;
        for (int i = tid; i < N; i += 64) {
            a += buf[i];
            b += buf[N + i];
            c += buf[2 * N + i];
            q += buf[3 * N + i];
            m = min(m, buf[4 * N + i]);
        }
        a = wave_sum64((long long)a);
        b = wave_sum64((long long)b);
        c = (uint32_t)wave_sum32((int)c);
        q = (uint32_t)wave_sum32((int)q);
        m = wave_min_u32(m);
        if (tid == 0) {
            if (a) atomicAdd((unsigned long long *)&s->relaxed, a);
            if (b) atomicAdd((unsigned long long *)&s->qedges, b);
            if (c) atomicAdd(&s->lowered, (int)c);
            if (q) atomicAdd(&s->qlen, (int)q);
            if (m != 0xFFFFFFFFu) atomicMin(&s->minw, m);
        }
    }
}

// Dense all-edge round over the destination-blocked edge list. A workgroup
// takes one tile (blk_relax_tile), then writes back the vertices it lowered
// and marks them with the round's stamp. A block with one tile belongs to
// one workgroup: plain stores, plain stamps, no global atomics in the whole
// round. Tiles of a split block (a high in-degree block) write back with a
// returning atomicMin and stamp with atomicExch, so each lowered vertex is
// appended once. The next frontier is left as one 16-bit mask per thread and
// tile (bit k: the thread's vertex of row k), counted into the slot; most
// dense rounds are followed by another, which reads no frontier, and a push
// round first builds its queue from the masks (k_frontier_from_bits). Writing
// the queue here cost 0.06-0.07 ms of a 1.2 ms heavy round at R-MAT scale 25:
// the workgroup had to wait for its returning queue reservation (one
// atomicAdd after a scan of its append counts) before the writes. Fills the
// round's counters.
template <class WT>
__global__ __launch_bounds__(kBlkThreads) __attribute__((amdgpu_waves_per_eu(8))) void k_blk_pass(
    const BlkTile *__restrict__ tiles, const BlkList L, uint32_t *dist, int64_t V,
    int32_t *stamp, int32_t tick_cur, const int64_t *row_ptr, uint16_t *app_bits, RoundSlot *slot)
{
    __shared__ uint32_t s_d[1 << kBlkBits];
    static_assert((1 << kBlkBits) >= 5 * kBlkThreads, "the counters' LDS fits the block's");
    RoundAcc acc;
    const BlkTile t = tiles[blockIdx.x];
    const int64_t base = (int64_t)t.blk << kBlkBits;
    const int nb = (int)min((int64_t)(1 << kBlkBits), V - base);
    acc.relaxed = blk_relax_tile<WT, true>(t, L, dist, base, nb, s_d);
    // write-back: the distance loads of the thread's 16 rows in flight together
    uint32_t lowm = 0;  // bit k: this lane lowered its vertex of row k
#pragma unroll
    for (int k = 0; k < kBlkPer; k++) {
        const int i = k * kBlkThreads + threadIdx.x;
        if (i < nb) {
            // below the value in global memory: this tile lowers v (for a
            // block of one tile that value is the one loaded; for a split
            // block another tile may lower it further before the returning
            // atomicMin, which then reports no change)
            const int32_t v = (int32_t)(base + i);
            const uint32_t nv = s_d[i];
            if (nv < dist[v]) {
                bool low = true;
                if (t.split) low = atomicMin(&dist[v], nv) > nv;
                else dist[v] = nv;  // only this workgroup writes the block
                if (low) {
                    lowm |= 1u << k;
                    acc.lowered++;
                    acc.minw = min(acc.minw, nv);
                }
            }
        }
    }
    // appends: out-degrees and stamps of the lowered vertices, four rows'
    // loads at a time (all sixteen held 64-bit offsets past the register
    // budget of two workgroups per CU, and spilled)
    uint32_t appm = 0;  // bit k: this lane appends its vertex of row k
#pragma unroll
    for (int k0 = 0; k0 < kBlkPer; k0 += 4) {
        if (((lowm >> k0) & 15u) == 0) continue;
        int64_t r0[4], r1[4];
#pragma unroll
        for (int j = 0; j < 4; j++) {
            r0[j] = r1[j] = 0;
            if ((lowm >> (k0 + j)) & 1u) {
                const int64_t v = base + (int64_t)(k0 + j) * kBlkThreads + threadIdx.x;
                r0[j] = row_ptr[v];
                r1[j] = row_ptr[v + 1];
            }
        }
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int64_t deg = r1[j] - r0[j];
            if (deg > 0) {
                const int32_t v = (int32_t)(base + (int64_t)(k0 + j) * kBlkThreads + threadIdx.x);
                bool app;
                if (t.split) {
                    app = atomicExch(&stamp[v], tick_cur) != tick_cur;
                } else {
                    stamp[v] = tick_cur;
                    app = true;
                }
                if (app) {
                    acc.qedges += deg;
                    appm |= 1u << (k0 + j);
                }
            }
        }
    }
    // the frontier stays in these masks (one store per thread, every tile,
    // so no stale bits survive) until a round that reads a queue needs it:
    // k_frontier_from_bits
    app_bits[(int64_t)blockIdx.x * kBlkThreads + threadIdx.x] = (uint16_t)appm;
    acc.appended = __popc(appm);
    flush_acc_lds(acc, slot, s_d);
}

// The queue of the frontier a dense round left in its masks: a workgroup
// takes kMaskTiles consecutive tiles (their masks stay in registers, 4 per
// thread per tile), reserves space for all their vertices with one atomicAdd
// and writes them. The order (tile, thread, row) differs from an appended
// queue's; no round depends on the order.
constexpr int kMaskTiles = 8;
static_assert(kBlkThreads == 4 * kBlock, "a thread reads 4 masks of a tile as one 64-bit word");

__global__ __launch_bounds__(kBlock) void k_frontier_from_bits(const BlkTile *__restrict__ tiles,
                                                               int32_t n_tiles,
                                                               const uint16_t *__restrict__ bits,
                                                               int32_t *q, int32_t *len)
{
    __shared__ int s_w[kBlock / 64];
    __shared__ int s_base;
    const int tid = threadIdx.x, lane = __lane_id(), w = tid >> 6;
    const int64_t t0 = (int64_t)blockIdx.x * kMaskTiles;
    uint64_t word[kMaskTiles];
    int cnt = 0;
#pragma unroll
    for (int i = 0; i < kMaskTiles; i++) {
        word[i] = 0;
        if (t0 + i < n_tiles)
            word[i] = reinterpret_cast<const uint64_t *>(bits)[(t0 + i) * kBlock + tid];
        cnt += __popcll(word[i]);
    }
    // exclusive scan of the counts over the workgroup
    int inc = cnt;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(inc, o);
        if (lane >= o) inc += y;
    }
    if (lane == 63) s_w[w] = inc;
    __syncthreads();
    if (tid == 0) {
        int tot = 0;
        for (int i = 0; i < kBlock / 64; i++) {
            const int c = s_w[i];
            s_w[i] = tot;
            tot += c;
        }
        s_base = tot > 0 ? atomicAdd(len, tot) : 0;
    }
    __syncthreads();
    int pos = s_base + s_w[w] + inc - cnt;
#pragma unroll
    for (int i = 0; i < kMaskTiles; i++) {
        uint64_t m = word[i];
        if (m == 0) continue;
        const int64_t base = (int64_t)tiles[t0 + i].blk << kBlkBits;
        while (m) {
            const int b = __builtin_ctzll(m);
            m &= m - 1;
            // mask j = 4 tid + b / 16 of the tile (thread j of k_blk_pass),
            // bit b % 16 = its row
            const int j = 4 * tid + (b >> 4), k = b & 15;
            q[pos++] = (int32_t)(base + (int64_t)k * kBlkThreads + j);
        }
    }
}

// Plain Bellman-Ford pass (bellman_ford.cu:62-99) over the destination-blocked
// edge list: the tile relaxation of the dense round without its frontier
// (no stamps, no appends, no counters), and the change flag protocol of
// k_el_pass -- pass p zeroes the flag of pass p+1, returns at once when pass
// p-1 changed nothing, and sets its own flag when it lowers a distance -- so
// passes are queued without a host round trip. A tile sets the flag only
// when its minimum for a vertex is below that vertex's distance in global
// memory, which its write then lowers: the flag never reports a change that
// did not happen, and a pass that leaves it clear relaxed every edge against
// the final distances: the fixpoint.
template <class WT>
__global__ __launch_bounds__(kBlkThreads) __attribute__((amdgpu_waves_per_eu(8))) void k_blk_plain(
    const BlkTile *__restrict__ tiles, const BlkList L, uint32_t *dist, int64_t V,
    int32_t *changed, int32_t pass)
{
    if (blockIdx.x == 0 && threadIdx.x == 0) changed[(pass + 1) & 63] = 0;
    if (pass > 0 && __hip_atomic_load(&changed[(pass - 1) & 63], __ATOMIC_RELAXED,
                                      __HIP_MEMORY_SCOPE_AGENT) == 0)
        return;
    __shared__ uint32_t s_d[1 << kBlkBits];
    const BlkTile t = tiles[blockIdx.x];
    const int64_t base = (int64_t)t.blk << kBlkBits;
    const int nb = (int)min((int64_t)(1 << kBlkBits), V - base);
    blk_relax_tile<WT, false>(t, L, dist, base, nb, s_d);
    bool low = false;
#pragma unroll
    for (int k = 0; k < kBlkPer; k++) {
        const int i = k * kBlkThreads + threadIdx.x;
        if (i < nb) {
            const uint32_t nv = s_d[i];
            if (nv < dist[base + i]) {
                low = true;
                if (t.split)  // the result of the min is not needed
                    __hip_atomic_fetch_min(&dist[base + i], nv, __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_AGENT);
                else
                    dist[base + i] = nv;  // only this workgroup writes the block
            }
        }
    }
    if (__ballot(low) != 0 && __lane_id() == 0) changed[pass & 63] = 1;
}

// ---------------------------------------------------------------------------
// Caller order from slot order (relabelled graphs): a gather of each output
// row through the vertex order.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_unpermute(const uint32_t *in, const int32_t *vperm,
                                                      int64_t V, int64_t rows, uint32_t *out)
{
    const int64_t n = V * rows;
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * kBlock) {
        const int64_t r = i / V, v = i - r * V;
        out[i] = in[r * V + vperm[v]];
    }
}

// parent_out[v] = caller id of the parent of v's slot (-1 stays -1)
__global__ __launch_bounds__(kBlock) void k_unpermute_parents(const int32_t *par,
                                                              const int32_t *vperm,
                                                              const int32_t *vinv, int64_t V,
                                                              int32_t *out)
{
    for (int64_t v = (int64_t)blockIdx.x * kBlock + threadIdx.x; v < V;
         v += (int64_t)gridDim.x * kBlock) {
        const int32_t p = par[vperm[v]];
        out[v] = p < 0 ? p : vinv[p];
    }
}

// ---------------------------------------------------------------------------
// Certificate helpers.
// ---------------------------------------------------------------------------
template <class WT>
__global__ __launch_bounds__(kBlock) void k_cert_edges(const int32_t *src, const int32_t *dst,
                                                       const uint32_t *w, int64_t E,
                                                       const uint32_t *dist, int64_t *viol)
{
    int64_t n = 0;
    for (int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x; e < E;
         e += (int64_t)gridDim.x * kBlock) {
        const uint32_t du = dist[src[e]];
        if (du == WT::INF) continue;
        if (WT::add(du, w[e]) < dist[dst[e]]) n++;
    }
    n = wave_sum64(n);
    if (__lane_id() == 0 && n) atomicAdd((unsigned long long *)viol, (unsigned long long)n);
}

__global__ void k_cert_root(Ctrl *c)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    zero_slot(&c->slot[0], 0u);
    zero_slot(&c->slot[1], 0xFFFFFFFFu);
}

// BFS roots of the tight-edge tree: every source (its own parent marks it
// visited during the BFS); counts sources whose distance is not 0.
__global__ __launch_bounds__(kBlock) void k_cert_roots(Ctrl *c, int32_t *parent, int32_t *q0,
                                                       const int64_t *row_ptr, const uint32_t *dist,
                                                       const int32_t *srcs, int32_t one_src,
                                                       int32_t n)
{
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const int32_t s = srcs ? srcs[i] : one_src;
    if (atomicCAS(&parent[s], -1, s) != -1) return;
    if (dist[s] != 0u) atomicAdd((unsigned long long *)&c->cert[1], 1ull);
    const int64_t deg = row_ptr[s + 1] - row_ptr[s];
    if (deg > 0) {
        q0[atomicAdd(&c->slot[0].qlen, 1)] = s;
        atomicAdd((unsigned long long *)&c->slot[0].qedges, (unsigned long long)deg);
    }
}

template <class WT>
__global__ __launch_bounds__(kBlock) void k_cert_count(const uint32_t *dist, int32_t *parent,
                                                       int32_t V, int64_t *cert)
{
    int64_t reach = 0, untight = 0, tree = 0;
    for (int64_t v = (int64_t)blockIdx.x * kBlock + threadIdx.x; v < V;
         v += (int64_t)gridDim.x * kBlock) {
        const bool fin = dist[v] != WT::INF;
        const int32_t p = parent[v];
        reach += fin;
        if (fin && p == -1) untight++;
        if (p != -1) tree++;
        if (p == v) parent[v] = -1;  // a root: only sources are their own parent
    }
    reach = wave_sum64(reach);
    untight = wave_sum64(untight);
    tree = wave_sum64(tree);
    if (__lane_id() == 0) {
        if (untight) atomicAdd((unsigned long long *)&cert[2], (unsigned long long)untight);
        if (reach) atomicAdd((unsigned long long *)&cert[5], (unsigned long long)reach);
        if (tree) atomicAdd((unsigned long long *)&cert[7], (unsigned long long)tree);
    }
}

inline unsigned grid_for(int64_t n, int64_t cap = 8192)
{
    int64_t b = (n + kBlock - 1) / kBlock;
    if (b < 1) b = 1;
    return (unsigned)(b < cap ? b : cap);
}

// ---------------------------------------------------------------------------
// Host side.
// ---------------------------------------------------------------------------
template <class WT, int DIR, int SRC, int PRED>
void launch_adv_k(int kidx, unsigned grid, hipStream_t s, const AdvArgs &a)
{
    switch (kidx) {
    case 0: k_advance<WT, DIR, SRC, PRED, 1><<<grid, kBlock, 0, s>>>(a); break;
    case 1: k_advance<WT, DIR, SRC, PRED, 2><<<grid, kBlock, 0, s>>>(a); break;
    case 2: k_advance<WT, DIR, SRC, PRED, 4><<<grid, kBlock, 0, s>>>(a); break;
    default: k_advance<WT, DIR, SRC, PRED, 8><<<grid, kBlock, 0, s>>>(a); break;
    }
    SSSP_LAUNCH_CHECK("k_advance");
}

enum class Kind { SparsePush, DensePull, SparseBfs };

template <class WT>
void launch_advance(sssp_graph *g, Kind kind, int kidx, AdvArgs &a)
{
    const int64_t TE = (int64_t)kBlock * kChunkK[kidx];
    a.ntiles = (a.T + TE - 1) / TE;
    if (a.ntiles <= 0 || a.n <= 0) return;
    int64_t cap = g->advance_blocks > 0 ? g->advance_blocks : (int64_t)g->cu_count * 8;
    cap = (cap + 7) / 8 * 8;
    int64_t nb = (a.ntiles + 7) / 8 * 8;
    if (nb > cap) nb = cap;
    const unsigned grid = (unsigned)nb;
    hipStream_t s = g->stream;
    switch (kind) {
    case Kind::SparsePush: launch_adv_k<WT, DIR_PUSH, SRC_SPARSE, PRED_NONE>(kidx, grid, s, a); break;
    case Kind::DensePull: launch_adv_k<WT, DIR_PULL, SRC_DENSE, PRED_UNSETTLED>(kidx, grid, s, a); break;
    case Kind::SparseBfs: launch_adv_k<WT, DIR_BFS, SRC_SPARSE, PRED_NONE>(kidx, grid, s, a); break;
    }
}

// Near-far threshold of a push round: entries with d < min over the frontier
// + delta are relaxed, the rest wait. Always above the minimum, so every push
// round relaxes at least the frontier's closest vertex.
uint32_t near_threshold(int wtype, uint32_t minF, double delta)
{
    if (!(delta > 0) || minF == 0xFFFFFFFFu) return 0xFFFFFFFFu;
    uint32_t thr;
    if (wtype == SSSP_W_F32) {
        float f;
        std::memcpy(&f, &minF, 4);
        f = f + (float)delta;
        std::memcpy(&thr, &f, 4);
        if (!(f < 3.0e38f)) thr = 0xFFFFFFFFu;
    } else {
        const uint64_t t = (uint64_t)minF + (uint64_t)(delta < 4e9 ? delta : 4e9);
        thr = t >= 0xFFFFFFFFull ? 0xFFFFFFFFu : (uint32_t)t;
    }
    return thr > minF ? thr : minF + 1;
}

// Builds the near entries of a sparse frontier q[0, n) -- n is an upper
// bound when dev_n is set (pile entries moved onto the queue) -- and, with a
// pile in the rule, moves far entries to pile buffer `nxt`.
void lb_prepare(sssp_graph *g, const int32_t *q, int32_t n, const int32_t *dev_n, int kidx,
                const NearRule &rule, int nxt)
{
    hipStream_t s = g->stream;
    int64_t nb = (n + kLbTile - 1) / kLbTile;
    if (nb < 1) nb = 1;
    if (nb + 1 > g->n_blk) throw Error{SSSP_ERR_INTERNAL, "lb scratch too small"};
    const int64_t TE = (int64_t)kBlock * kChunkK[kidx];
    k_lb_reduce<<<(unsigned)nb, kBlock, 0, s>>>(q, n, dev_n, g->row_ptr, g->dist, rule, g->blk_a,
                                                g->blk_b);
    SSSP_LAUNCH_CHECK("k_lb_reduce");
    scan_blocks_pair_inplace(g->blk_a, g->blk_b, nb, s);
    LbWriteArgs w{q, n, dev_n, g->row_ptr, g->dist, rule, g->blk_a, g->blk_b, nb, g->lb_v,
                  g->lb_start, g->lb_scan, g->lb_tile, TE, g->ctrl, g->farst, g->pile[nxt], nxt};
    k_lb_write<<<(unsigned)nb, kBlock, 0, s>>>(w);
    SSSP_LAUNCH_CHECK("k_lb_write");
}

template <class WT>
AdvArgs base_args(sssp_graph *g, int32_t r, int32_t tick0)
{
    AdvArgs a{};
    a.row_ptr = g->row_ptr;
    a.dist = g->dist;
    a.stamp = g->stamp;
    a.parent = g->parent;
    a.tick_prev = tick0 + r;  // stamp of the frontier this round consumes
    a.tick_cur = tick0 + r + 1;
    a.qnext = g->queue[(r + 1) & 1];
    a.slot = &g->ctrl->slot[(r + 1) & 1];
    return a;
}

// Queues the frontier of round r when a dense round left it in its append
// masks (frontier_bits).
void frontier_to_queue(sssp_graph *g, int32_t r)
{
    if (!g->frontier_bits) return;
    hipStream_t s = g->stream;
    k_fill<int32_t><<<1, 64, 0, s>>>(&g->ctrl->mat_n, 1, 0);
    SSSP_LAUNCH_CHECK("k_fill mat_n");
    const unsigned grid = (unsigned)((g->n_btile + kMaskTiles - 1) / kMaskTiles);
    k_frontier_from_bits<<<grid, kBlock, 0, s>>>(g->b_tile, g->n_btile, g->b_app,
                                                 g->queue[r & 1], &g->ctrl->mat_n);
    SSSP_LAUNCH_CHECK("k_frontier_from_bits");
    g->frontier_bits = false;
}

// Runs one relaxation round in the given mode. `pile` is the pile buffer the
// round reads; the host mirror of the control block holds its counters.
// Returns true when a push round wrote the next pile (the caller swaps).
template <class WT>
bool run_round(sssp_graph *g, int mode, int kidx, int32_t r, int32_t tick0, const RoundSlot &cur,
               bool verify)
{
    AdvArgs a = base_args<WT>(g, r, tick0);
    hipStream_t s = g->stream;
    const int p = g->pile_cur;
    const Ctrl &h = *g->ctrl_host;
    const int32_t pn = h.pile_n[p];
    const bool snap_ok = r > 0 && g->snap_round == r - 1;
    g->snap_round = -1;  // any round but a dense one changes distances behind the snapshot
    g->last_round_dense = false;
    // a round that reads the frontier as a queue builds it first (push);
    // every round but k_blk_pass's leaves the next frontier queued
    if (mode != SSSP_MODE_PUSH || verify) g->frontier_bits = false;
    if (mode != SSSP_MODE_PUSH && !verify && pn > 0) {
        k_pile_flush<<<grid_for(pn, 4096), kBlock, 0, s>>>(g->pile[p], pn, g->stamp, a.tick_prev,
                                                           g->farst, g->ctrl, p);
        SSSP_LAUNCH_CHECK("k_pile_flush");
    }
    if (verify || mode == SSSP_MODE_ALL) {
        const double fe = (double)cur.qedges + (pn > 0 ? (double)h.pile_edges[p] : 0.0);
        const double unreached =
            cur.sample_n > 0 ? (double)cur.sample_unreached / cur.sample_n : 0.0;
        if (verify || fe >= g->dense_all_frac * (double)g->E || unreached >= g->dense_unreached) {
            // dense all-edge round: with most edges active -- a large
            // frontier, or most of the graph still unreached so that the sweep
            // lowers most vertices it meets -- relaxing every reached source
            // without per-edge appends costs less than the edge loop's
            // appends: over the destination-blocked list (LDS minima, no
            // global atomics, the frontier marked by the same kernel), or
            // the plain pass followed by a compaction of the lowered vertices
            // against a snapshot
            g->last_round_dense = true;
            if (g->dense_kernel == 0 && g->n_btile > 0) {
                k_blk_pass<WT><<<(unsigned)g->n_btile, kBlkThreads, 0, s>>>(
                    g->b_tile, blk_list(g), g->dist, g->V, g->stamp, a.tick_cur,
                    g->row_ptr, g->b_app, a.slot);
                SSSP_LAUNCH_CHECK("k_blk_pass");
                g->frontier_bits = true;
                return false;
            }
            // the snapshot still holds the distances a dense round r - 1 left
            // (its compaction updated it) unless anything else ran since
            if (!snap_ok)
                SSSP_HIP(hipMemcpyAsync(g->snap, g->dist, 4 * (size_t)g->V,
                                        hipMemcpyDeviceToDevice, s));
            const int64_t nb = std::max<int64_t>(8, (int64_t)g->cu_count * 8);
            const int64_t per_block = ((g->E >> 2) + nb - 1) / nb;
            k_el_pass<WT, uint32_t, 4, true><<<(unsigned)nb, kBlock, 0, s>>>(
                g->x_src, g->x_dst, g->x_w, g->E, g->dist, g->ctrl->changed, 0, per_block,
                g->xgoff);
            SSSP_LAUNCH_CHECK("k_el_pass (dense round)");
            k_lowered_compact<<<grid_for(g->V, (int64_t)g->cu_count * 8), kBlock, 0, s>>>(
                g->dist, g->snap, g->stamp, a.tick_cur, g->row_ptr, g->V, g->E, a.qnext, a.slot);
            SSSP_LAUNCH_CHECK("k_lowered_compact");
            g->snap_round = r;
            return false;
        }
        // all-edge round streamed over the src-sorted edge list; the chunk
        // bucket sets the grid (2, 4, 8 or 16 workgroups per CU)
        const int64_t nblocks = (int64_t)g->cu_count * (2 << kidx);
        if (verify)
            k_el_active<WT, PRED_REACHED><<<(unsigned)nblocks, kBlock, 0, s>>>(
                g->x_src, g->x_dst, g->x_w, g->E, g->dist, g->stamp, a.tick_prev, a.tick_cur,
                g->row_ptr, a.qnext, a.slot, g->xgoff);
        else
            k_el_active<WT, PRED_STAMP><<<(unsigned)nblocks, kBlock, 0, s>>>(
                g->x_src, g->x_dst, g->x_w, g->E, g->dist, g->stamp, a.tick_prev, a.tick_cur,
                g->row_ptr, a.qnext, a.slot, g->xgoff);
        SSSP_LAUNCH_CHECK("k_el_active");
        return false;
    }
    switch (mode) {
    case SSSP_MODE_PUSH: {
        frontier_to_queue(g, r);
        int32_t *q = g->queue[r & 1];
        const bool piled = g->delta > 0;
        NearRule rule{0xFFFFFFFFu, 0, nullptr, a.tick_cur};
        int32_t nbound = cur.qlen;
        int64_t ebound = cur.qedges;
        if (piled) {
            const uint32_t minF = cur.minw < h.pile_min[p] ? cur.minw : h.pile_min[p];
            rule.thr = near_threshold(g->wtype, minF, g->delta);
            rule.hub_degree = g->hub_degree;
            rule.farst = g->farst;
            if (pn > 0) {
                k_pile_filter<<<grid_for(pn, 4096), kBlock, 0, s>>>(
                    g->pile[p], pn, g->dist, rule.thr, g->farst, g->stamp, a.tick_prev, a.tick_cur,
                    q, g->row_ptr, g->pile[p ^ 1], g->ctrl, p ^ 1);
                SSSP_LAUNCH_CHECK("k_pile_filter");
                nbound += pn;
                ebound += h.pile_edges[p];
            }
        }
        lb_prepare(g, q, nbound, piled ? &g->ctrl->lb_n : nullptr, kidx, rule, p ^ 1);
        // a large frontier marks its lowered vertices in the bitmap
        const bool bits = g->fbits && g->push_bits_edges > 0 && !piled &&
                          ebound >= g->push_bits_edges;
        a.fbits = bits ? g->fbits : nullptr;
        a.ent_v = g->lb_v;
        a.ent_start = g->lb_start;
        a.ent_scan = g->lb_scan;
        a.tile = g->lb_tile;
        a.n = nbound;       // upper bounds for the grid; the kernel reads the
        a.T = ebound;       // compacted counts from the device
        a.dev_n = &g->ctrl->n_compact;
        a.dev_T = &g->ctrl->lb_total;
        a.adj = g->el_dst;
        a.wts = g->el_w;
        launch_advance<WT>(g, Kind::SparsePush, kidx, a);
        if (bits) {
            k_bits_frontier<<<(unsigned)g->cu_count * 8, kBlock, 0, s>>>(
                g->fbits, g->V, g->row_ptr, g->stamp, a.tick_cur, a.qnext, a.slot);
            SSSP_LAUNCH_CHECK("k_bits_frontier");
        }
        return piled;
    }
    case SSSP_MODE_PULL: {
        a.ent_v = g->nzi_v;
        a.ent_scan = g->nzi_scan;
        a.tile = g->nzi_tile[kidx];
        a.n = g->n_nzi;
        a.T = g->E;
        a.adj = g->csc_src;
        a.wts = g->csc_w;
        // settled test against the minimum over the frontier and the pile
        a.minF = cur.minw < h.pile_min[p] ? cur.minw : h.pile_min[p];
        launch_advance<WT>(g, Kind::DensePull, kidx, a);
        return false;
    }
    default: throw Error{SSSP_ERR_INTERNAL, "bad round mode"};
    }
}

void read_slot(sssp_graph *g, int32_t r)
{
    hipStream_t s = g->stream;
    Ctrl *d = g->ctrl;
    Ctrl *h = g->ctrl_host;
    SSSP_HIP(hipMemcpyAsync(&h->slot[r & 1], &d->slot[r & 1], sizeof(RoundSlot),
                            hipMemcpyDeviceToHost, s));
    SSSP_HIP(hipMemcpyAsync(&h->dec_mode, &d->dec_mode,
                            sizeof(Ctrl) - offsetof(Ctrl, dec_mode), hipMemcpyDeviceToHost, s));
    SSSP_HIP(hipStreamSynchronize(s));
}

template <class WT>
void solve_plain(sssp_graph *g, sssp_stats *st)
{
    hipStream_t s = g->stream;
    const int64_t E = g->E;
    const int64_t ng = E >> 2;
    int64_t nblocks = (int64_t)g->cu_count * 8;
    if (nblocks < 8) nblocks = 8;
    const int64_t per_block = (ng + nblocks - 1) / nblocks;
    const bool blocked = plain_uses_blocked(g);
    int32_t pass = 0;
    // with trace_events every pass is bracketed by events (one pass per batch)
    const int batch = g->trace_events ? 1 : 4;
    int32_t done_at = -1;
    while (done_at < 0) {
        for (int i = 0; i < batch; i++) {
            if (g->trace_events) SSSP_HIP(hipEventRecord(g->ev_r0, s));
            if (blocked) {
                k_blk_plain<WT><<<(unsigned)g->n_btile, kBlkThreads, 0, s>>>(
                    g->b_tile, blk_list(g), g->dist, g->V, g->ctrl->changed,
                    pass + i);
                SSSP_LAUNCH_CHECK("k_blk_plain");
            } else {
                k_el_pass<WT, uint32_t, 4, true><<<(unsigned)nblocks, kBlock, 0, s>>>(
                    g->x_src, g->x_dst, g->x_w, E, g->dist, g->ctrl->changed, pass + i,
                    per_block, g->xgoff);
                SSSP_LAUNCH_CHECK("k_el_pass");
            }
            if (g->trace_events) SSSP_HIP(hipEventRecord(g->ev_r1, s));
        }
        SSSP_HIP(hipMemcpyAsync(g->ctrl_host->changed, g->ctrl->changed, sizeof(int32_t) * 64,
                                hipMemcpyDeviceToHost, s));
        SSSP_HIP(hipStreamSynchronize(s));
        if (g->trace_events) {
            TraceRow tr{};
            tr.mode = SSSP_MODE_ALL;
            tr.frontier = g->V;
            tr.frontier_edges = E;
            SSSP_HIP(hipEventElapsedTime(&tr.ms, g->ev_r0, g->ev_r1));
            g->trace.push_back(tr);
        }
        for (int i = 0; i < batch; i++) {
            if (g->ctrl_host->changed[(pass + i) & 63] == 0) {
                done_at = pass + i;
                break;
            }
        }
        pass += batch;
        if (done_at < 0 && pass > g->max_rounds)
            throw Error{SSSP_ERR_NOT_CONVERGED, "plain Bellman-Ford did not converge within max_rounds"};
        if (done_at < 0 && pass >= g->V + 1 && pass > 4)
            throw Error{SSSP_ERR_NOT_CONVERGED, "plain Bellman-Ford exceeded V passes"};
    }
    st->iterations = done_at + 1;
    st->n_all = done_at + 1;
    st->verify_passes = 1;  // the last pass changed nothing: the fixpoint check
    st->edges_relaxed = (int64_t)(done_at + 1) * E;
}

// Waits until the decision kernel has written round `seq` into the host
// mirror: a spin on the mirror's sequence number (coherent host memory),
// which sees the write about a microsecond after it lands -- a stream
// synchronize wakes the host up later, and the GPU idles until the round's
// kernels are launched. After kSpinPolls polls it falls back to the stream
// synchronize, which also reports a failed kernel.
constexpr int64_t kSpinPolls = (int64_t)1 << 22;

void wait_decision(sssp_graph *g, int32_t seq)
{
    volatile int32_t *p = &g->ctrl_host->dec_seq;
    for (int64_t i = 0; g->decide_spin && i < kSpinPolls; i++) {
        if (*p == seq) {
            std::atomic_thread_fence(std::memory_order_acquire);
            return;
        }
        if ((i & 1023) == 1023 && hipStreamQuery(g->stream) != hipErrorNotReady) break;
    }
    SSSP_HIP(hipStreamSynchronize(g->stream));
}

// Tail rounds from round r on, tail_group queued per host check, until one
// does not run (empty frontier, or more than tail_edges out-edges again) or
// the round cap. Returns the index of the first round not run here; its input
// slot is in the host mirror.
template <class WT>
int32_t run_tail(sssp_graph *g, int32_t r, int32_t tick0, int32_t r_cap, sssp_stats *st,
                 int64_t &relaxed)
{
    hipStream_t s = g->stream;
    const int group = g->trace_events ? 1 : std::max(1, g->tail_group);
    // the list kernel's grid covers 2^17 frontier vertices per pass; the
    // relaxation's, tail_blocks workgroups (0 = 2 per CU) looping over tiles
    // (measured at R-MAT 25 x 16 and 22 x 32: 2 per CU 1-2 % faster than 4)
    const unsigned grid = (unsigned)std::max(1, g->cu_count * 2);
    const int64_t nb = g->tail_blocks > 0 ? g->tail_blocks : (int64_t)g->cu_count * 2;
    const unsigned agrid_small = (unsigned)std::max<int64_t>(8, (nb + 7) / 8 * 8);
    TailArgs a{};
    a.ctrl = g->ctrl;
    a.mirror = g->ctrl_host_dev;
    a.row_ptr = g->row_ptr;
    a.adj = g->el_dst;
    a.wts = g->el_w;
    a.dist = g->dist;
    a.stamp = g->stamp;
    a.tail_edges = g->tail_edges;
    a.first = 1;
    const Ctrl &h = *g->ctrl_host;
    const int32_t r0 = r;
    for (;;) {
        const int n = (int)std::min<int64_t>(group, (int64_t)r_cap - r);
        if (n <= 0) break;
        const RoundSlot in = h.slot[r & 1];  // the frontier the first queued round consumes
        const bool big = in.qedges > kTailBigEdges;
        const int kidx = big ? kTailBigChunk : kTailChunk;
        const int64_t TE = (int64_t)kBlock * kChunkK[kidx];
        const unsigned agrid = big ? (unsigned)g->cu_count * 8 : agrid_small;
        if (g->trace_events) SSSP_HIP(hipEventRecord(g->ev_r0, s));
        for (int i = 0; i < n; i++) {
            a.r = r + i;
            a.q = g->queue[a.r & 1];
            a.qnext = g->queue[(a.r + 1) & 1];
            a.tick_cur = tick0 + a.r + 1;
            a.seq = ++g->dec_seq;
            k_tail_list<WT><<<grid, kBlock, 0, s>>>(a, g->lb_v, g->lb_start, g->lb_scan,
                                                     g->lb_tile, TE);
            SSSP_LAUNCH_CHECK("k_tail_list");
            AdvArgs adv = base_args<WT>(g, a.r, tick0);
            adv.ent_v = g->lb_v;
            adv.ent_start = g->lb_start;
            adv.ent_scan = g->lb_scan;
            adv.tile = g->lb_tile;
            adv.adj = g->el_dst;
            adv.wts = g->el_w;
            adv.dev_nt = &g->ctrl->tail_list;
            adv.n = 1;  // the counts are read on the device
            adv.T = 1;
            adv.ntiles = 1;
            launch_adv_k<WT, DIR_PUSH, SRC_SPARSE, PRED_NONE>(kidx, agrid, s, adv);
            k_tail_close<<<1, 64, 0, s>>>(a);
            SSSP_LAUNCH_CHECK("k_tail_close");
            a.first = 0;
        }
        if (g->trace_events) SSSP_HIP(hipEventRecord(g->ev_r1, s));
        wait_decision(g, a.seq);
        if (h.tail_stop >= 0) {
            r = h.tail_stop;
            break;
        }
        if (g->trace_events) {
            float ms = 0.f;
            SSSP_HIP(hipEventSynchronize(g->ev_r1));
            SSSP_HIP(hipEventElapsedTime(&ms, g->ev_r0, g->ev_r1));
            TraceRow tr{};
            tr.mode = SSSP_MODE_TAIL;
            tr.frontier = in.qlen;
            tr.frontier_edges = in.qedges;
            tr.ms = ms;
            g->trace.push_back(tr);
        }
        r += n;
    }
    // totals of the rounds run here; the last one's relaxations are in the
    // slot the next decided round reads, which counts them
    const int32_t done = r > r0 ? h.tail_rounds : 0;
    if (done > 0) relaxed += h.tail_relaxed - h.slot[r & 1].relaxed;
    if (!g->trace_events) {
        // one row per round (sizes and times are only traced with
        // trace_events: rounds queued together are not told apart)
        TraceRow tr{};
        tr.mode = SSSP_MODE_TAIL;
        tr.frontier = tr.frontier_edges = -1;
        for (int32_t i = 0; i < done; i++) g->trace.push_back(tr);
    }
    st->iterations += done;
    st->n_push += done;
    return r;
}

// One round r: bookkeeping and decision (k_round_open, one launch), host
// read-back of the frontier counters, then the round itself -- or, for a
// decided push round over a frontier with few out-edges, the tail rounds
// (run_tail, below r_cap). Returns the number of round indices used, 0
// (having run nothing) when the solve is over (no frontier, no pile,
// fixpoint check done or off). force_mode >= 0 replaces the decision (not the
// fixpoint check) and runs one round.
template <class WT>
int32_t step_round(sssp_graph *g, int32_t r, int sched, int32_t tick0, sssp_stats *st,
                   bool &verified, int64_t &relaxed, int32_t r_cap, int force_mode = -1,
                   int force_chunk = 0)
{
    hipStream_t s = g->stream;
    const bool learned = (sched == SSSP_SCHED_LEARNED);
    const int ns = g->n_sample;
    DecideArgs da{};
    da.r = r;
    da.sched = sched;
    da.V = g->V;
    da.E = (double)g->E;
    da.has_csc = g->has_csc ? 1 : 0;
    da.push_edge_frac = (float)g->push_edge_frac;
    da.pull_frac = (float)g->pull_frac;
    da.push_chunk = g->push_chunk;
    da.all_chunk = g->all_chunk;
    da.pull_chunk = g->pull_chunk;
    sssp_mlp *net = g->solving_bfs ? g->bfs_mlp : g->mlp;
    da.mlp = (learned && net) ? mlp_device_params(net) : nullptr;
    da.pile = g->pile_cur;
    da.mirror = g->ctrl_host_dev;
    da.seq = ++g->dec_seq;
    k_round_open<<<grid_for(ns, 64), kBlock, 0, s>>>(g->ctrl, r, g->pile_cur, g->sample_v,
                                                     g->has_csc ? g->sample_in : nullptr, ns,
                                                     g->dist, WT::INF, da);
    SSSP_LAUNCH_CHECK("k_round_open");
    wait_decision(g, da.seq);
    const RoundSlot cur = g->ctrl_host->slot[r & 1];
    const int32_t piled = g->ctrl_host->pile_n[g->pile_cur];
    const int64_t piled_edges = g->ctrl_host->pile_edges[g->pile_cur];
    relaxed += cur.relaxed;
    bool verify = false;
    if (cur.qlen == 0 && piled == 0) {
        if (!g->verify || verified) return 0;
        verify = true;
    }
    int mode, kidx;
    switch (sched) {
    case SSSP_SCHED_PUSH: mode = SSSP_MODE_PUSH; kidx = g->push_chunk; break;
    case SSSP_SCHED_ALL: mode = SSSP_MODE_ALL; kidx = g->all_chunk; break;
    case SSSP_SCHED_PULL:
        mode = g->has_csc ? SSSP_MODE_PULL : SSSP_MODE_ALL;
        kidx = g->pull_chunk;
        break;
    default:
        mode = g->ctrl_host->dec_mode;
        kidx = g->ctrl_host->dec_chunk;
        if (g->ctrl_host->dec_mlp) st->mlp_rounds++;
        break;
    }
    if (force_mode >= 0) {
        mode = force_mode;
        kidx = force_chunk;
    }
    if (mode == SSSP_MODE_PULL && !g->has_csc) mode = SSSP_MODE_ALL;
    if (kidx < 0 || kidx >= kChunkCount) kidx = 1;
    if (!verify && cur.qlen == 0 && mode != SSSP_MODE_PUSH && force_mode < 0) {
        // only the pile is pending: a push round drains it
        mode = SSSP_MODE_PUSH;
        kidx = g->push_chunk;
    }
    if (verify) {
        mode = SSSP_MODE_ALL;
        kidx = g->all_chunk;
        st->verify_passes++;
        verified = true;
    } else {
        verified = false;
    }
    // the tail rounds take over once the frontier, having had more than
    // tail_edges out-edges, is down to at most that (a frontier still growing
    // from the source may hold hubs, whose edges one wave would walk alone)
    if (cur.qedges + piled_edges > g->tail_edges) g->tail_armed = true;
    if (!verify && mode == SSSP_MODE_PUSH && force_mode < 0 && g->tail_edges > 0 &&
        g->tail_armed && (sched == SSSP_SCHED_HEURISTIC || sched == SSSP_SCHED_LEARNED) &&
        piled == 0 && g->delta <= 0.0 && cur.qlen > 0 && cur.qedges <= g->tail_edges &&
        r + 1 < r_cap) {
        frontier_to_queue(g, r);
        const int32_t r1 = run_tail<WT>(g, r, tick0, r_cap, st, relaxed);
        if (r1 > r) return r1 - r;
        // (the first tail round always runs: its frontier was checked here)
    }
    if (g->trace_events) SSSP_HIP(hipEventRecord(g->ev_r0, s));
    if (run_round<WT>(g, mode, kidx, r, tick0, cur, verify)) g->pile_cur ^= 1;
    float ms = 0.f;
    if (g->trace_events) {
        SSSP_HIP(hipEventRecord(g->ev_r1, s));
        SSSP_HIP(hipEventSynchronize(g->ev_r1));
        SSSP_HIP(hipEventElapsedTime(&ms, g->ev_r0, g->ev_r1));
    }
    TraceRow tr{};
    tr.mode = verify ? SSSP_MODE_CHECK : g->last_round_dense ? SSSP_MODE_DENSE : mode;
    tr.chunk = kidx;
    tr.frontier = (int64_t)cur.qlen + piled;
    tr.frontier_edges = verify ? g->E : cur.qedges + piled_edges;
    tr.ms = ms;
    tr.unsettled = cur.sample_n > 0 ? (float)cur.sample_unsettled / cur.sample_n : 0.f;
    g->trace.push_back(tr);
    st->iterations++;
    if (mode == SSSP_MODE_PUSH) st->n_push++;
    else if (mode == SSSP_MODE_ALL) st->n_all++;
    else st->n_pull++;
    return 1;
}

// Returns the number of round indices the solve used (its stamp ticks end
// at tick0 + that number).
template <class WT>
int32_t solve_rounds(sssp_graph *g, int sched, int32_t tick0, sssp_stats *st)
{
    bool verified = false;
    int64_t relaxed = 0;
    g->pile_cur = 0;
    g->tail_armed = false;
    g->frontier_bits = false;
    int32_t r = 0;
    for (;;) {
        if (r >= g->max_rounds)
            throw Error{SSSP_ERR_NOT_CONVERGED, "solve did not converge within max_rounds"};
        const int32_t used = step_round<WT>(g, r, sched, tick0, st, verified, relaxed,
                                            g->max_rounds);
        if (used == 0) break;
        r += used;
    }
    st->edges_relaxed = relaxed;
    return r;
}

void upload_sources(sssp_graph *g, const int32_t *sources, int32_t n)
{
    if (n > g->src_cap) {
        dfree(g->src_list);
        g->src_list = nullptr;
        g->src_list = dmalloc<int32_t>((size_t)n);
        g->src_cap = n;
    }
    SSSP_HIP(hipMemcpyAsync(g->src_list, sources, 4 * (size_t)n, hipMemcpyHostToDevice, g->stream));
}

template <class WT>
void solve_t(sssp_graph *g, const int32_t *sources, int32_t n_src, int sched, sssp_stats *st)
{
    hipStream_t s = g->stream;
    if (g->tick > (1 << 30) - g->max_rounds - 4) {
        reset_stamps(g);
        g->tick = 1;
    }
    const int32_t tick0 = g->tick;
    const int32_t source = sources[0];
    if (n_src > 1) upload_sources(g, sources, n_src);
    g->trace.clear();
    g->snap_round = -1;
    SSSP_HIP(hipEventRecord(g->ev_a, s));
    k_fill<uint32_t><<<grid_for(g->V), kBlock, 0, s>>>(g->dist, g->V, WT::INF);
    SSSP_LAUNCH_CHECK("k_fill dist");
    k_init_solve<<<1, 64, 0, s>>>(g->ctrl, g->dist, g->stamp, g->queue[0], g->row_ptr,
                                  n_src > 1 ? -1 : source, tick0);
    SSSP_LAUNCH_CHECK("k_init_solve");
    if (n_src > 1) {
        k_init_sources<<<grid_for(n_src, 1 << 20), kBlock, 0, s>>>(
            g->ctrl, g->dist, g->stamp, g->queue[0], g->row_ptr, g->src_list, n_src, tick0);
        SSSP_LAUNCH_CHECK("k_init_sources");
    }
    // Below plain_below_edges the rule and the network hand the whole solve
    // to the plain pass loop: its passes (the blocked pass from 64 tiles)
    // then cost less than the rounds with their fixed per-round cost.
    // Measured (learned, ms): R-MAT 21 x 16 (2^25 edges) 2.98 plain against
    // 3.66 in rounds, BFS 0.79 against 0.95; at 2^26 edges about even (R-MAT
    // 22 x 16 5.12 / 5.00, R-MAT 21 x 32 4.46 / 4.71); R-MAT 22 x 32 (2^27)
    // 7.1 against 5.6 in rounds.
    const bool small = (sched == SSSP_SCHED_HEURISTIC || sched == SSSP_SCHED_LEARNED) &&
                       n_src == 1 && g->E < g->plain_below_edges;
    if (sched == SSSP_SCHED_PLAIN || small) {
        solve_plain<WT>(g, st);
        g->tick = tick0 + 1;
    } else {
        if (g->pile_dirty) {
            // an earlier solve stopped with vertices still piled
            k_fill<int32_t><<<grid_for(g->V), kBlock, 0, s>>>(g->farst, g->V, 0);
            SSSP_LAUNCH_CHECK("pile state reset");
        }
        g->pile_dirty = true;
        int32_t used = 0;
        try {
            used = solve_rounds<WT>(g, sched, tick0, st);
        } catch (...) {
            g->tick = tick0 + g->max_rounds + 2;
            throw;
        }
        g->pile_dirty = false;
        g->tick = tick0 + used + 2;
    }
    // the result in the caller's vertex order is part of the solve: timed
    if (g->vperm) unpermute_rows(g->dist, g->vperm, g->V, 1, g->dout, s);
    SSSP_HIP(hipEventRecord(g->ev_b, s));
    SSSP_HIP(hipEventSynchronize(g->ev_b));
    float ms = 0.f;
    SSSP_HIP(hipEventElapsedTime(&ms, g->ev_a, g->ev_b));
    st->t_solve_ms = ms;
    st->gteps = ms > 0 ? (double)g->E / (ms * 1e-3) / 1e9 : 0.0;
    // reached vertices, counted after the timed region
    k_count_reached<WT><<<grid_for(g->V, 2048), kBlock, 0, s>>>(g->dist, g->V, &g->ctrl->n_reached);
    SSSP_LAUNCH_CHECK("k_count_reached");
    int64_t nr = 0;
    SSSP_HIP(hipMemcpyAsync(&nr, &g->ctrl->n_reached, sizeof(nr), hipMemcpyDeviceToHost, s));
    SSSP_HIP(hipStreamSynchronize(s));
    st->reached = (int32_t)nr;
}

template <class WT>
void certify_t(sssp_graph *g, const int32_t *sources, int32_t n_src, int64_t *counts)
{
    if (n_src > 1) upload_sources(g, sources, n_src);
    hipStream_t s = g->stream;
    Ctrl *c = g->ctrl;
    k_fill<int64_t><<<1, 64, 0, s>>>(c->cert, 8, 0);
    k_cert_edges<WT><<<grid_for(g->E), kBlock, 0, s>>>(g->el_src, g->el_dst, g->el_w, g->E, g->dist,
                                                       &c->cert[0]);
    SSSP_LAUNCH_CHECK("k_cert_edges");
    k_fill<int32_t><<<grid_for(g->V), kBlock, 0, s>>>(g->parent, g->V, -1);
    k_cert_root<<<1, 64, 0, s>>>(c);
    k_cert_roots<<<grid_for(n_src, 1 << 20), kBlock, 0, s>>>(c, g->parent, g->queue[0], g->row_ptr,
                                                              g->dist, n_src > 1 ? g->src_list : nullptr,
                                                              sources[0], n_src);
    SSSP_LAUNCH_CHECK("k_cert_roots");
    for (int32_t r = 0;; r++) {
        if (r > g->V + 2) throw Error{SSSP_ERR_INTERNAL, "certificate BFS did not terminate"};
        k_round_begin<<<1, kBlock, 0, s>>>(c, r, -1, g->sample_v, nullptr, 0, g->dist,
                                           0xFFFFFFFFu);
        SSSP_LAUNCH_CHECK("k_round_begin (bfs)");
        read_slot(g, r);
        const RoundSlot cur = g->ctrl_host->slot[r & 1];
        if (cur.qlen == 0) break;
        AdvArgs a = base_args<WT>(g, r, 0);
        lb_prepare(g, g->queue[r & 1], cur.qlen, nullptr, 2, NearRule{0xFFFFFFFFu, 0, nullptr, 0}, 0);
        a.ent_v = g->lb_v;
        a.ent_start = g->lb_start;
        a.ent_scan = g->lb_scan;
        a.tile = g->lb_tile;
        a.n = cur.qlen;
        a.T = cur.qedges;
        a.adj = g->el_dst;
        a.wts = g->el_w;
        launch_advance<WT>(g, Kind::SparseBfs, 2, a);
    }
    k_cert_count<WT><<<grid_for(g->V), kBlock, 0, s>>>(g->dist, g->parent, g->V, c->cert);
    SSSP_LAUNCH_CHECK("k_cert_count");
    SSSP_HIP(hipMemcpyAsync(counts, c->cert, sizeof(int64_t) * 8, hipMemcpyDeviceToHost, s));
    SSSP_HIP(hipStreamSynchronize(s));
    // the BFS tree counts the source's own marker once: tree size = vertices
    // with a parent + the source
}

// Timing collection with rollouts. At every round r of a solve the state is
// saved (distances, stamps, pile states, the round's queue and pile, the
// control block) and each mode x chunk bucket c is tried: round r runs c and
// the rounds after it follow the hand-written rule (or, rollout_policy = 1,
// the network set on the graph) to the end of the solve;
// times[r][c] is the best of two device timings of that whole remainder (host
// round trips included, as in a real solve). The solve then advances by the best c. The
// labels are thus one step of policy improvement over the rule: a round that
// costs more now but leaves less work scores by the total.
// The cost is quadratic in the solve's round count R (each of R rounds runs
// 12 choices x 2 timings of up to R remaining rounds), so a far-pile solve of
// hundreds of rounds runs hundreds of thousands of them: the host counts
// every round it launches and gives up with SSSP_ERR_NOT_CONVERGED past
// g->collect_budget (between rounds, so every kernel has ended normally).
constexpr int kRolloutReps = 2;

template <class WT>
void collect_t(sssp_graph *g, int32_t source, int32_t cap, float *feats, float *times,
               int32_t *rows)
{
    hipStream_t s = g->stream;
    if (g->tick > (1 << 30) - 2 * g->max_rounds - 4) {
        reset_stamps(g);
        g->tick = 1;
    }
    if (g->pile_dirty) {
        k_fill<int32_t><<<grid_for(g->V), kBlock, 0, s>>>(g->farst, g->V, 0);
        SSSP_LAUNCH_CHECK("pile state reset");
    }
    const int32_t tick0 = g->tick;
    const size_t V = (size_t)g->V;
    struct Saved {
        void *p[6] = {};
        ~Saved()
        {
            for (void *x : p) dfree(x);
        }
    } sv;
    uint32_t *sd = dmalloc<uint32_t>(V);
    sv.p[0] = sd;
    int32_t *sst = dmalloc<int32_t>(V);
    sv.p[1] = sst;
    int32_t *sfa = dmalloc<int32_t>(V);
    sv.p[2] = sfa;
    int32_t *sq = dmalloc<int32_t>(V + 1);
    sv.p[3] = sq;
    int32_t *sp = dmalloc<int32_t>(V);
    sv.p[4] = sp;
    Ctrl *sc = dmalloc<Ctrl>(1);
    sv.p[5] = sc;
    const int trace_was = g->trace_events;
    g->trace_events = 0;
    g->trace.clear();
    g->pile_dirty = true;
    g->pile_cur = 0;
    g->snap_round = -1;
    k_fill<uint32_t><<<grid_for(g->V), kBlock, 0, s>>>(g->dist, g->V, WT::INF);
    k_init_solve<<<1, 64, 0, s>>>(g->ctrl, g->dist, g->stamp, g->queue[0], g->row_ptr, source, tick0);
    SSSP_LAUNCH_CHECK("collect init");
    const int nt = kModes * kChunkCount;
    // the policy the rounds after the tried one follow: the rule, or (one more
    // step of policy improvement) the network set on the graph
    const int cont = (g->rollout_policy == 1 && (g->solving_bfs ? g->bfs_mlp : g->mlp))
                         ? SSSP_SCHED_LEARNED
                         : SSSP_SCHED_HEURISTIC;
    int32_t row = 0, r = 0, r_hi = 0;
    bool verified = false;
    int64_t relaxed = 0;
    sssp_stats st{};
    g->collect_rounds = 0;
    g->tail_armed = false;
    g->frontier_bits = false;
    auto charge = [&] {
        if (++g->collect_rounds > g->collect_budget) {
            SSSP_HIP(hipStreamSynchronize(s));
            g->tick = tick0 + 2 * g->max_rounds + 2;
            g->trace_events = trace_was;
            throw Error{SSSP_ERR_NOT_CONVERGED,
                        "collect: round budget exhausted (" + std::to_string(g->collect_budget) +
                            " rounds; raise collect_budget)"};
        }
    };
    auto copy = [&](void *dst, const void *src, size_t bytes) {
        SSSP_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, s));
    };
    for (;; r++) {
        if (r >= g->max_rounds) {
            g->tick = tick0 + 2 * g->max_rounds + 2;
            g->trace_events = trace_was;
            throw Error{SSSP_ERR_NOT_CONVERGED, "collect: too many rounds"};
        }
        const int p = g->pile_cur;
        frontier_to_queue(g, r);  // the saved state holds the frontier as a queue
        copy(sd, g->dist, 4 * V);
        copy(sst, g->stamp, 4 * V);
        copy(sfa, g->farst, 4 * V);
        copy(sq, g->queue[r & 1], 4 * (V + 1));
        copy(sp, g->pile[p], 4 * V);
        copy(sc, g->ctrl, sizeof(Ctrl));
        const bool armed = g->tail_armed;
        auto restore = [&] {
            g->snap_round = -1;
            g->tail_armed = armed;
            g->frontier_bits = false;  // queued before the state was saved
            copy(g->dist, sd, 4 * V);
            copy(g->stamp, sst, 4 * V);
            copy(g->farst, sfa, 4 * V);
            copy(g->queue[r & 1], sq, 4 * (V + 1));
            copy(g->pile[p], sp, 4 * V);
            copy(g->ctrl, sc, sizeof(Ctrl));
            g->pile_cur = p;
        };
        float t[kModes * kChunkCount];
        float f[kIn];
        float best = 3.0e38f;
        int bm = -1, bc = 0;
        bool over = false, verify_round = false;
        for (int mode = 0; mode < kModes && !over; mode++) {
            for (int c = 0; c < kChunkCount; c++) {
                float ms = -1.f;
                if (!(mode == SSSP_MODE_PULL && !g->has_csc)) {
                    // best of kRolloutReps timings from the same state
                    for (int rep = 0; rep < kRolloutReps; rep++) {
                        restore();
                        SSSP_HIP(hipEventRecord(g->ev_a, s));
                        bool ver = verified;  // the rollout continues this solve
                        int64_t rel = 0;
                        sssp_stats tmp{};
                        charge();
                        if (!step_round<WT>(g, r, SSSP_SCHED_HEURISTIC, tick0, &tmp, ver, rel,
                                            2 * g->max_rounds, mode, c)) {
                            over = true;  // nothing pending at round r: the solve is complete
                            break;
                        }
                        if (bm < 0 && rep == 0) {
                            for (int i = 0; i < kIn; i++) f[i] = g->ctrl_host->feat[i];
                            // a fixpoint check runs whatever the choice: no label
                            verify_round = g->trace.back().mode == 3;
                        }
                        int32_t r2 = r + 1;
                        for (;;) {
                            if (r2 >= 2 * g->max_rounds) {
                                g->tick = tick0 + 2 * g->max_rounds + 2;
                                g->trace_events = trace_was;
                                throw Error{SSSP_ERR_NOT_CONVERGED,
                                            "collect: rollout did not converge"};
                            }
                            charge();
                            const int32_t used = step_round<WT>(g, r2, cont, tick0, &tmp, ver,
                                                                rel, 2 * g->max_rounds);
                            if (used == 0) break;
                            r2 += used;
                        }
                        r_hi = r2 > r_hi ? r2 : r_hi;
                        SSSP_HIP(hipEventRecord(g->ev_b, s));
                        SSSP_HIP(hipEventSynchronize(g->ev_b));
                        float one = 0.f;
                        SSSP_HIP(hipEventElapsedTime(&one, g->ev_a, g->ev_b));
                        ms = (ms < 0.f || one < ms) ? one : ms;
                    }
                    if (over) break;
                    if (ms < best) {
                        best = ms;
                        bm = mode;
                        bc = c;
                    }
                }
                t[mode * kChunkCount + c] = ms;
            }
        }
        restore();
        if (over) break;
        charge();
        if (!step_round<WT>(g, r, SSSP_SCHED_HEURISTIC, tick0, &st, verified, relaxed,
                            2 * g->max_rounds, bm, bc))
            break;
        if (!verify_round && row < cap) {
            for (int i = 0; i < kIn; i++) feats[(size_t)row * kIn + i] = f[i];
            for (int i = 0; i < nt; i++) times[(size_t)row * nt + i] = t[i];
            row++;
        }
    }
    SSSP_HIP(hipStreamSynchronize(s));
    g->trace.clear();
    g->trace_events = trace_was;
    g->pile_dirty = false;
    g->tick = tick0 + (r_hi > r ? r_hi : r) + 2;
    *rows = row;
}

}  // namespace

void reset_stamps(sssp_graph *g)
{
    k_stamp_init<<<grid_for(g->V), kBlock, 0, g->stream>>>(g->stamp, g->row_ptr, g->V);
    SSSP_LAUNCH_CHECK("k_stamp_init");
    k_fill<int32_t><<<grid_for(g->V), kBlock, 0, g->stream>>>(g->farst, g->V, 0);
    SSSP_LAUNCH_CHECK("pile state reset");
    g->pile_dirty = false;
}

void sample_in_degrees(sssp_graph *g)
{
    dfree(g->sample_in);
    g->sample_in = nullptr;
    if (!g->has_csc || g->n_sample <= 0) return;
    g->sample_in = dmalloc<int32_t>((size_t)g->n_sample);
    k_sample_in<<<(unsigned)((g->n_sample + kBlock - 1) / kBlock), kBlock, 0, g->stream>>>(
        g->sample_v, g->n_sample, g->col_ptr, g->sample_in);
    SSSP_LAUNCH_CHECK("k_sample_in");
}

void unpermute_rows(const uint32_t *in, const int32_t *vperm, int64_t V, int64_t rows,
                    uint32_t *out, hipStream_t s)
{
    if (V * rows == 0) return;
    k_unpermute<<<grid_for(V * rows), kBlock, 0, s>>>(in, vperm, V, rows, out);
    SSSP_LAUNCH_CHECK("k_unpermute");
}

void parents_out(sssp_graph *g, int32_t *parent_host)
{
    hipStream_t s = g->stream;
    if (!g->vperm) {
        SSSP_HIP(hipMemcpyAsync(parent_host, g->parent, 4 * (size_t)g->V, hipMemcpyDeviceToHost, s));
        SSSP_HIP(hipStreamSynchronize(s));
        return;
    }
    int32_t *tmp = dmalloc<int32_t>((size_t)g->V);
    struct Free {
        void *p;
        ~Free() { dfree(p); }
    } fr{tmp};
    k_unpermute_parents<<<grid_for(g->V), kBlock, 0, s>>>(g->parent, g->vperm, g->vinv, g->V, tmp);
    SSSP_LAUNCH_CHECK("k_unpermute_parents");
    SSSP_HIP(hipMemcpyAsync(parent_host, tmp, 4 * (size_t)g->V, hipMemcpyDeviceToHost, s));
    SSSP_HIP(hipStreamSynchronize(s));
}

void solve(sssp_graph *g, const int32_t *sources, int32_t n_src, int sched, sssp_stats *st)
{
    if (g->wtype == SSSP_W_F32) solve_t<WF32>(g, sources, n_src, sched, st);
    else solve_t<WU32>(g, sources, n_src, sched, st);
}

// Marks the rounds in flight as a BFS (its own scheduler network) for the
// lifetime of the guard.
struct BfsScope {
    sssp_graph *g;
    explicit BfsScope(sssp_graph *gg) : g(gg) { g->solving_bfs = true; }
    ~BfsScope() { g->solving_bfs = false; }
};

void bfs(sssp_graph *g, int32_t source, int sched, sssp_stats *st)
{
    BfsScope scope(g);
    solve_t<WUnit>(g, &source, 1, sched, st);
}

void certify_levels(sssp_graph *g, int32_t source, int64_t *counts)
{
    certify_t<WUnit>(g, &source, 1, counts);
}

void collect_timings(sssp_graph *g, int32_t source, int32_t cap, float *feats, float *times,
                     int32_t *rows)
{
    if (g->collect_kind == 1) {
        BfsScope scope(g);
        collect_t<WUnit>(g, source, cap, feats, times, rows);
    } else if (g->wtype == SSSP_W_F32) {
        collect_t<WF32>(g, source, cap, feats, times, rows);
    } else {
        collect_t<WU32>(g, source, cap, feats, times, rows);
    }
}

void certify(sssp_graph *g, const int32_t *sources, int32_t n_src, int64_t *counts)
{
    if (g->wtype == SSSP_W_F32) certify_t<WF32>(g, sources, n_src, counts);
    else certify_t<WU32>(g, sources, n_src, counts);
}

// ---------------------------------------------------------------------------
// Reference-shaped wrappers over caller-owned device arrays.
// ---------------------------------------------------------------------------
template <class WT, class DT>
void ref_init(DT *d, int32_t V, int32_t source)
{
    k_fill<DT><<<grid_for(V), kBlock>>>(d, V, WT::INF);
    SSSP_LAUNCH_CHECK("init_distances");
    k_fill<DT><<<1, 64>>>(d + source, 1, (DT)0);
    SSSP_LAUNCH_CHECK("init_distances source");
    SSSP_HIP(hipDeviceSynchronize());
}

template <class WT, class DT>
void ref_bf(const int32_t *src, const int32_t *dst, const DT *w, int32_t V, int64_t E, DT *d,
            int32_t *passes_out, int32_t *changed_buf)
{
    int dev = 0;
    SSSP_HIP(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    SSSP_HIP(hipGetDeviceProperties(&prop, dev));
    int32_t *changed = changed_buf ? changed_buf : dmalloc<int32_t>(64);
    int32_t hc[64];
    struct Free {
        void *p;
        ~Free() { dfree(p); }
    } fr{changed_buf ? nullptr : changed};
    // pass 0 reads no flag and sets flag 0 only on a change: all 64 start at 0
    k_fill<int32_t><<<1, 64>>>(changed, 64, 0);
    SSSP_LAUNCH_CHECK("clear change flags");
    int64_t nblocks = (int64_t)prop.multiProcessorCount * 8;
    const bool vec = sizeof(DT) == 4 && (reinterpret_cast<uintptr_t>(src) % 16 == 0) &&
                     (reinterpret_cast<uintptr_t>(dst) % 16 == 0) &&
                     (reinterpret_cast<uintptr_t>(w) % 16 == 0);
    const int64_t units = vec ? (E >> 2) : E;
    const int64_t per_block = (units + nblocks - 1) / nblocks;
    int32_t pass = 0, done_at = -1;
    while (done_at < 0) {
        for (int i = 0; i < 4; i++) {
            if (vec)
                k_el_pass<WT, DT, 4, false><<<(unsigned)nblocks, kBlock>>>(
                    src, dst, w, E, d, changed, pass + i, per_block, nullptr);
            else
                k_el_pass<WT, DT, 1, false><<<(unsigned)nblocks, kBlock>>>(
                    src, dst, w, E, d, changed, pass + i, per_block, nullptr);
            SSSP_LAUNCH_CHECK("bellman_ford pass");
        }
        SSSP_HIP(hipMemcpy(hc, changed, sizeof(hc), hipMemcpyDeviceToHost));
        for (int i = 0; i < 4; i++)
            if (hc[(pass + i) & 63] == 0) {
                done_at = pass + i;
                break;
            }
        pass += 4;
        // the reference loops at most V times (bellman_ford.cu:146)
        if (done_at < 0 && pass >= V + 4)
            throw Error{SSSP_ERR_NOT_CONVERGED, "Bellman-Ford did not converge within V passes"};
    }
    if (passes_out) *passes_out = done_at + 1;
}

__global__ __launch_bounds__(kBlock) void k_levels_to_visited(const uint32_t *lev, int64_t V,
                                                              uint8_t *vis)
{
    for (int64_t v = (int64_t)blockIdx.x * kBlock + threadIdx.x; v < V;
         v += (int64_t)gridDim.x * kBlock)
        vis[v] = lev[v] != WUnit::INF ? 1 : 0;
}

void levels_to_visited(const uint32_t *lev, int32_t V, uint8_t *vis, hipStream_t s)
{
    k_levels_to_visited<<<grid_for(V), kBlock, 0, s>>>(lev, V, vis);
    SSSP_LAUNCH_CHECK("k_levels_to_visited");
}

// The reference's bfs_wrapper (bfs.cu:44-85) over caller-owned device edge
// arrays: all-edge passes until one marks nothing. The passes run on levels
// (unit-weight Bellman-Ford, the dst array stands in for the unused weights)
// and end in the visited flags, 1 for every vertex reachable from source.
void ref_bfs(const int32_t *src, const int32_t *dst, int32_t V, int64_t E, uint8_t *visited,
             int32_t source, int32_t *passes_out)
{
    uint32_t *lev = dmalloc<uint32_t>((size_t)V);
    struct Free {
        void *p;
        ~Free() { dfree(p); }
    } fr{lev};
    ref_init<WUnit, uint32_t>(lev, V, source);
    ref_bf<WUnit, uint32_t>(src, dst, reinterpret_cast<const uint32_t *>(dst), V, E, lev,
                            passes_out);
    levels_to_visited(lev, V, visited, nullptr);
    SSSP_HIP(hipDeviceSynchronize());
}

// ---------------------------------------------------------------------------
// Scratch of the allocation-free f32 wrapper: the 64-slot ring of change flags
// (bytes 0-255) and the two validation counters (bytes 256-271), in a
// file-scope device array instead of memory allocated per call. The scratch
// does not grow with the graph, so every size the allocating wrapper accepts
// fits. Each device holds its own copy of the array at its own address, looked
// up once per device (with that device current) by hipGetSymbolAddress; host
// code passes that address to kernels and never the array's name, which in
// host code only stands in for the symbol. Calls on one device share its
// array: calls that overlap on one device race on the flags.
// ---------------------------------------------------------------------------
constexpr size_t kNoAllocScratchBytes = 512;
constexpr size_t kScratchFlags = 0;  // int32_t[64]
constexpr size_t kScratchBad = 256;  // int64_t[2]
static_assert(kScratchBad + 2 * sizeof(int64_t) <= kNoAllocScratchBytes, "scratch layout");
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

constexpr int kScratchDevices = 64;
static std::mutex g_scratch_mu;
static unsigned char *g_scratch_ptr[kScratchDevices];

static unsigned char *scratch_on_current_device()
{
    int dev = 0;
    SSSP_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kScratchDevices)
        throw Error{SSSP_ERR_ARG, "device ordinal " + std::to_string(dev) +
                                      " beyond the allocation-free wrapper's table of " +
                                      std::to_string(kScratchDevices)};
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            throw Error{SSSP_ERR_HIP, std::string("hipGetSymbolAddress(scratch) failed: ") +
                                          hipGetErrorString(e)};
        g_scratch_ptr[dev] = static_cast<unsigned char *>(p);
    }
    return g_scratch_ptr[dev];
}

void ref_bf_noalloc_f32(const int32_t *src, const int32_t *dst, const uint32_t *w, int32_t V,
                        int64_t E, uint32_t *d, int32_t *passes_out)
{
    unsigned char *p = scratch_on_current_device();
    validate_edges(src, dst, w, SSSP_W_F32, E, V, nullptr,
                   reinterpret_cast<int64_t *>(p + kScratchBad));
    ref_bf<WF32, uint32_t>(src, dst, w, V, E, d, passes_out,
                           reinterpret_cast<int32_t *>(p + kScratchFlags));
}

template void ref_init<WF32, uint32_t>(uint32_t *, int32_t, int32_t);
template void ref_init<WU32, uint32_t>(uint32_t *, int32_t, int32_t);
template void ref_init<WF64, unsigned long long>(unsigned long long *, int32_t, int32_t);
template void ref_bf<WF32, uint32_t>(const int32_t *, const int32_t *, const uint32_t *, int32_t,
                                     int64_t, uint32_t *, int32_t *, int32_t *);
template void ref_bf<WU32, uint32_t>(const int32_t *, const int32_t *, const uint32_t *, int32_t,
                                     int64_t, uint32_t *, int32_t *, int32_t *);
template void ref_bf<WF64, unsigned long long>(const int32_t *, const int32_t *,
                                               const unsigned long long *, int32_t, int64_t,
                                               unsigned long long *, int32_t *, int32_t *);

}  // namespace sssp
