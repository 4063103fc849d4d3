// layout.hip -- builds the device layouts of a graph in HBM:
//   * the edge list sorted by src, stable (el_src/el_dst/el_w) with row
//     offsets row_ptr[V+1] (int64, as compressed_adjacency_list.h:13 keeps
//     `long long *edge_ptrs`) -- this is the CSR of the reference
//     (graph.hpp:316-361: per-source order = edge-list order);
//   * the CSC (in-edges grouped by dst, stable) for pull rounds;
//   * static lists of vertices with out-/in-degree > 0 and their tile tables.
// Sorting is a stable LSD radix sort on the vertex id with the original edge
// index as payload: per 4096-key tile an LDS digit histogram, one device-wide
// scan of the (digit, tile) counts, then a scatter that ranks keys inside the
// tile with 8 ballots per wave (stable by construction).
#include "sssp_graph.hpp"

#include <algorithm>

namespace sssp {

namespace {

constexpr int kScanItems = 8;
constexpr int kScanTile = kBlock * kScanItems;  // 2048 elements per block
constexpr int kRsItems = 16;
constexpr int kRsTile = kBlock * kRsItems;  // 4096 keys per tile

// --------------------------------------------------------------------------
// Generic exclusive scan of int64, three phases.
// --------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_scan_reduce(const int64_t *in, int64_t n, int64_t *blk)
{
    __shared__ int64_t s_warp[kBlock / 64];
    const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
    int64_t s = 0;
#pragma unroll
    for (int i = 0; i < kScanItems; i++)
        if (base + i < n) s += in[base + i];
    int64_t tot;
    (void)block_excl_scan_i64(s, &tot, s_warp);
    if (threadIdx.x == 0) blk[blockIdx.x] = tot;
}

// single block: exclusive scan of blk[0..nb) in place, blk[nb] = total
__device__ __forceinline__ void scan_blocks_body(int64_t *blk, int64_t nb)
{
    __shared__ int64_t s_warp[kBlock / 64];
    int64_t carry = 0;
    for (int64_t base = 0; base < nb; base += kBlock) {
        int64_t i = base + threadIdx.x;
        int64_t x = i < nb ? blk[i] : 0;
        int64_t tot;
        int64_t ex = block_excl_scan_i64(x, &tot, s_warp);
        if (i < nb) blk[i] = carry + ex;
        carry += tot;
    }
    if (threadIdx.x == 0) blk[nb] = carry;
}

__global__ __launch_bounds__(kBlock) void k_scan_blocks(int64_t *blk, int64_t nb)
{
    scan_blocks_body(blk, nb);
}

// two arrays of nb + 1 values in one launch: workgroup 0 scans a, 1 scans b
__global__ __launch_bounds__(kBlock) void k_scan_blocks_pair(int64_t *a, int64_t *b, int64_t nb)
{
    scan_blocks_body(blockIdx.x == 0 ? a : b, nb);
}

__global__ __launch_bounds__(kBlock) void k_scan_down(const int64_t *in, int64_t *out, int64_t n,
                                                      const int64_t *blk)
{
    __shared__ int64_t s_warp[kBlock / 64];
    const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
    int64_t v[kScanItems];
    int64_t s = 0;
#pragma unroll
    for (int i = 0; i < kScanItems; i++) {
        v[i] = base + i < n ? in[base + i] : 0;
        s += v[i];
    }
    int64_t tot;
    int64_t run = block_excl_scan_i64(s, &tot, s_warp) + blk[blockIdx.x];
#pragma unroll
    for (int i = 0; i < kScanItems; i++) {
        if (base + i < n) out[base + i] = run;
        run += v[i];
    }
}

// --------------------------------------------------------------------------
// Input validation: ids in [0, V), weights >= 0 and not NaN.
// --------------------------------------------------------------------------
// ---------------------------------------------------------------------------
// Vertex order (slots by descending in- plus out-degree).
// ---------------------------------------------------------------------------
// tot[v] = in- plus out-degree, out[v] = out-degree
__global__ __launch_bounds__(kBlock) void k_degrees(const int32_t *src, const int32_t *dst,
                                                    int64_t E, uint32_t *tot, uint32_t *out)
{
    for (int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x; e < E;
         e += (int64_t)gridDim.x * kBlock) {
        const int32_t s = src[e];
        atomicAdd(&tot[s], 1u);
        atomicAdd(&out[s], 1u);
        atomicAdd(&tot[dst[e]], 1u);
    }
}

__global__ __launch_bounds__(kBlock) void k_max_u32(const uint32_t *x, int64_t n, uint32_t *mx)
{
    uint32_t m = 0;
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * kBlock)
        m = max(m, x[i]);
    for (int o = 32; o > 0; o >>= 1) m = max(m, (uint32_t)__shfl_xor(m, o));
    if (__lane_id() == 0 && m) atomicMax(mx, m);
}

// ascending sort key = descending count (counts above kOrderCap tie), of
// vertex at[i] (or of vertex i when `at` is null)
constexpr uint32_t kOrderCap = 1u << 30;
__global__ __launch_bounds__(kBlock) void k_order_keys(const uint32_t *cnt, const uint32_t *at,
                                                       int64_t V, uint32_t *key)
{
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < V;
         i += (int64_t)gridDim.x * kBlock)
        key[i] = kOrderCap - min(cnt[at ? at[i] : i], kOrderCap);
}

// slot i holds vertex by_out[pos[i]] -> vinv, and its inverse vperm
__global__ __launch_bounds__(kBlock) void k_invert_order(const uint32_t *by_out,
                                                         const uint32_t *pos, int64_t V,
                                                         int32_t *vperm, int32_t *vinv)
{
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < V;
         i += (int64_t)gridDim.x * kBlock) {
        const int32_t v = (int32_t)by_out[pos[i]];
        vinv[i] = v;
        vperm[v] = (int32_t)i;
    }
}

__global__ __launch_bounds__(kBlock) void k_relabel_edges(const int32_t *src, const int32_t *dst,
                                                          int64_t E, const int32_t *vperm,
                                                          int32_t *osrc, int32_t *odst)
{
    for (int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x; e < E;
         e += (int64_t)gridDim.x * kBlock) {
        osrc[e] = vperm[src[e]];
        odst[e] = vperm[dst[e]];
    }
}

__global__ void k_zero_i64(int64_t *p, int n)
{
    if ((int)threadIdx.x < n) p[threadIdx.x] = 0;
}

template <class W>
__global__ __launch_bounds__(kBlock) void k_validate(const int32_t *src, const int32_t *dst,
                                                     const W *w, int64_t E, int32_t V,
                                                     int64_t *bad)
{
    int64_t nb_id = 0, nb_w = 0;
    for (int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x; e < E;
         e += (int64_t)gridDim.x * kBlock) {
        int32_t s = src[e], d = dst[e];
        if (s < 0 || s >= V || d < 0 || d >= V) nb_id++;
        if (w) {
            const W x = w[e];
            if (!(x >= (W)0)) nb_w++;  // negative or NaN
        }
    }
    nb_id = wave_sum64(nb_id);
    nb_w = wave_sum64(nb_w);
    if (__lane_id() == 0) {
        if (nb_id) atomicAdd((unsigned long long *)&bad[0], (unsigned long long)nb_id);
        if (nb_w) atomicAdd((unsigned long long *)&bad[1], (unsigned long long)nb_w);
    }
}

// --------------------------------------------------------------------------
// Stable LSD radix sort of (key, value) by 8-bit digits.
// counts layout: counts[digit * ntiles + tile] (digit-major) so that one
// exclusive scan yields every tile's output base per digit, in tile order.
// --------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_rs_hist(const uint32_t *keys, int64_t n, int shift,
                                                    int64_t *counts, int64_t ntiles)
{
    __shared__ uint32_t h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    const int64_t base = (int64_t)blockIdx.x * kRsTile;
#pragma unroll 4
    for (int i = 0; i < kRsItems; i++) {
        int64_t idx = base + i * kBlock + threadIdx.x;
        if (idx < n) atomicAdd(&h[(keys[idx] >> shift) & 255u], 1u);
    }
    __syncthreads();
    counts[(int64_t)threadIdx.x * ntiles + blockIdx.x] = h[threadIdx.x];
}

__global__ __launch_bounds__(kBlock) void k_rs_scatter(const uint32_t *keys_in,
                                                       const uint32_t *vals_in,
                                                       uint32_t *keys_out, uint32_t *vals_out,
                                                       int64_t n, int shift,
                                                       const int64_t *offsets, int64_t ntiles)
{
    __shared__ int64_t s_off[256];
    __shared__ uint32_t s_cnt[kBlock / 64][256];
    const int tid = threadIdx.x, lane = __lane_id(), wv = tid >> 6;
    s_off[tid] = offsets[(int64_t)tid * ntiles + blockIdx.x];
#pragma unroll
    for (int w = 0; w < kBlock / 64; w++) s_cnt[w][tid] = 0;
    __syncthreads();
    const int64_t base = (int64_t)blockIdx.x * kRsTile;
    for (int r = 0; r < kRsItems; r++) {
        const int64_t idx = base + (int64_t)r * kBlock + tid;
        const bool valid = idx < n;
        uint32_t key = valid ? keys_in[idx] : 0u;
        uint32_t val = valid ? (vals_in ? vals_in[idx] : (uint32_t)idx) : 0u;
        uint32_t dg = (key >> shift) & 255u;
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; b++) {
            bool bit = (dg >> b) & 1u;
            uint64_t bm = __ballot(bit);
            peers &= bit ? bm : ~bm;
        }
        const uint32_t rank = __popcll(peers & lanemask_lt());
        if (valid && (peers >> lane) == 1ull) s_cnt[wv][dg] = __popcll(peers);
        __syncthreads();
        if (valid) {
            uint32_t pre = 0;
            for (int w = 0; w < wv; w++) pre += s_cnt[w][dg];
            int64_t pos = s_off[dg] + pre + rank;
            keys_out[pos] = key;
            vals_out[pos] = val;
        }
        __syncthreads();
        uint32_t add = 0;
#pragma unroll
        for (int w = 0; w < kBlock / 64; w++) {
            add += s_cnt[w][tid];
            s_cnt[w][tid] = 0;
        }
        s_off[tid] += add;
        __syncthreads();
    }
}

// --------------------------------------------------------------------------
// Layout assembly kernels.
// --------------------------------------------------------------------------
template <class W>
__global__ __launch_bounds__(kBlock) void k_gather_pair(const uint32_t *perm, int64_t E,
                                                        const int32_t *ids, const W *w,
                                                        int32_t *ids_out, uint32_t *w_out)
{
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < E;
         i += (int64_t)gridDim.x * kBlock) {
        uint32_t p = perm[i];
        ids_out[i] = ids[p];
        W x = w[p];
        uint32_t b;
        if constexpr (sizeof(W) == 4) {
            b = *reinterpret_cast<const uint32_t *>(&x);
        } else {
            b = 0;
        }
        w_out[i] = b;
    }
}

__global__ __launch_bounds__(kBlock) void k_copy_i32(const uint32_t *in, int32_t *out, int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * kBlock)
        out[i] = (int32_t)in[i];
}

// XCD group of an edge: a multiplicative hash of its destination's 64-vertex
// block, so each XCD's L2 serves a disjoint, work-balanced eighth of the
// distance array (R-MAT's per-bit skew makes any bit-sliced split uneven).
__device__ __forceinline__ uint32_t xcd_group(int32_t dst)
{
    return (((uint32_t)dst >> 6) * 0x9E3779B1u) >> 29;
}

__global__ __launch_bounds__(kBlock) void k_group_keys(const int32_t *dst, int64_t E, uint32_t *keys,
                                                       int64_t *gcount)
{
    __shared__ int64_t h[8];
    if (threadIdx.x < 8) h[threadIdx.x] = 0;
    __syncthreads();
    for (int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x; e < E;
         e += (int64_t)gridDim.x * kBlock) {
        const uint32_t k = xcd_group(dst[e]);
        keys[e] = k;
        atomicAdd((unsigned long long *)&h[k], 1ull);
    }
    __syncthreads();
    if (threadIdx.x < 8 && h[threadIdx.x])
        atomicAdd((unsigned long long *)&gcount[threadIdx.x], (unsigned long long)h[threadIdx.x]);
}

// out[goff[x] + j] = in[perm[cbase[x] + j]] for j < count[x]; the rest of
// each group's region (up to 3 slots, so every group is a whole number of
// 16-byte vectors) is a weight-0 self loop on vertex 0, which never changes
// a distance.
__global__ __launch_bounds__(kBlock) void k_group_gather(const uint32_t *perm, const int64_t *goff,
                                                         const int64_t *cbase, const int32_t *src,
                                                         const int32_t *dst, const uint32_t *w,
                                                         int32_t *xs, int32_t *xd, uint32_t *xw)
{
    const int64_t Et = goff[8];
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < Et;
         i += (int64_t)gridDim.x * kBlock) {
        int x = 0;
        while (x < 7 && i >= goff[x + 1]) x++;
        const int64_t j = i - goff[x];
        const int64_t cnt = cbase[x + 1] - cbase[x];
        if (j < cnt) {
            const uint32_t p = perm[cbase[x] + j];
            xs[i] = src[p];
            xd[i] = dst[p];
            xw[i] = w[p];
        } else {
            xs[i] = 0;
            xd[i] = 0;
            xw[i] = 0u;
        }
    }
}

__global__ __launch_bounds__(kBlock) void k_blk_keys(const int32_t *dst, int64_t E, uint32_t *keys)
{
    for (int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x; e < E;
         e += (int64_t)gridDim.x * kBlock)
        keys[e] = (uint32_t)dst[e] >> kBlkBits;
}

// Sorted position j of block b = skeys[j] goes to boff[b] + (j - bptr[b]);
// perm[j] is its position in the src-sorted list, whose order is kept.
__global__ __launch_bounds__(kBlock) void k_blk_gather(const uint32_t *skeys, const uint32_t *perm,
                                                       int64_t E, const int64_t *bptr,
                                                       const int64_t *boff, const int32_t *src,
                                                       const int32_t *dst, const uint32_t *w,
                                                       int32_t *bs, uint16_t *bd, uint32_t *bw)
{
    for (int64_t j = (int64_t)blockIdx.x * kBlock + threadIdx.x; j < E;
         j += (int64_t)gridDim.x * kBlock) {
        const uint32_t b = skeys[j];
        const int64_t o = boff[b] + (j - bptr[b]);
        const uint32_t p = perm[j];
        bs[o] = src[p];
        bd[o] = (uint16_t)(dst[p] - (int32_t)(b << kBlkBits));
        bw[o] = w[p];
    }
}

// The padding slots of each block (up to kBlkChunk - 1) copy the block's last
// edge: relaxing it again changes nothing, and its source delta is 0.
__global__ __launch_bounds__(kBlock) void k_blk_pad(const int64_t *bptr, const int64_t *boff,
                                                    int32_t nblk, int32_t *bs, uint16_t *bd,
                                                    uint32_t *bw)
{
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < (int64_t)nblk * kBlkChunk;
         i += (int64_t)gridDim.x * kBlock) {
        const int32_t b = (int32_t)(i / kBlkChunk);
        const int64_t last = boff[b] + (bptr[b + 1] - bptr[b]) - 1;
        const int64_t o = last + 1 + i % kBlkChunk;
        if (o < boff[b + 1]) {  // only blocks with edges have slots
            bs[o] = bs[last];
            bd[o] = bd[last];
            bw[o] = bw[last];
        }
    }
}

// Source deltas of one chunk (one wave, 4 slots per lane, in slot order):
// d[k] = source of slot k - source of the slot before it, 0 for the first.
__device__ __forceinline__ void chunk_deltas(const int32_t *bs, int64_t c, int32_t (&s)[4],
                                             uint32_t (&d)[4])
{
    typedef int32_t v4i __attribute__((ext_vector_type(4)));
    const v4i v = reinterpret_cast<const v4i *>(bs)[c * (kBlkChunk / 4) + __lane_id()];
    s[0] = v.x;
    s[1] = v.y;
    s[2] = v.z;
    s[3] = v.w;
    const int32_t prev = __shfl_up(s[3], 1);
    d[0] = __lane_id() == 0 ? 0u : (uint32_t)(s[0] - prev);
    d[1] = (uint32_t)(s[1] - s[0]);
    d[2] = (uint32_t)(s[2] - s[1]);
    d[3] = (uint32_t)(s[3] - s[2]);
}

// wide[c] = 1 when chunk c has a source gap that does not fit the packed
// delta (sources ascend within a block, so every delta is >= 0)
__global__ __launch_bounds__(kBlock) void k_blk_chunk_wide(const int32_t *bs, int64_t nch,
                                                           uint8_t *wide)
{
    const int64_t c = ((int64_t)blockIdx.x * kBlock + threadIdx.x) >> 6;
    if (c >= nch) return;  // wave-uniform
    int32_t s[4];
    uint32_t d[4];
    chunk_deltas(bs, c, s, d);
    bool big = false;
    for (int k = 0; k < 4; k++) big |= d[k] >= (1u << kBlkDeltaBits);
    const uint64_t m = __ballot(big);
    if (__lane_id() == 0) wide[c] = m != 0;
}

// Packs chunk c: widx[c] < 0 -> deltas in the words and the first source in
// cbase; else the destinations alone in the words, the sources whole in
// wide[widx[c] * kBlkChunk ...] and cbase = -(1 + widx[c]).
__global__ __launch_bounds__(kBlock) void k_blk_pack(const int32_t *bs, const uint16_t *bd,
                                                     const int32_t *widx, int64_t nch,
                                                     uint32_t *pk, int32_t *cbase, int32_t *wide)
{
    const int64_t c = ((int64_t)blockIdx.x * kBlock + threadIdx.x) >> 6;
    if (c >= nch) return;  // wave-uniform
    int32_t s[4];
    uint32_t d[4];
    chunk_deltas(bs, c, s, d);
    const int32_t wi = widx[c];
    const int64_t o = c * kBlkChunk + 4 * __lane_id();
    for (int k = 0; k < 4; k++)
        pk[o + k] = (wi < 0 ? d[k] << kBlkBits : 0u) | (uint32_t)bd[o + k];
    if (wi >= 0)
        for (int k = 0; k < 4; k++) wide[(int64_t)wi * kBlkChunk + 4 * __lane_id() + k] = s[k];
    if (__lane_id() == 0) cbase[c] = wi < 0 ? s[0] : -(1 + wi);
}

// off[b * (nfine + 1) + k] = the first slot of block b whose source is at
// least k << fine_bits (sources ascend within a block's padded run), aligned
// down to its chunk; k = nfine gives the end of the run.
__global__ __launch_bounds__(kBlock) void k_blk_range_offsets(const int32_t *bs, const int64_t *boff,
                                                              int32_t nblk, int fine_bits,
                                                              int32_t nfine, int64_t *off)
{
    const int64_t n = (int64_t)nblk * (nfine + 1);
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * kBlock) {
        const int32_t b = (int32_t)(i / (nfine + 1));
        const int32_t k = (int32_t)(i % (nfine + 1));
        const int64_t b0 = boff[b], b1 = boff[b + 1];
        int64_t pos = b1;
        if (k < nfine) {
            const int64_t key = (int64_t)k << fine_bits;
            int64_t lo = b0, hi = b1;  // first position with bs >= key
            while (lo < hi) {
                const int64_t mid = (lo + hi) >> 1;
                if ((int64_t)bs[mid] < key) lo = mid + 1;
                else hi = mid;
            }
            pos = b0 + (lo - b0) / kBlkChunk * kBlkChunk;
        }
        off[i] = pos;
    }
}

__global__ __launch_bounds__(kBlock) void k_m_keys(const int32_t *dst, int64_t E, uint32_t *keys)
{
    for (int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x; e < E;
         e += (int64_t)gridDim.x * kBlock)
        keys[e] = (uint32_t)dst[e] >> kMBits;
}

// record j = the edge at sorted position j (perm: its position in the
// src-sorted list, whose order is kept within a block)
__global__ __launch_bounds__(kBlock) void k_m_gather(const uint32_t *skeys, const uint32_t *perm,
                                                     int64_t E, const int32_t *src,
                                                     const int32_t *dst, const uint32_t *w,
                                                     MRec *out)
{
    for (int64_t j = (int64_t)blockIdx.x * kBlock + threadIdx.x; j < E;
         j += (int64_t)gridDim.x * kBlock) {
        const uint32_t p = perm[j];
        out[j] = MRec{src[p], w[p], dst[p] - (int32_t)(skeys[j] << kMBits), 0};
    }
}

// row_ptr[v] = first position whose sorted key is >= v; row_ptr[V] = E.
// Position i starts the vertices (keys[i-1], keys[i]]; it writes them when
// they are few. A long run of keyless vertices (on a degree-ordered graph
// every isolated vertex sits at the end) is left at -1 for
// k_rowptr_search, so no thread loops over millions of vertices.
constexpr int64_t kRowptrRun = 64;
__global__ __launch_bounds__(kBlock) void k_rowptr_from_sorted(const uint32_t *keys, int64_t E,
                                                               int32_t V, int64_t *row_ptr)
{
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i <= E;
         i += (int64_t)gridDim.x * kBlock) {
        int64_t lo = (i == 0) ? 0 : (int64_t)keys[i - 1] + 1;  // first vertex not yet placed
        int64_t hi = (i == E) ? (int64_t)V : (int64_t)keys[i];  // last vertex starting at i
        if (hi - lo < kRowptrRun)
            for (int64_t v = lo; v <= hi; v++) row_ptr[v] = i;
    }
}

// the entries k_rowptr_from_sorted left at -1: lower bound of v in the keys
__global__ __launch_bounds__(kBlock) void k_rowptr_search(const uint32_t *keys, int64_t E,
                                                          int32_t V, int64_t *row_ptr)
{
    for (int64_t v = (int64_t)blockIdx.x * kBlock + threadIdx.x; v <= V;
         v += (int64_t)gridDim.x * kBlock) {
        if (row_ptr[v] >= 0) continue;
        int64_t lo = 0, hi = E;  // first position with key >= v
        while (lo < hi) {
            const int64_t mid = (lo + hi) >> 1;
            if ((int64_t)keys[mid] < v) lo = mid + 1;
            else hi = mid;
        }
        row_ptr[v] = lo;
    }
}

__global__ __launch_bounds__(kBlock) void k_fill_i64(int64_t *p, int64_t n, int64_t val)
{
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * kBlock)
        p[i] = val;
}


__global__ __launch_bounds__(kBlock) void k_nz_flags(const int64_t *ptr, int32_t V, int64_t *flags)
{
    for (int64_t v = (int64_t)blockIdx.x * kBlock + threadIdx.x; v < V;
         v += (int64_t)gridDim.x * kBlock)
        flags[v] = ptr[v + 1] > ptr[v] ? 1 : 0;
}

__global__ __launch_bounds__(kBlock) void k_nz_compact(const int64_t *ptr, int32_t V,
                                                       const int64_t *pos, int32_t *nz_v,
                                                       int64_t *nz_scan, int64_t n_nz, int64_t E)
{
    for (int64_t v = (int64_t)blockIdx.x * kBlock + threadIdx.x; v < V;
         v += (int64_t)gridDim.x * kBlock) {
        if (ptr[v + 1] > ptr[v]) {
            int64_t j = pos[v];
            nz_v[j] = (int32_t)v;
            nz_scan[j] = ptr[v];
        }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) nz_scan[n_nz] = E;
}

// tile_out[t] = the entry whose edge range contains edge t * TE; tile_out[ntiles] = n.
__global__ __launch_bounds__(kBlock) void k_tile_table(const int64_t *scan, int32_t n, int64_t TE,
                                                       int64_t ntiles, int32_t *tile_out)
{
    for (int64_t j = (int64_t)blockIdx.x * kBlock + threadIdx.x; j < n;
         j += (int64_t)gridDim.x * kBlock) {
        int64_t a = scan[j], b = scan[j + 1];
        for (int64_t t = (a + TE - 1) / TE; t * TE < b; t++) tile_out[t] = (int32_t)j;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) tile_out[ntiles] = n;
}

inline int grid_for(int64_t n, int64_t cap = 8192)
{
    int64_t b = (n + kBlock - 1) / kBlock;
    if (b < 1) b = 1;
    return (int)(b < cap ? b : cap);
}

void rowptr_from_sorted(const uint32_t *keys, int64_t E, int32_t V, int64_t *row_ptr,
                        hipStream_t s)
{
    k_fill_i64<<<grid_for((int64_t)V + 1), kBlock, 0, s>>>(row_ptr, (int64_t)V + 1, -1);
    k_rowptr_from_sorted<<<grid_for(E + 1), kBlock, 0, s>>>(keys, E, V, row_ptr);
    k_rowptr_search<<<grid_for((int64_t)V + 1), kBlock, 0, s>>>(keys, E, V, row_ptr);
    SSSP_LAUNCH_CHECK("rowptr_from_sorted");
}

// Sorts by key; returns the permutation (original indices) in *perm_out
// (one of the two value buffers) and sorted keys in *keys_out.
void radix_sort_pairs(const int32_t *keys_in, int64_t n, int32_t max_key, uint32_t *kA,
                      uint32_t *kB, uint32_t *vA, uint32_t *vB, int64_t *counts, int64_t *blk_a,
                      int64_t *blk_b, int64_t n_blk, hipStream_t s, uint32_t **keys_out,
                      uint32_t **perm_out)
{
    int bits = 0;
    while (bits < 31 && ((int64_t)1 << bits) <= (int64_t)max_key) bits++;
    int passes = (bits + 7) / 8;
    if (passes < 1) passes = 1;
    const int64_t ntiles = (n + kRsTile - 1) / kRsTile;
    const uint32_t *kin = reinterpret_cast<const uint32_t *>(keys_in);
    const uint32_t *vin = nullptr;  // iota on the first pass
    uint32_t *kout = kA, *vout = vA;
    for (int p = 0; p < passes; p++) {
        const int shift = 8 * p;
        if (ntiles > 0) {
            k_rs_hist<<<(unsigned)ntiles, kBlock, 0, s>>>(kin, n, shift, counts, ntiles);
            SSSP_LAUNCH_CHECK("k_rs_hist");
            device_exclusive_scan_i64(counts, counts, 256 * ntiles, blk_a, blk_b, n_blk, s);
            k_rs_scatter<<<(unsigned)ntiles, kBlock, 0, s>>>(kin, vin, kout, vout, n, shift,
                                                              counts, ntiles);
            SSSP_LAUNCH_CHECK("k_rs_scatter");
        }
        kin = kout;
        vin = vout;
        kout = (kout == kA) ? kB : kA;
        vout = (vout == vA) ? vB : vA;
    }
    *keys_out = const_cast<uint32_t *>(kin);
    *perm_out = const_cast<uint32_t *>(vin);
}

void build_static_list(const int64_t *ptr, int32_t V, int64_t E, int64_t *tmp_flags,
                       int64_t *tmp_pos, int64_t *blk_a, int64_t *blk_b, int64_t n_blk,
                       hipStream_t s, int32_t **nz_v, int64_t **nz_scan, int32_t *n_nz,
                       int32_t **tiles)
{
    k_nz_flags<<<grid_for(V), kBlock, 0, s>>>(ptr, V, tmp_flags);
    SSSP_LAUNCH_CHECK("k_nz_flags");
    device_exclusive_scan_i64(tmp_flags, tmp_pos, V, blk_a, blk_b, n_blk, s);
    int64_t last_pos = 0, last_flag = 0;
    if (V > 0) {
        SSSP_HIP(hipMemcpyAsync(&last_pos, tmp_pos + V - 1, 8, hipMemcpyDeviceToHost, s));
        SSSP_HIP(hipMemcpyAsync(&last_flag, tmp_flags + V - 1, 8, hipMemcpyDeviceToHost, s));
        SSSP_HIP(hipStreamSynchronize(s));
    }
    const int64_t n = last_pos + last_flag;
    *n_nz = (int32_t)n;
    *nz_v = dmalloc<int32_t>((size_t)n + 1);
    *nz_scan = dmalloc<int64_t>((size_t)n + 1);
    k_nz_compact<<<grid_for(V), kBlock, 0, s>>>(ptr, V, tmp_pos, *nz_v, *nz_scan, n, E);
    SSSP_LAUNCH_CHECK("k_nz_compact");
    for (int c = 0; c < kChunkCount; c++) {
        const int64_t TE = (int64_t)kBlock * kChunkK[c];
        const int64_t ntiles = (E + TE - 1) / TE;
        tiles[c] = dmalloc<int32_t>((size_t)ntiles + 1);
        build_tile_table(*nz_scan, (int32_t)n, E, (int32_t)TE, tiles[c], s);
    }
}

}  // namespace

void scan_blocks_pair_inplace(int64_t *a, int64_t *b, int64_t nb, hipStream_t s)
{
    k_scan_blocks_pair<<<2, kBlock, 0, s>>>(a, b, nb);
    SSSP_LAUNCH_CHECK("k_scan_blocks_pair");
}

int64_t scan_blocks_needed(int64_t n) { return (n + kScanTile - 1) / kScanTile + 2; }

void device_exclusive_scan_i64(const int64_t *in, int64_t *out, int64_t n, int64_t *blk_a,
                               int64_t *blk_b, int64_t n_blk_cap, hipStream_t s)
{
    (void)blk_b;
    if (n <= 0) return;
    const int64_t nb = (n + kScanTile - 1) / kScanTile;
    if (nb + 1 > n_blk_cap) throw Error{SSSP_ERR_INTERNAL, "scan scratch too small"};
    k_scan_reduce<<<(unsigned)nb, kBlock, 0, s>>>(in, n, blk_a);
    SSSP_LAUNCH_CHECK("k_scan_reduce");
    k_scan_blocks<<<1, kBlock, 0, s>>>(blk_a, nb);
    SSSP_LAUNCH_CHECK("k_scan_blocks");
    k_scan_down<<<(unsigned)nb, kBlock, 0, s>>>(in, out, n, blk_a);
    SSSP_LAUNCH_CHECK("k_scan_down");
}

void build_tile_table(const int64_t *scan, int32_t n, int64_t total, int32_t tile_edges,
                      int32_t *tile_out, hipStream_t s)
{
    const int64_t ntiles = (total + tile_edges - 1) / tile_edges;
    k_tile_table<<<grid_for(n > 0 ? n : 1), kBlock, 0, s>>>(scan, n, tile_edges, ntiles, tile_out);
    SSSP_LAUNCH_CHECK("k_tile_table");
}

// Packed 12-byte edge records (src, dst, 4-byte weight) of the binary
// edge-list file -> the three edge arrays.
__global__ __launch_bounds__(kBlock) void k_split_records(const uint32_t *__restrict__ rec, int64_t n,
                                                         int32_t *src, int32_t *dst, uint32_t *w)
{
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * kBlock) {
        src[i] = (int32_t)rec[3 * i];
        dst[i] = (int32_t)rec[3 * i + 1];
        w[i] = rec[3 * i + 2];
    }
}

void split_records(const uint32_t *rec, int64_t n, int32_t *src, int32_t *dst, uint32_t *w,
                   hipStream_t s)
{
    if (n <= 0) return;
    k_split_records<<<grid_for(n), kBlock, 0, s>>>(rec, n, src, dst, w);
    SSSP_LAUNCH_CHECK("k_split_records");
}

void build_multi_layout(sssp_graph *g)
{
    if (g->m_rec || g->E == 0) return;
    hipStream_t s = g->stream;
    const int64_t E = g->E;
    const int32_t nblk = (int32_t)(((int64_t)g->V + (1 << kMBits) - 1) >> kMBits);
    const int64_t rs_tiles = (E + kRsTile - 1) / kRsTile;
    uint32_t *keys = dmalloc<uint32_t>((size_t)E);
    uint32_t *kA = dmalloc<uint32_t>((size_t)E), *kB = dmalloc<uint32_t>((size_t)E);
    uint32_t *vA = dmalloc<uint32_t>((size_t)E), *vB = dmalloc<uint32_t>((size_t)E);
    int64_t *counts = dmalloc<int64_t>((size_t)std::max<int64_t>(256 * rs_tiles, (int64_t)nblk + 1));
    int64_t *bptr = dmalloc<int64_t>((size_t)nblk + 1);
    struct Scratch {
        void *p[7];
        ~Scratch()
        {
            for (void *x : p) dfree(x);
        }
    } scratch{{keys, kA, kB, vA, vB, counts, bptr}};
    k_m_keys<<<grid_for(E), kBlock, 0, s>>>(g->el_dst, E, keys);
    SSSP_LAUNCH_CHECK("k_m_keys");
    uint32_t *skeys = nullptr, *perm = nullptr;
    radix_sort_pairs(reinterpret_cast<const int32_t *>(keys), E, nblk - 1, kA, kB, vA, vB, counts,
                     g->blk_a, g->blk_b, g->n_blk, s, &skeys, &perm);
    MRec *rec = dmalloc<MRec>((size_t)E);
    k_m_gather<<<grid_for(E), kBlock, 0, s>>>(skeys, perm, E, g->el_src, g->el_dst, g->el_w, rec);
    SSSP_LAUNCH_CHECK("k_m_gather");
    rowptr_from_sorted(skeys, E, nblk, bptr, s);
    std::vector<int64_t> bp((size_t)nblk + 1);
    SSSP_HIP(hipMemcpyAsync(bp.data(), bptr, 8 * bp.size(), hipMemcpyDeviceToHost, s));
    SSSP_HIP(hipStreamSynchronize(s));
    std::vector<BlkTile> tiles;
    for (int32_t b = 0; b < nblk; b++) {
        const int64_t c = bp[(size_t)b + 1] - bp[(size_t)b];
        const int64_t nt = (c + kMTileEdges - 1) / kMTileEdges;
        for (int64_t k = 0; k < nt; k++) {
            const int64_t e0 = bp[(size_t)b] + k * kMTileEdges;
            tiles.push_back(BlkTile{b, nt > 1 ? 1 : 0, e0, std::min(e0 + kMTileEdges, bp[(size_t)b + 1])});
        }
    }
    BlkTile *td = dmalloc<BlkTile>(std::max<size_t>(tiles.size(), 1));
    SSSP_HIP(hipMemcpyAsync(td, tiles.data(), sizeof(BlkTile) * tiles.size(), hipMemcpyHostToDevice, s));
    SSSP_HIP(hipStreamSynchronize(s));
    g->m_rec = rec;
    g->m_tile = td;
    g->n_mtile = (int32_t)tiles.size();
}

// Edge-list check on the device: every id in [0, V) and (wtype >= 0) every
// weight >= 0 and not NaN (unsigned weights always are); throws SSSP_ERR_ARG.
// The relaxations rely on it: distances are compared as unsigned bit patterns,
// which orders non-negative IEEE values only. The two counters live in
// `bad_buf` when the caller brings one (device memory, 16 bytes), else in a
// buffer allocated for the call; a kernel clears them either way.
void validate_edges(const int32_t *src, const int32_t *dst, const void *w, int wtype, int64_t E,
                    int32_t V, hipStream_t s, int64_t *bad_buf)
{
    if (E <= 0) return;
    int64_t *bad = bad_buf ? bad_buf : dmalloc<int64_t>(2);
    struct Free {
        void *p;
        ~Free() { dfree(p); }
    } fr{bad_buf ? nullptr : bad};
    k_zero_i64<<<1, 64, 0, s>>>(bad, 2);
    SSSP_LAUNCH_CHECK("k_zero_i64");
    if (wtype == SSSP_W_F32)
        k_validate<float><<<grid_for(E), kBlock, 0, s>>>(src, dst, static_cast<const float *>(w), E,
                                                         V, bad);
    else if (wtype == SSSP_W_F64)
        k_validate<double><<<grid_for(E), kBlock, 0, s>>>(src, dst, static_cast<const double *>(w),
                                                          E, V, bad);
    else if (wtype == SSSP_W_U32)
        k_validate<uint32_t><<<grid_for(E), kBlock, 0, s>>>(src, dst,
                                                            static_cast<const uint32_t *>(w), E, V,
                                                            bad);
    else
        k_validate<uint32_t><<<grid_for(E), kBlock, 0, s>>>(src, dst, nullptr, E, V, bad);
    SSSP_LAUNCH_CHECK("k_validate");
    int64_t hb[2] = {0, 0};
    SSSP_HIP(hipMemcpyAsync(hb, bad, 16, hipMemcpyDeviceToHost, s));
    SSSP_HIP(hipStreamSynchronize(s));
    if (hb[0]) throw Error{SSSP_ERR_ARG, std::to_string(hb[0]) + " edges have a vertex id outside [0, V)"};
    if (hb[1]) throw Error{SSSP_ERR_ARG, std::to_string(hb[1]) + " edges have a negative or NaN weight"};
}

// The tiles of the destination-blocked list: runs of at most `cap` edge slots
// (a multiple of kBlkChunk) of one block, `split` when a block has several.
// Without source ranges (R.stride == 0) a tile is a run of its block's slots;
// order 0 = smallest tiles first (stable by block), 1 = by block, 2 = largest
// first, 3 = by descending block. With source ranges each block's run is cut
// at the (chunk-aligned) slots where its sources reach multiples of the range
// size, and the tiles are dispatched range by range (by block within one):
// the workgroups running at one time then gather d[src] from one range of
// the distance array, which an XCD's L2 holds. bo: padded block offsets,
// nblk + 1 values.
std::vector<BlkTile> make_blk_tiles(const std::vector<int64_t> &bo, int64_t cap, int order,
                                    const BlkRanges &R)
{
    struct T {
        BlkTile t;
        int32_t range;
    };
    std::vector<T> tiles;
    const int32_t nblk = (int32_t)bo.size() - 1;
    const bool ranged = R.stride > 0 && !R.off.empty();
    const int32_t nr = ranged ? (R.n_fine + R.stride - 1) / R.stride : 1;
    for (int32_t b = 0; b < nblk; b++) {
        const size_t first = tiles.size();
        for (int32_t r = 0; r < nr; r++) {
            int64_t lo = bo[(size_t)b], hi = bo[(size_t)b + 1];
            if (ranged) {
                const size_t row = (size_t)b * (size_t)(R.n_fine + 1);
                lo = R.off[row + (size_t)std::min(r * R.stride, R.n_fine)];
                hi = R.off[row + (size_t)std::min((r + 1) * R.stride, R.n_fine)];
            }
            for (int64_t e0 = lo; e0 < hi; e0 += cap)
                tiles.push_back(T{BlkTile{b, 0, e0, std::min(e0 + cap, hi)}, r});
        }
        if (tiles.size() - first > 1)
            for (size_t i = first; i < tiles.size(); i++) tiles[i].t.split = 1;
    }
    if (ranged)
        std::stable_sort(tiles.begin(), tiles.end(),
                         [](const T &a, const T &b) { return a.range < b.range; });
    else if (order == 0 || order == 2)
        std::stable_sort(tiles.begin(), tiles.end(), [order](const T &a, const T &b) {
            const int64_t la = a.t.e1 - a.t.e0, lb = b.t.e1 - b.t.e0;
            return order == 0 ? la < lb : la > lb;
        });
    else if (order == 3)
        std::stable_sort(tiles.begin(), tiles.end(),
                         [](const T &a, const T &b) { return a.t.blk > b.t.blk; });
    std::vector<BlkTile> out;
    out.reserve(tiles.size());
    for (const T &x : tiles) out.push_back(x.t);
    return out;
}

void retile_blocked(sssp_graph *g, int64_t cap, int order, int range_bits)
{
    if (g->b_off_host.empty()) return;
    BlkRanges R = g->b_ranges;
    R.stride = 0;
    if (range_bits > 0) {
        if (R.off.empty() || range_bits < R.fine_bits || range_bits > 31)
            throw Error{SSSP_ERR_ARG, "blk_range_bits: source ranges unavailable for this size"};
        R.stride = 1 << (range_bits - R.fine_bits);
    }
    std::vector<BlkTile> tiles = make_blk_tiles(g->b_off_host, cap, order, R);
    BlkTile *td = dmalloc<BlkTile>(std::max<size_t>(tiles.size(), 1));
    SSSP_HIP(hipMemcpyAsync(td, tiles.data(), sizeof(BlkTile) * tiles.size(), hipMemcpyHostToDevice,
                            g->stream));
    SSSP_HIP(hipStreamSynchronize(g->stream));
    uint16_t *ad = dmalloc<uint16_t>(std::max<size_t>(tiles.size(), 1) * kBlkThreads);
    dfree(g->b_tile);
    dfree(g->b_app);
    g->b_tile = td;
    g->b_app = ad;
    g->frontier_bits = false;
    g->n_btile = (int32_t)tiles.size();
    g->blk_tile_cap = cap;
    g->blk_tile_order = order;
    g->blk_range_bits = range_bits;
}

void build_layouts(sssp_graph *g, const int32_t *d_src, const int32_t *d_dst, const uint32_t *d_w,
                   bool build_csc)
{
    hipStream_t s = g->stream;
    const int64_t E = g->E;
    const int32_t V = g->V;

    validate_edges(d_src, d_dst, d_w, g->wtype, E, V, s);

    // scratch for the sorts (of E edges, and of V vertices for the order)
    const int64_t n_sort = std::max<int64_t>(E, (int64_t)V);
    const int64_t rs_tiles = (n_sort + kRsTile - 1) / kRsTile;
    const int64_t scan_n = std::max<int64_t>((int64_t)V + 1, 256 * rs_tiles);
    g->n_blk = scan_blocks_needed(scan_n);
    g->blk_a = dmalloc<int64_t>((size_t)g->n_blk);
    g->blk_b = dmalloc<int64_t>((size_t)g->n_blk);
    uint32_t *kA = dmalloc<uint32_t>((size_t)n_sort), *kB = dmalloc<uint32_t>((size_t)n_sort);
    uint32_t *vA = dmalloc<uint32_t>((size_t)n_sort), *vB = dmalloc<uint32_t>((size_t)n_sort);
    int64_t *counts = dmalloc<int64_t>((size_t)std::max<int64_t>(256 * rs_tiles, (int64_t)V + 1));
    int64_t *tmp_pos = dmalloc<int64_t>((size_t)V + 1);
    struct Scratch {
        void *p[8];
        ~Scratch()
        {
            for (void *x : p) dfree(x);
        }
    } scratch{{kA, kB, vA, vB, counts, tmp_pos, nullptr, nullptr}};

    // ---- vertex order: slot = rank by descending in- plus out-degree, ties
    // by descending out-degree, then by vertex id (two stable sorts: by
    // out-degree, then by total degree); the edges are renamed into slots
    // before any layout is built. On R-MAT graphs the hubs are vertex ids
    // with few one bits, scattered over the whole id range, so the distance
    // gathers of a dense round touch a cache line per hub; in slot order the
    // hubs' distances are one dense prefix, and among the many vertices of
    // equal small degree the more often gathered come first. Orders measured
    // at R-MAT scale 25 (dense round per launch, learned solve): ids 2.04 /
    // 21.3 ms; total degree, ties by id 1.70 / 20.0; out-degree 1.79 / 20.5;
    // in-degree 1.68 / 20.1; total degree, ties by out-degree 1.61 / 19.0.
    // It pays from 2^23 vertices on skewed graphs (learned R-MAT GTEPS with /
    // without: scale 23 21.4 / 19.0, 24 24.5 / 22.0, 25 +12 %), not below or
    // without hubs (scale 22 x 16 14.6 / 14.6, 22 x 32 20.3 / 22.7, uniform
    // 2^24 x 16 7.9 / 8.2): by default the layouts use it when V >= 2^23 and
    // the largest total degree is at least 16 times the mean.
    uint32_t *tot = nullptr, *outd = nullptr;
    if (g->relabel_mode > 0 || (g->relabel_mode == 0 && V >= (1 << 23))) {
        if (E > 0 && V > 1) {
            tot = dmalloc<uint32_t>((size_t)V);
            scratch.p[6] = tot;
            outd = dmalloc<uint32_t>((size_t)V);
            scratch.p[7] = outd;
            SSSP_HIP(hipMemsetAsync(tot, 0, 4 * (size_t)V, s));
            SSSP_HIP(hipMemsetAsync(outd, 0, 4 * (size_t)V, s));
            k_degrees<<<grid_for(E), kBlock, 0, s>>>(d_src, d_dst, E, tot, outd);
            SSSP_LAUNCH_CHECK("k_degrees");
        }
        g->relabel = tot != nullptr;
        if (g->relabel && g->relabel_mode == 0) {
            uint32_t *mx = dmalloc<uint32_t>(1);
            struct F1 {
                void *p;
                ~F1() { dfree(p); }
            } f1{mx};
            SSSP_HIP(hipMemsetAsync(mx, 0, 4, s));
            k_max_u32<<<grid_for(V), kBlock, 0, s>>>(tot, V, mx);
            SSSP_LAUNCH_CHECK("k_max_u32");
            uint32_t hmax = 0;
            SSSP_HIP(hipMemcpyAsync(&hmax, mx, 4, hipMemcpyDeviceToHost, s));
            SSSP_HIP(hipStreamSynchronize(s));
            // mean total degree = 2E / V
            g->relabel = (double)hmax * (double)V >= 16.0 * 2.0 * (double)E;
        }
    }
    if (g->relabel) {
        uint32_t *okeys = dmalloc<uint32_t>((size_t)V);
        uint32_t *by_out = dmalloc<uint32_t>((size_t)V);
        struct F0 {
            void *p[2];
            ~F0()
            {
                for (void *x : p) dfree(x);
            }
        } f0{{okeys, by_out}};
        uint32_t *sk = nullptr, *ord = nullptr;
        k_order_keys<<<grid_for(V), kBlock, 0, s>>>(outd, nullptr, V, okeys);
        SSSP_LAUNCH_CHECK("k_order_keys out");
        radix_sort_pairs(reinterpret_cast<const int32_t *>(okeys), V, (int32_t)kOrderCap, kA, kB,
                         vA, vB, counts, g->blk_a, g->blk_b, g->n_blk, s, &sk, &ord);
        SSSP_HIP(hipMemcpyAsync(by_out, ord, 4 * (size_t)V, hipMemcpyDeviceToDevice, s));
        k_order_keys<<<grid_for(V), kBlock, 0, s>>>(tot, by_out, V, okeys);
        SSSP_LAUNCH_CHECK("k_order_keys total");
        radix_sort_pairs(reinterpret_cast<const int32_t *>(okeys), V, (int32_t)kOrderCap, kA, kB,
                         vA, vB, counts, g->blk_a, g->blk_b, g->n_blk, s, &sk, &ord);
        g->vperm = dmalloc<int32_t>((size_t)V);
        g->vinv = dmalloc<int32_t>((size_t)V);
        k_invert_order<<<grid_for(V), kBlock, 0, s>>>(by_out, ord, V, g->vperm, g->vinv);
        SSSP_LAUNCH_CHECK("k_invert_order");
        // the renamed edges take the degree arrays' places in the scratch
        // list: both are dead once the order exists
        SSSP_HIP(hipStreamSynchronize(s));
        dfree(tot);
        dfree(outd);
        scratch.p[6] = scratch.p[7] = nullptr;
        int32_t *rs = dmalloc<int32_t>((size_t)E);
        scratch.p[6] = rs;
        int32_t *rd = dmalloc<int32_t>((size_t)E);
        scratch.p[7] = rd;
        k_relabel_edges<<<grid_for(E), kBlock, 0, s>>>(d_src, d_dst, E, g->vperm, rs, rd);
        SSSP_LAUNCH_CHECK("k_relabel_edges");
        g->vperm_host.resize((size_t)V);
        SSSP_HIP(hipMemcpyAsync(g->vperm_host.data(), g->vperm, 4 * (size_t)V,
                                hipMemcpyDeviceToHost, s));
        SSSP_HIP(hipStreamSynchronize(s));
        d_src = rs;
        d_dst = rd;
    } else {
        dfree(tot);
        dfree(outd);
        scratch.p[6] = scratch.p[7] = nullptr;
    }

    // ---- CSR / src-sorted edge list ----
    uint32_t *skeys = nullptr, *perm = nullptr;
    radix_sort_pairs(d_src, E, V > 0 ? V - 1 : 0, kA, kB, vA, vB, counts, g->blk_a, g->blk_b,
                     g->n_blk, s, &skeys, &perm);
    g->el_src = dmalloc<int32_t>((size_t)E);
    g->el_dst = dmalloc<int32_t>((size_t)E);
    g->el_w = dmalloc<uint32_t>((size_t)E);
    g->row_ptr = dmalloc<int64_t>((size_t)V + 1);
    if (E > 0) {
        k_copy_i32<<<grid_for(E), kBlock, 0, s>>>(skeys, g->el_src, E);
        k_gather_pair<uint32_t><<<grid_for(E), kBlock, 0, s>>>(perm, E, d_dst, d_w, g->el_dst, g->el_w);
        SSSP_LAUNCH_CHECK("csr gather");
    }
    rowptr_from_sorted(skeys, E, V, g->row_ptr, s);
    build_static_list(g->row_ptr, V, E, counts, tmp_pos, g->blk_a, g->blk_b, g->n_blk, s,
                      &g->nzo_v, &g->nzo_scan, &g->n_nzo, g->nzo_tile);

    // ---- XCD-partitioned edge list: stable split of the src-sorted list into
    // 8 destination groups (src order kept inside each group) ----
    {
        int64_t *gcnt = dmalloc<int64_t>(8);
        struct F {
            void *p;
            ~F() { dfree(p); }
        } fg{gcnt};
        SSSP_HIP(hipMemsetAsync(gcnt, 0, 64, s));
        if (E > 0) {
            k_group_keys<<<grid_for(E), kBlock, 0, s>>>(g->el_dst, E, kA, gcnt);
            SSSP_LAUNCH_CHECK("k_group_keys");
        }
        int64_t hc[8];
        SSSP_HIP(hipMemcpyAsync(hc, gcnt, 64, hipMemcpyDeviceToHost, s));
        SSSP_HIP(hipStreamSynchronize(s));
        int64_t hoff[9], hbase[9];
        hoff[0] = hbase[0] = 0;
        for (int x = 0; x < 8; x++) {
            hbase[x + 1] = hbase[x] + hc[x];
            hoff[x + 1] = hoff[x] + (hc[x] + 3) / 4 * 4;
        }
        g->xgoff_host.assign(hoff, hoff + 9);
        uint32_t *gperm = nullptr, *gkeys = nullptr;
        // one radix pass on the 3-bit group key; values = positions in the
        // src-sorted list (stable)
        radix_sort_pairs(reinterpret_cast<const int32_t *>(kA), E, 7, kB, kB, vA, vB, counts,
                         g->blk_a, g->blk_b, g->n_blk, s, &gkeys, &gperm);
        int64_t *dgo = dmalloc<int64_t>(18);
        struct F2 {
            void *p;
            ~F2() { dfree(p); }
        } fg2{dgo};
        SSSP_HIP(hipMemcpyAsync(dgo, hoff, 72, hipMemcpyHostToDevice, s));
        SSSP_HIP(hipMemcpyAsync(dgo + 9, hbase, 72, hipMemcpyHostToDevice, s));
        g->x_src = dmalloc<int32_t>((size_t)hoff[8] + 4);
        g->x_dst = dmalloc<int32_t>((size_t)hoff[8] + 4);
        g->x_w = dmalloc<uint32_t>((size_t)hoff[8] + 4);
        g->xgoff = dmalloc<int64_t>(9);
        SSSP_HIP(hipMemcpyAsync(g->xgoff, hoff, 72, hipMemcpyHostToDevice, s));
        if (hoff[8] > 0) {
            k_group_gather<<<grid_for(hoff[8]), kBlock, 0, s>>>(gperm, dgo, dgo + 9, g->el_src,
                                                                g->el_dst, g->el_w, g->x_src,
                                                                g->x_dst, g->x_w);
            SSSP_LAUNCH_CHECK("k_group_gather");
        }
        SSSP_HIP(hipStreamSynchronize(s));
    }

    // ---- destination-blocked edge list ----
    if (E > 0) {
        const int32_t nblk = (int32_t)(((int64_t)V + (1 << kBlkBits) - 1) >> kBlkBits);
        uint32_t *keys = dmalloc<uint32_t>((size_t)E);
        int64_t *bptr = dmalloc<int64_t>((size_t)nblk + 1);
        int64_t *boff = dmalloc<int64_t>((size_t)nblk + 1);
        struct F3 {
            void *p[3];
            ~F3()
            {
                for (void *x : p) dfree(x);
            }
        } f3{{keys, bptr, boff}};
        k_blk_keys<<<grid_for(E), kBlock, 0, s>>>(g->el_dst, E, keys);
        SSSP_LAUNCH_CHECK("k_blk_keys");
        uint32_t *bkeys = nullptr, *bperm = nullptr;
        radix_sort_pairs(reinterpret_cast<const int32_t *>(keys), E, nblk - 1, kA, kB, vA, vB,
                         counts, g->blk_a, g->blk_b, g->n_blk, s, &bkeys, &bperm);
        rowptr_from_sorted(bkeys, E, nblk, bptr, s);
        std::vector<int64_t> bp((size_t)nblk + 1), bo((size_t)nblk + 1);
        SSSP_HIP(hipMemcpyAsync(bp.data(), bptr, 8 * bp.size(), hipMemcpyDeviceToHost, s));
        SSSP_HIP(hipStreamSynchronize(s));
        bo[0] = 0;
        // Without hubs (the largest block's in-edges under twice the mean:
        // uniform random graphs) appending all-edge rounds lose to dense ones
        // from a smaller share of the graph on; skewed graphs (R-MAT) keep
        // the appends longer. Measured: ER 2^24 x 16 34.8 ms at 0.4 against
        // 37.5 at 0.15, R-MAT 22 x 32 6.1 ms at 0.15 against 6.5 at 0.4.
        int64_t max_blk = 0;
        for (int32_t b = 0; b < nblk; b++)
            max_blk = std::max(max_blk, bp[(size_t)b + 1] - bp[(size_t)b]);
        const bool no_hubs = (double)max_blk * nblk < 2.0 * (double)E;
        if (no_hubs) g->dense_all_frac = 0.4;
        for (int32_t b = 0; b < nblk; b++)
            bo[(size_t)b + 1] = bo[(size_t)b] +
                                (bp[(size_t)b + 1] - bp[(size_t)b] + kBlkChunk - 1) / kBlkChunk * kBlkChunk;
        // Tile cap: 2^17 edges up to 2^29 edges, doubled while more than
        // kBlkTileCount tiles of the cap fit in the graph -- the 2.08 G-edge
        // R-MAT graph ran 50.9 GTEPS with 2^19-edge tiles, 46.7 with 2^17;
        // R-MAT 25 x 16 (2^29 edges), learned / plain ms, with the dense
        // round's stream loaded a step ahead: 14.9 / 22.3 at 2^17, 15.6 /
        // 23.3 at 2^18, 17.0 / 26.3 at 2^19 (before that change 2^18 was the
        // faster: 19.2 against 19.6 ms). A power of two, so every tile holds
        // whole chunks.
        int64_t tcap = kBlkTileEdges;
        while (tcap * kBlkTileCount < E) tcap <<= 1;
        g->b_off_host = bo;
        std::vector<BlkTile> tiles = make_blk_tiles(bo, tcap, 0, BlkRanges{});
        g->blk_tile_cap = tcap;
        g->blk_tile_order = 0;
        const int64_t Ep = bo[(size_t)nblk];
        const int64_t nch = Ep / kBlkChunk;
        SSSP_HIP(hipMemcpyAsync(boff, bo.data(), 8 * bo.size(), hipMemcpyHostToDevice, s));
        // whole sources and destination offsets first, packed below
        int32_t *bs = dmalloc<int32_t>((size_t)Ep);
        uint16_t *bd = dmalloc<uint16_t>((size_t)Ep);
        uint8_t *wflag = dmalloc<uint8_t>((size_t)nch);
        int32_t *widx = dmalloc<int32_t>((size_t)nch);
        struct F4 {
            void *p[4];
            ~F4()
            {
                for (void *x : p) dfree(x);
            }
        } f4{{bs, bd, wflag, widx}};
        g->b_w = dmalloc<uint32_t>((size_t)Ep);
        g->n_btile = (int32_t)tiles.size();
        g->b_edges = Ep;
        g->b_tile = dmalloc<BlkTile>(tiles.size());
        g->b_app = dmalloc<uint16_t>(tiles.size() * kBlkThreads);
        SSSP_HIP(hipMemcpyAsync(g->b_tile, tiles.data(), sizeof(BlkTile) * tiles.size(),
                                hipMemcpyHostToDevice, s));
        k_blk_gather<<<grid_for(E), kBlock, 0, s>>>(bkeys, bperm, E, bptr, boff, g->el_src,
                                                    g->el_dst, g->el_w, bs, bd, g->b_w);
        SSSP_LAUNCH_CHECK("k_blk_gather");
        k_blk_pad<<<grid_for((int64_t)nblk * kBlkChunk), kBlock, 0, s>>>(bptr, boff, nblk, bs, bd,
                                                                         g->b_w);
        SSSP_LAUNCH_CHECK("k_blk_pad");
        const unsigned chunk_grid = (unsigned)((nch * 64 + kBlock - 1) / kBlock);
        k_blk_chunk_wide<<<chunk_grid, kBlock, 0, s>>>(bs, nch, wflag);
        SSSP_LAUNCH_CHECK("k_blk_chunk_wide");
        std::vector<uint8_t> wf((size_t)nch);
        SSSP_HIP(hipMemcpyAsync(wf.data(), wflag, (size_t)nch, hipMemcpyDeviceToHost, s));
        SSSP_HIP(hipStreamSynchronize(s));
        std::vector<int32_t> wi((size_t)nch);
        int64_t nw = 0;
        for (int64_t c = 0; c < nch; c++) wi[(size_t)c] = wf[(size_t)c] ? (int32_t)nw++ : -1;
        SSSP_HIP(hipMemcpyAsync(widx, wi.data(), 4 * (size_t)nch, hipMemcpyHostToDevice, s));
        g->b_pk = dmalloc<uint32_t>((size_t)Ep);
        g->b_cbase = dmalloc<int32_t>((size_t)nch);
        g->b_wide = nw ? dmalloc<int32_t>((size_t)(nw * kBlkChunk)) : nullptr;
        g->b_nwide = nw;
        k_blk_pack<<<chunk_grid, kBlock, 0, s>>>(bs, bd, widx, nch, g->b_pk, g->b_cbase, g->b_wide);
        SSSP_LAUNCH_CHECK("k_blk_pack");
        // source-range cuts of every block at 2^kBlkRangeFineBits granularity
        // (for tiles dispatched range by range), when the table is small
        const int64_t nfine = ((int64_t)V + (1 << kBlkRangeFineBits) - 1) >> kBlkRangeFineBits;
        if ((int64_t)nblk * (nfine + 1) <= kBlkRangeMaxEntries) {
            const int64_t n_off = (int64_t)nblk * (nfine + 1);
            int64_t *roff = dmalloc<int64_t>((size_t)n_off);
            struct F1 {
                void *p;
                ~F1() { dfree(p); }
            } f1{roff};
            k_blk_range_offsets<<<grid_for(n_off), kBlock, 0, s>>>(bs, boff, nblk, kBlkRangeFineBits,
                                                                  (int32_t)nfine, roff);
            SSSP_LAUNCH_CHECK("k_blk_range_offsets");
            g->b_ranges.off.resize((size_t)n_off);
            SSSP_HIP(hipMemcpyAsync(g->b_ranges.off.data(), roff, 8 * (size_t)n_off,
                                    hipMemcpyDeviceToHost, s));
            g->b_ranges.fine_bits = kBlkRangeFineBits;
            g->b_ranges.n_fine = (int32_t)nfine;
        }
        SSSP_HIP(hipStreamSynchronize(s));
        // Without hubs the d[src] gathers of a tile spread over the whole
        // distance array (degree order cannot gather them), so a graph whose
        // distances outgrow the L2 several times over runs its tiles by source
        // range, 8 ranges (uniform random graphs, ms learned / plain / dense
        // round: 2^24 x 16 31.7 / 71.7 / 2.96 by block, 23.6 / 36.5 / 1.57 in
        // 8 ranges, 26.4 / 37.1 / 1.83 in 16, 23.9 / 41.6 / 1.75 in 4; 2^22 x 16
        // 7.75 / 11.1 / 0.38 by block, 7.09 / 7.93 / 0.38 in 8 ranges). Graphs
        // with hubs lose by it (R-MAT 25 x 16: 17.2 / 24.7 ms by block, 18.5 /
        // 26.4 in 8 ranges).
        if (no_hubs && V >= (1 << kBlkRangeMinVertexBits) && !g->b_ranges.off.empty()) {
            int lg = 0;
            while (((int64_t)1 << lg) < (int64_t)V) lg++;
            retile_blocked(g, tcap, 0, std::max(kBlkRangeFineBits, lg - 3));
        }
    }

    // ---- CSC ----
    if (build_csc) {
        radix_sort_pairs(d_dst, E, V > 0 ? V - 1 : 0, kA, kB, vA, vB, counts, g->blk_a, g->blk_b,
                         g->n_blk, s, &skeys, &perm);
        g->csc_src = dmalloc<int32_t>((size_t)E);
        g->csc_w = dmalloc<uint32_t>((size_t)E);
        g->col_ptr = dmalloc<int64_t>((size_t)V + 1);
        if (E > 0) {
            k_gather_pair<uint32_t><<<grid_for(E), kBlock, 0, s>>>(perm, E, d_src, d_w, g->csc_src, g->csc_w);
            SSSP_LAUNCH_CHECK("csc gather");
        }
        rowptr_from_sorted(skeys, E, V, g->col_ptr, s);
        build_static_list(g->col_ptr, V, E, counts, tmp_pos, g->blk_a, g->blk_b, g->n_blk, s,
                          &g->nzi_v, &g->nzi_scan, &g->n_nzi, g->nzi_tile);
        g->has_csc = true;
    }
    SSSP_HIP(hipStreamSynchronize(s));
}

}  // namespace sssp
