// capi.cpp -- the extern "C" boundary of libsssp_mi355x.so: handle lifetime,
// uploads, status codes, and the reference-shaped wrappers.
#include <cmath>
#include <cstring>
#include <string>

#include "mlp_dev.hpp"
#include "sssp_graph.hpp"

namespace sssp {

namespace {
thread_local std::string g_last_error;
}

int g_fail_next_alloc = 0;

void set_error(const std::string &msg) { g_last_error = msg; }

int fail(int code, const std::string &msg)
{
    g_last_error = msg;
    return code;
}

namespace {

void set_device(int dev)
{
    int n = 0;
    SSSP_HIP(hipGetDeviceCount(&n));
    SSSP_REQUIRE(dev >= 0 && dev < n, "device index out of range (" + std::to_string(dev) +
                                          " of " + std::to_string(n) + ")");
    SSSP_HIP(hipSetDevice(dev));
}

void graph_free(sssp_graph *g)
{
    if (!g) return;
    (void)hipSetDevice(g->device);
    if (g->stream) (void)hipStreamSynchronize(g->stream);
    void *ptrs[] = {g->row_ptr, g->el_src, g->el_dst, g->el_w, g->col_ptr, g->csc_src, g->csc_w,
                    g->nzo_v, g->nzo_scan, g->nzi_v, g->nzi_scan, g->sample_v, g->sample_in, g->dist, g->stamp,
                    g->queue[0], g->queue[1], g->lb_v, g->lb_start, g->lb_scan, g->lb_tile,
                    g->blk_a, g->blk_b, g->parent, g->ctrl, g->bdist, g->bscratch,
                    g->src_list, g->btasks, g->bcounter, g->bround, g->x_src, g->x_dst,
                    g->x_w, g->xgoff, g->pile[0], g->pile[1], g->farst, g->snap,
                    g->b_pk, g->b_cbase, g->b_wide, g->b_w, g->b_tile, g->b_app, g->fbits, g->m_rec, g->m_tile, g->mdist,
                    g->vperm, g->vinv, g->dout != g->dist ? g->dout : nullptr, g->msrc, g->mflags};
    if (g->batch_exec) (void)hipGraphExecDestroy((hipGraphExec_t)g->batch_exec);
    for (auto &pair : g->mexec)
        for (void *e : pair)
            if (e) (void)hipGraphExecDestroy((hipGraphExec_t)e);
    for (void *p : ptrs) dfree(p);
    for (int c = 0; c < kChunkCount; c++) {
        dfree(g->nzo_tile[c]);
        dfree(g->nzi_tile[c]);
    }
    if (g->ctrl_host) (void)hipHostFree(g->ctrl_host);
    hipEvent_t evs[] = {g->ev_a, g->ev_b, g->ev_r0, g->ev_r1};
    for (hipEvent_t e : evs)
        if (e) (void)hipEventDestroy(e);
    if (g->stream) (void)hipStreamDestroy(g->stream);
    delete g;
}

size_t wbytes(int wtype) { return wtype == SSSP_W_F64 ? 8 : 4; }

}  // namespace

}  // namespace sssp

using namespace sssp;

extern "C" {

const char *sssp_last_error(void) { return g_last_error.c_str(); }

int sssp_version(void) { return 10000; }

int sssp_debug_fail_next_alloc(void)
{
    g_fail_next_alloc = 1;
    return SSSP_OK;
}

int sssp_device_count(int32_t *count)
{
    return guard([&] {
        SSSP_REQUIRE(count, "null count");
        int n = 0;
        SSSP_HIP(hipGetDeviceCount(&n));
        *count = n;
    });
}

int sssp_device_info(int32_t device, char *name, int32_t name_len, int32_t *cu_count,
                     int64_t *hbm_bytes, int32_t *l2_bytes)
{
    return guard([&] {
        set_device(device);
        hipDeviceProp_t p;
        SSSP_HIP(hipGetDeviceProperties(&p, device));
        if (name && name_len > 0) {
            std::strncpy(name, p.gcnArchName, (size_t)name_len - 1);
            name[name_len - 1] = 0;
        }
        if (cu_count) *cu_count = p.multiProcessorCount;
        if (hbm_bytes) *hbm_bytes = (int64_t)p.totalGlobalMem;
        if (l2_bytes) *l2_bytes = p.l2CacheSize;
    });
}

int sssp_device_props(int32_t device, char *name, int32_t name_len, int64_t *free_bytes,
                      int64_t *total_bytes, int32_t *max_threads_per_block, int32_t *max_grid)
{
    return guard([&] {
        int n = 0, prev = 0;
        SSSP_HIP(hipGetDeviceCount(&n));
        SSSP_REQUIRE(device >= 0 && device < n, "device index out of range");
        SSSP_HIP(hipGetDevice(&prev));
        struct Restore {
            int d;
            ~Restore() { (void)hipSetDevice(d); }
        } restore{prev};
        SSSP_HIP(hipSetDevice(device));
        hipDeviceProp_t p;
        SSSP_HIP(hipGetDeviceProperties(&p, device));
        if (name && name_len > 0) {
            std::strncpy(name, p.name, (size_t)name_len - 1);
            name[name_len - 1] = 0;
        }
        size_t fr = 0, tot = 0;
        if (free_bytes || total_bytes) SSSP_HIP(hipMemGetInfo(&fr, &tot));
        if (free_bytes) *free_bytes = (int64_t)fr;
        if (total_bytes) *total_bytes = (int64_t)tot;
        if (max_threads_per_block) *max_threads_per_block = p.maxThreadsPerBlock;
        if (max_grid)
            for (int i = 0; i < 3; i++) max_grid[i] = p.maxGridSize[i];
    });
}

}  // extern "C"

namespace sssp {
namespace {

sssp_graph *graph_begin(int32_t weight_type, int32_t vertices, int64_t edges, int32_t device)
{
    SSSP_REQUIRE(weight_type == SSSP_W_F32 || weight_type == SSSP_W_U32,
                 "graph weight type must be SSSP_W_F32 or SSSP_W_U32");
    SSSP_REQUIRE(vertices >= 1, "graph needs at least one vertex");
    SSSP_REQUIRE(edges >= 0 && edges < ((int64_t)1 << 31) - 1, "edge count must be in [0, 2^31 - 1)");
    set_device(device);
    sssp_graph *g = new sssp_graph();
    g->device = device;
    g->wtype = weight_type;
    g->V = vertices;
    g->E = edges;
    hipDeviceProp_t p;
    SSSP_HIP(hipGetDeviceProperties(&p, device));
    g->cu_count = p.multiProcessorCount > 0 ? p.multiProcessorCount : 256;
    SSSP_HIP(hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking));
    return g;
}

// Builds the layouts from edge arrays already in device memory and allocates
// the solve state.
void graph_finish(sssp_graph *g, const int32_t *d_src, const int32_t *d_dst, const uint32_t *d_w,
                  uint32_t flags)
{
    hipStream_t s = g->stream;
    SSSP_REQUIRE(!((flags & SSSP_NO_RELABEL) && (flags & SSSP_RELABEL)),
                 "SSSP_RELABEL and SSSP_NO_RELABEL exclude each other");
    g->relabel_mode = (flags & SSSP_RELABEL) ? 1 : (flags & SSSP_NO_RELABEL) ? -1 : 0;
    build_layouts(g, d_src, d_dst, d_w, !(flags & SSSP_NO_CSC));
    const size_t V = (size_t)g->V;
    g->dist = dmalloc<uint32_t>(V);
    g->dout = g->relabel ? dmalloc<uint32_t>(V) : g->dist;
    g->stamp = dmalloc<int32_t>(V);
    g->pile[0] = dmalloc<int32_t>(V);
    g->pile[1] = dmalloc<int32_t>(V);
    g->farst = dmalloc<int32_t>(V);
    g->snap = dmalloc<uint32_t>(V);
    reset_stamps(g);
    g->queue[0] = dmalloc<int32_t>(V + 1);
    g->queue[1] = dmalloc<int32_t>(V + 1);
    g->lb_v = dmalloc<int32_t>(V + 1);
    {
        const size_t nfw = 2 * (((size_t)V + 63) / 64) + 2;
        g->fbits = dmalloc<uint32_t>(nfw);
        SSSP_HIP(hipMemsetAsync(g->fbits, 0, 4 * nfw, s));
    }
    g->lb_start = dmalloc<int64_t>(V + 1);
    g->lb_scan = dmalloc<int64_t>(V + 2);
    g->lb_tile = dmalloc<int32_t>((size_t)(g->E / kBlock) + 4);
    g->parent = dmalloc<int32_t>(V);
    g->ctrl = dmalloc<Ctrl>(1);
    SSSP_HIP(hipMemsetAsync(g->ctrl, 0, sizeof(Ctrl), s));
    // coherent and mapped: the decision kernel writes the round's counters
    // straight into it, so a round costs one stream synchronize, no copies
    SSSP_HIP(hipHostMalloc((void **)&g->ctrl_host, sizeof(Ctrl),
                           hipHostMallocMapped | hipHostMallocCoherent));
    std::memset(g->ctrl_host, 0, sizeof(Ctrl));
    SSSP_HIP(hipHostGetDevicePointer((void **)&g->ctrl_host_dev, g->ctrl_host, 0));
    // fixed-stride vertex sample for the unsettled-fraction estimate
    g->n_sample = g->V < 16384 ? g->V : 16384;
    std::vector<int32_t> sv((size_t)g->n_sample);
    for (int32_t i = 0; i < g->n_sample; i++)
        sv[(size_t)i] = (int32_t)(((int64_t)i * g->V) / g->n_sample);
    g->sample_v = dmalloc<int32_t>((size_t)g->n_sample);
    SSSP_HIP(hipMemcpyAsync(g->sample_v, sv.data(), 4 * sv.size(), hipMemcpyHostToDevice, s));
    sample_in_degrees(g);
    SSSP_HIP(hipEventCreate(&g->ev_a));
    SSSP_HIP(hipEventCreate(&g->ev_b));
    SSSP_HIP(hipEventCreate(&g->ev_r0));
    SSSP_HIP(hipEventCreate(&g->ev_r1));
    SSSP_HIP(hipStreamSynchronize(s));
    // resident arrays: the three edge orders and the CSC, row offsets, the
    // per-vertex arrays (dist, stamp, piles, pile state, snapshot, parent,
    // queues, load-balancing entries, static entry lists)
    const int64_t Ei = g->E, Vi = g->V;
    const int64_t xe = g->xgoff_host.empty() ? 0 : g->xgoff_host.back();
    g->device_bytes = Ei * 12 + xe * 12 + g->b_edges * 8 + g->b_edges / kBlkChunk * 4 +
                      g->b_nwide * kBlkChunk * 4 + (int64_t)g->n_btile * (24 + kBlkThreads * 2) +
                      (Vi + 1) * 8 + (g->has_csc ? Ei * 8 + (Vi + 1) * 8 : 0) + Vi * 4 * 7 +
                      (Vi + 1) * 4 * 3 + (Vi + 1) * 8 * 2 + Vi * 24 + (g->relabel ? Vi * 12 : 0);
}

struct DevBuf {
    void *p[3] = {nullptr, nullptr, nullptr};
    ~DevBuf()
    {
        for (void *x : p) dfree(x);
    }
};

}  // namespace
}  // namespace sssp

extern "C" {

int sssp_graph_create(int32_t weight_type, const int32_t *src, const int32_t *dst,
                      const void *weights, int32_t vertices, int64_t edges, int32_t device,
                      uint32_t flags, sssp_graph **out)
{
    sssp_graph *g = nullptr;
    int rc = guard([&] {
        SSSP_REQUIRE(out, "null output handle");
        *out = nullptr;
        SSSP_REQUIRE(edges == 0 || (src && dst && weights), "null edge array");
        g = graph_begin(weight_type, vertices, edges, device);
        hipStream_t s = g->stream;
        const int32_t *d_src = src, *d_dst = dst;
        const uint32_t *d_w = static_cast<const uint32_t *>(weights);
        DevBuf up;
        if (!(flags & SSSP_INPUT_DEVICE) && edges > 0) {
            int32_t *u_src = dmalloc<int32_t>((size_t)edges);
            up.p[0] = u_src;
            int32_t *u_dst = dmalloc<int32_t>((size_t)edges);
            up.p[1] = u_dst;
            uint32_t *u_w = dmalloc<uint32_t>((size_t)edges);
            up.p[2] = u_w;
            SSSP_HIP(hipMemcpyAsync(u_src, src, 4 * (size_t)edges, hipMemcpyHostToDevice, s));
            SSSP_HIP(hipMemcpyAsync(u_dst, dst, 4 * (size_t)edges, hipMemcpyHostToDevice, s));
            SSSP_HIP(hipMemcpyAsync(u_w, weights, 4 * (size_t)edges, hipMemcpyHostToDevice, s));
            d_src = u_src;
            d_dst = u_dst;
            d_w = u_w;
        }
        graph_finish(g, d_src, d_dst, d_w, flags);
        *out = g;
    });
    if (rc != SSSP_OK && g) graph_free(g);
    return rc;
}

int sssp_graph_create_from_bin(const char *path, int32_t directed, int32_t weight_type,
                               int32_t device, uint32_t flags, sssp_graph **out)
{
    sssp_graph *g = nullptr;
    FILE *f = nullptr;
    struct Host {
        void *p[2] = {nullptr, nullptr};
        hipEvent_t ev[2] = {nullptr, nullptr};
        ~Host()
        {
            for (hipEvent_t e : ev)
                if (e) (void)hipEventDestroy(e);
            for (void *x : p)
                if (x) (void)hipHostFree(x);
        }
    };
    int rc = guard([&] {
        SSSP_REQUIRE(out && path, "null argument");
        *out = nullptr;
        f = std::fopen(path, "rb");
        if (!f) throw Error{SSSP_ERR_IO, std::string("cannot open ") + path};
        uint8_t dir = 0;
        int32_t V = 0;
        int64_t E = 0;
        if (std::fread(&dir, 1, 1, f) != 1 || std::fread(&V, 4, 1, f) != 1 ||
            std::fread(&E, 8, 1, f) != 1)
            throw Error{SSSP_ERR_IO, std::string("truncated header in ") + path};
        if (V < 1 || E < 0) throw Error{SSSP_ERR_IO, std::string("corrupt header in ") + path};
        if ((dir != 0) != (directed != 0))
            throw Error{SSSP_ERR_ARG, "loading a directed graph as undirected (or undirected as directed)"};
        if (std::fseek(f, 0, SEEK_END) != 0)
            throw Error{SSSP_ERR_IO, std::string("cannot seek in ") + path};
        const long size = std::ftell(f);
        const int64_t body = (int64_t)size - 13;
        if (E > 0 && body == 16 * E)
            throw Error{SSSP_ERR_ARG, "file has 8-byte weights; device graphs hold 4-byte weights"};
        if (body != 12 * E)
            throw Error{SSSP_ERR_IO, std::string("edge records do not match the edge count in ") + path};
        if (std::fseek(f, 13, SEEK_SET) != 0)
            throw Error{SSSP_ERR_IO, std::string("cannot seek in ") + path};
        g = graph_begin(weight_type, V, E, device);
        hipStream_t s = g->stream;
        DevBuf arr, stage;
        int32_t *d_src = dmalloc<int32_t>((size_t)(E > 0 ? E : 1));
        arr.p[0] = d_src;
        int32_t *d_dst = dmalloc<int32_t>((size_t)(E > 0 ? E : 1));
        arr.p[1] = d_dst;
        uint32_t *d_w = dmalloc<uint32_t>((size_t)(E > 0 ? E : 1));
        arr.p[2] = d_w;
        const int64_t chunk = E < ((int64_t)1 << 22) ? (E > 0 ? E : 1) : ((int64_t)1 << 22);
        Host h;
        for (int b = 0; b < 2; b++) {
            stage.p[b] = dmalloc<uint32_t>((size_t)(3 * chunk));
            SSSP_HIP(hipHostMalloc(&h.p[b], (size_t)(12 * chunk), hipHostMallocDefault));
            SSSP_HIP(hipEventCreate(&h.ev[b]));
        }
        // host reads of chunk k+1 overlap the copy and split of chunk k
        try {
            for (int64_t e0 = 0, k = 0; e0 < E; e0 += chunk, k++) {
                const int b = (int)(k & 1);
                const int64_t n = E - e0 < chunk ? E - e0 : chunk;
                if (k >= 2) SSSP_HIP(hipEventSynchronize(h.ev[b]));  // buffer b free again
                if (std::fread(h.p[b], 12, (size_t)n, f) != (size_t)n)
                    throw Error{SSSP_ERR_IO, std::string("truncated edge records in ") + path};
                uint32_t *dst_stage = static_cast<uint32_t *>(stage.p[b]);
                SSSP_HIP(hipMemcpyAsync(dst_stage, h.p[b], (size_t)(12 * n), hipMemcpyHostToDevice,
                                        s));
                split_records(dst_stage, n, d_src + e0, d_dst + e0, d_w + e0, s);
                SSSP_HIP(hipEventRecord(h.ev[b], s));
            }
        } catch (...) {
            (void)hipStreamSynchronize(s);  // no copy may still read the pinned buffers
            throw;
        }
        std::fclose(f);
        f = nullptr;
        SSSP_HIP(hipStreamSynchronize(s));
        graph_finish(g, d_src, d_dst, d_w, flags & ~SSSP_INPUT_DEVICE);
        *out = g;
    });
    if (f) std::fclose(f);
    if (rc != SSSP_OK && g) graph_free(g);
    return rc;
}

int sssp_graph_create_rmat(int32_t scale, int32_t edge_factor, int32_t a, int32_t b, int32_t c,
                           int32_t d, uint32_t base_seed, int32_t virtual_threads, int32_t device,
                           uint32_t flags, int32_t *src_out, int32_t *dst_out, float *w_out,
                           sssp_graph **out)
{
    sssp_graph *g = nullptr;
    int rc = guard([&] {
        SSSP_REQUIRE(out, "null output handle");
        *out = nullptr;
        SSSP_REQUIRE(scale >= 1 && scale <= 30, "rmat: scale must be in [1, 30]");
        SSSP_REQUIRE(edge_factor >= 1, "rmat: edge_factor must be >= 1");
        SSSP_REQUIRE(a >= 0 && b >= 0 && c >= 0 && d >= 0 && a + b + c + d <= 100,
                     "rmat: probabilities are integer percents summing to <= 100");
        SSSP_REQUIRE(virtual_threads >= 1, "rmat: virtual_threads must be >= 1");
        const int64_t E = ((int64_t)1 << scale) * edge_factor;
        g = graph_begin(SSSP_W_F32, 1 << scale, E, device);
        hipStream_t s = g->stream;
        DevBuf gen;
        int32_t *d_src = dmalloc<int32_t>((size_t)E);
        gen.p[0] = d_src;
        int32_t *d_dst = dmalloc<int32_t>((size_t)E);
        gen.p[1] = d_dst;
        float *d_w = dmalloc<float>((size_t)E);
        gen.p[2] = d_w;
        rmat_device(scale, edge_factor, a, b, c, d, base_seed, virtual_threads, d_src, d_dst, d_w, s);
        if (src_out) SSSP_HIP(hipMemcpyAsync(src_out, d_src, 4 * (size_t)E, hipMemcpyDeviceToHost, s));
        if (dst_out) SSSP_HIP(hipMemcpyAsync(dst_out, d_dst, 4 * (size_t)E, hipMemcpyDeviceToHost, s));
        if (w_out) SSSP_HIP(hipMemcpyAsync(w_out, d_w, 4 * (size_t)E, hipMemcpyDeviceToHost, s));
        graph_finish(g, d_src, d_dst, reinterpret_cast<const uint32_t *>(d_w), flags);
        *out = g;
    });
    if (rc != SSSP_OK && g) graph_free(g);
    return rc;
}

int sssp_graph_destroy(sssp_graph *g)
{
    return guard([&] {
        SSSP_REQUIRE(g, "null graph");
        graph_free(g);
    });
}

int sssp_graph_info(const sssp_graph *g, int32_t *vertices, int64_t *edges, int32_t *weight_type,
                    int64_t *device_bytes)
{
    return guard([&] {
        SSSP_REQUIRE(g, "null graph");
        if (vertices) *vertices = g->V;
        if (edges) *edges = g->E;
        if (weight_type) *weight_type = g->wtype;
        if (device_bytes) *device_bytes = g->device_bytes;
    });
}

int sssp_graph_vertex_order(const sssp_graph *g, int32_t *slot_of_vertex)
{
    return guard([&] {
        SSSP_REQUIRE(g && slot_of_vertex, "null argument");
        for (int32_t v = 0; v < g->V; v++) slot_of_vertex[v] = slot_of(g, v);
    });
}

int sssp_graph_download_layout(const sssp_graph *g, int64_t *row_ptr, int32_t *el_src,
                               int32_t *el_dst, void *el_w, int64_t *col_ptr, int32_t *csc_src,
                               void *csc_w)
{
    return guard([&] {
        SSSP_REQUIRE(g, "null graph");
        set_device(g->device);
        const size_t V1 = (size_t)g->V + 1, E = (size_t)g->E;
        hipStream_t s = g->stream;
        if (row_ptr) SSSP_HIP(hipMemcpyAsync(row_ptr, g->row_ptr, 8 * V1, hipMemcpyDeviceToHost, s));
        if (el_src) SSSP_HIP(hipMemcpyAsync(el_src, g->el_src, 4 * E, hipMemcpyDeviceToHost, s));
        if (el_dst) SSSP_HIP(hipMemcpyAsync(el_dst, g->el_dst, 4 * E, hipMemcpyDeviceToHost, s));
        if (el_w) SSSP_HIP(hipMemcpyAsync(el_w, g->el_w, 4 * E, hipMemcpyDeviceToHost, s));
        if (col_ptr || csc_src || csc_w) SSSP_REQUIRE(g->has_csc, "graph was built without CSC");
        if (col_ptr) SSSP_HIP(hipMemcpyAsync(col_ptr, g->col_ptr, 8 * V1, hipMemcpyDeviceToHost, s));
        if (csc_src) SSSP_HIP(hipMemcpyAsync(csc_src, g->csc_src, 4 * E, hipMemcpyDeviceToHost, s));
        if (csc_w) SSSP_HIP(hipMemcpyAsync(csc_w, g->csc_w, 4 * E, hipMemcpyDeviceToHost, s));
        SSSP_HIP(hipStreamSynchronize(s));
    });
}

int sssp_solve(sssp_graph *g, int32_t source, int32_t schedule, void *dist_out, sssp_stats *stats)
{
    return guard([&] {
        SSSP_REQUIRE(g, "null graph");
        SSSP_REQUIRE(source >= 0 && source < g->V, "source vertex out of range");
        SSSP_REQUIRE(schedule >= SSSP_SCHED_PLAIN && schedule <= SSSP_SCHED_PULL, "unknown schedule");
        set_device(g->device);
        sssp_stats st;
        std::memset(&st, 0, sizeof(st));
        g->have_result = false;
        g->bfs_result = false;
        const int32_t slot = slot_of(g, source);
        solve(g, &slot, 1, schedule, &st);
        g->last_source = source;
        g->last_sched = schedule;
        g->have_result = true;
        if (dist_out)
            SSSP_HIP(hipMemcpy(dist_out, g->dout, 4 * (size_t)g->V, hipMemcpyDeviceToHost));
        if (stats) *stats = st;
    });
}

int sssp_bfs(sssp_graph *g, int32_t source, int32_t schedule, uint8_t *visited_out,
             uint32_t *levels_out, sssp_stats *stats)
{
    return guard([&] {
        SSSP_REQUIRE(g, "null graph");
        SSSP_REQUIRE(source >= 0 && source < g->V, "source vertex out of range");
        SSSP_REQUIRE(schedule >= SSSP_SCHED_PLAIN && schedule <= SSSP_SCHED_PULL, "unknown schedule");
        set_device(g->device);
        sssp_stats st;
        std::memset(&st, 0, sizeof(st));
        g->have_result = false;
        g->bfs_result = false;
        bfs(g, slot_of(g, source), schedule, &st);
        g->last_source = source;
        g->last_sched = schedule;
        g->bfs_result = true;
        if (levels_out)
            SSSP_HIP(hipMemcpy(levels_out, g->dout, 4 * (size_t)g->V, hipMemcpyDeviceToHost));
        if (visited_out) {
            uint8_t *dv = dmalloc<uint8_t>((size_t)g->V);
            DevBuf own;
            own.p[0] = dv;
            levels_to_visited(g->dout, g->V, dv, g->stream);
            SSSP_HIP(hipStreamSynchronize(g->stream));
            SSSP_HIP(hipMemcpy(visited_out, dv, (size_t)g->V, hipMemcpyDeviceToHost));
        }
        if (stats) *stats = st;
    });
}

int sssp_bfs_edges(const int32_t *d_src, const int32_t *d_dst, int32_t vertices, int64_t edges,
                   uint8_t *d_visited, int32_t source, int32_t *passes)
{
    return guard([&] {
        SSSP_REQUIRE(d_visited && vertices >= 1 && edges >= 0 && (edges == 0 || (d_src && d_dst)),
                     "bad arguments");
        SSSP_REQUIRE(source >= 0 && source < vertices, "source vertex out of range");
        validate_edges(d_src, d_dst, nullptr, -1, edges, vertices, nullptr);
        ref_bfs(d_src, d_dst, vertices, edges, d_visited, source, passes);
    });
}

int sssp_get_distances(sssp_graph *g, void *dist_out)
{
    return guard([&] {
        SSSP_REQUIRE(g && dist_out, "null argument");
        SSSP_REQUIRE(g->have_result, "no solve result on this graph");
        set_device(g->device);
        SSSP_HIP(hipMemcpy(dist_out, g->dout, 4 * (size_t)g->V, hipMemcpyDeviceToHost));
    });
}

int sssp_get_trace(sssp_graph *g, int32_t cap, int32_t *modes, int64_t *frontier,
                   int64_t *frontier_edges, float *round_ms, int32_t *rounds)
{
    return guard([&] {
        SSSP_REQUIRE(g && rounds, "null argument");
        const int32_t n = (int32_t)g->trace.size();
        *rounds = n;
        for (int32_t i = 0; i < n && i < cap; i++) {
            const TraceRow &t = g->trace[(size_t)i];
            if (modes) modes[i] = t.mode;
            if (frontier) frontier[i] = t.frontier;
            if (frontier_edges) frontier_edges[i] = t.frontier_edges;
            if (round_ms) round_ms[i] = t.ms;
        }
    });
}

int sssp_certify(sssp_graph *g, int32_t source, int32_t *parent_out, int64_t *counts,
                 int32_t *valid)
{
    return guard([&] {
        SSSP_REQUIRE(g && counts, "null argument");
        SSSP_REQUIRE(g->have_result || g->bfs_result, "no solve result on this graph");
        SSSP_REQUIRE(source >= 0 && source < g->V, "source vertex out of range");
        set_device(g->device);
        const int32_t slot = slot_of(g, source);
        if (g->bfs_result) certify_levels(g, slot, counts);  // hop levels: every edge weighs 1
        else certify(g, &slot, 1, counts);
        if (parent_out) parents_out(g, parent_out);
        if (valid) *valid = (counts[0] == 0 && counts[1] == 0 && counts[2] == 0) ? 1 : 0;
    });
}

int sssp_set_param(sssp_graph *g, const char *name, double value)
{
    return guard([&] {
        SSSP_REQUIRE(g && name, "null argument");
        const std::string n(name);
        auto chunk = [&](int &dst) {
            SSSP_REQUIRE(value >= 0 && value < kChunkCount, "chunk bucket must be in [0, 4)");
            dst = (int)value;
        };
        if (n == "push_chunk") chunk(g->push_chunk);
        else if (n == "all_chunk") chunk(g->all_chunk);
        else if (n == "pull_chunk") chunk(g->pull_chunk);
        else if (n == "advance_blocks") {
            SSSP_REQUIRE(value >= 0 && value <= 1 << 20, "advance_blocks out of range");
            g->advance_blocks = (int)value;
        } else if (n == "verify") g->verify = value != 0;
        else if (n == "max_rounds") {
            // round ticks of a solve (and twice that for a timing collection)
            // must stay below 2^30
            SSSP_REQUIRE(value >= 1 && value <= (double)(1 << 28), "max_rounds must be in [1, 2^28]");
            g->max_rounds = (int)value;
        } else if (n == "push_edge_frac") g->push_edge_frac = value;
        else if (n == "pull_frac") g->pull_frac = value;
        else if (n == "delta") {
            SSSP_REQUIRE(value >= 0, "delta must be >= 0");
            g->delta = value;
        } else if (n == "collect_kind") {
            SSSP_REQUIRE(value == 0 || value == 1, "collect_kind must be 0 (SSSP) or 1 (BFS)");
            g->collect_kind = (int)value;
        } else if (n == "collect_budget") {
            SSSP_REQUIRE(value >= 1 && value <= 9.0e15, "collect_budget must be >= 1");
            g->collect_budget = (int64_t)value;
        } else if (n == "rollout_policy") {
            SSSP_REQUIRE(value == 0 || value == 1, "rollout_policy must be 0 (rule) or 1 (network)");
            g->rollout_policy = (int)value;
        } else if (n == "dense_kernel") {
            SSSP_REQUIRE(value == 0 || value == 1, "dense_kernel must be 0 or 1");
            g->dense_kernel = (int)value;
        } else if (n == "blk_tile_cap") {
            SSSP_REQUIRE(value >= (double)kBlkChunk && value <= 2147483648.0 &&
                             std::fmod(value, (double)kBlkChunk) == 0,
                         "blk_tile_cap must be a multiple of 256 in [256, 2^31]");
            retile_blocked(g, (int64_t)value, g->blk_tile_order, g->blk_range_bits);
        } else if (n == "blk_tile_order") {
            SSSP_REQUIRE(value == 0 || value == 1 || value == 2 || value == 3,
                         "blk_tile_order must be 0, 1, 2 or 3");
            retile_blocked(g, g->blk_tile_cap, (int)value, g->blk_range_bits);
        } else if (n == "blk_range_bits") {
            SSSP_REQUIRE(value == 0 || (value >= kBlkRangeFineBits && value <= 31),
                         "blk_range_bits must be 0 or in [18, 31]");
            retile_blocked(g, g->blk_tile_cap, g->blk_tile_order, (int)value);
        } else if (n == "decide_spin") {
            SSSP_REQUIRE(value == 0 || value == 1, "decide_spin must be 0 or 1");
            g->decide_spin = (int)value;
        } else if (n == "batch_capture") {
            SSSP_REQUIRE(value == 0 || value == 1, "batch_capture must be 0 or 1");
            g->batch_capture = (int)value;
        } else if (n == "plain_kernel") {
            SSSP_REQUIRE(value == 0 || value == 1 || value == 2, "plain_kernel must be 0, 1 or 2");
            g->plain_kernel = (int)value;
        } else if (n == "dense_unreached") {
            SSSP_REQUIRE(value >= 0 && value <= 2, "dense_unreached must be in [0, 2]");
            g->dense_unreached = value;
        } else if (n == "dense_all_frac") {
            SSSP_REQUIRE(value >= 0 && value <= 1e9, "dense_all_frac must be >= 0");
            g->dense_all_frac = value;
        } else if (n == "plain_below_edges") {
            SSSP_REQUIRE(value >= 0 && value <= 9.2e18, "plain_below_edges must be >= 0");
            g->plain_below_edges = (int64_t)value;
        } else if (n == "hub_degree") {
            SSSP_REQUIRE(value >= 0 && value <= 2147483647.0, "hub_degree must be in [0, 2^31)");
            g->hub_degree = (int)value;
        } else if (n == "tail_edges") {
            SSSP_REQUIRE(value >= 0 && value <= 9.2e18, "tail_edges must be >= 0");
            g->tail_edges = (int64_t)value;
        } else if (n == "tail_group") {
            SSSP_REQUIRE(value >= 1 && value <= 64, "tail_group must be in [1, 64]");
            g->tail_group = (int)value;
        } else if (n == "push_bits_edges") {
            SSSP_REQUIRE(value >= 0 && value <= 9.2e18, "push_bits_edges must be >= 0");
            g->push_bits_edges = (int64_t)value;
        } else if (n == "tail_blocks") {
            SSSP_REQUIRE(value >= 0 && value <= 65536, "tail_blocks must be in [0, 65536]");
            g->tail_blocks = (int)value;
        } else if (n == "trace_events") g->trace_events = value != 0;
        else throw Error{SSSP_ERR_ARG, "unknown parameter '" + n + "'"};
    });
}

int sssp_get_param(const sssp_graph *g, const char *name, double *value)
{
    return guard([&] {
        SSSP_REQUIRE(g && name && value, "null argument");
        const std::string n(name);
        if (n == "push_chunk") *value = g->push_chunk;
        else if (n == "all_chunk") *value = g->all_chunk;
        else if (n == "pull_chunk") *value = g->pull_chunk;
        else if (n == "advance_blocks") *value = g->advance_blocks;
        else if (n == "verify") *value = g->verify;
        else if (n == "max_rounds") *value = g->max_rounds;
        else if (n == "push_edge_frac") *value = g->push_edge_frac;
        else if (n == "pull_frac") *value = g->pull_frac;
        else if (n == "delta") *value = g->delta;
        else if (n == "hub_degree") *value = g->hub_degree;
        else if (n == "plain_below_edges") *value = (double)g->plain_below_edges;
        else if (n == "dense_all_frac") *value = g->dense_all_frac;
        else if (n == "dense_unreached") *value = g->dense_unreached;
        else if (n == "dense_kernel") *value = g->dense_kernel;
        else if (n == "batch_capture") *value = g->batch_capture;
        else if (n == "decide_spin") *value = g->decide_spin;
        else if (n == "blk_tile_cap") *value = (double)g->blk_tile_cap;
        else if (n == "blk_tile_order") *value = g->blk_tile_order;
        else if (n == "blk_tiles") *value = g->n_btile;
        else if (n == "blk_wide_chunks") *value = (double)g->b_nwide;
        else if (n == "blk_range_bits") *value = g->blk_range_bits;
        else if (n == "plain_kernel") *value = g->plain_kernel;
        else if (n == "plain_blocked") *value = plain_uses_blocked(g) ? 1 : 0;
        else if (n == "tail_edges") *value = (double)g->tail_edges;
        else if (n == "tail_group") *value = g->tail_group;
        else if (n == "tail_blocks") *value = g->tail_blocks;
        else if (n == "push_bits_edges") *value = (double)g->push_bits_edges;
        else if (n == "multi_passes") *value = g->multi_passes;
        else if (n == "rollout_policy") *value = g->rollout_policy;
        else if (n == "collect_kind") *value = g->collect_kind;
        else if (n == "collect_budget") *value = (double)g->collect_budget;
        else if (n == "collect_rounds") *value = (double)g->collect_rounds;
        else if (n == "trace_events") *value = g->trace_events;
        else if (n == "cu_count") *value = g->cu_count;
        else throw Error{SSSP_ERR_ARG, "unknown parameter '" + n + "'"};
    });
}

// ---- reference-shaped wrappers --------------------------------------------
int sssp_init_distances_f32(float *d_dist, int32_t vertices, int32_t source)
{
    return guard([&] {
        SSSP_REQUIRE(d_dist && vertices >= 1 && source >= 0 && source < vertices, "bad arguments");
        ref_init<WF32, uint32_t>(reinterpret_cast<uint32_t *>(d_dist), vertices, source);
    });
}

int sssp_init_distances_u32(uint32_t *d_dist, int32_t vertices, int32_t source)
{
    return guard([&] {
        SSSP_REQUIRE(d_dist && vertices >= 1 && source >= 0 && source < vertices, "bad arguments");
        ref_init<WU32, uint32_t>(d_dist, vertices, source);
    });
}

int sssp_init_distances_f64(double *d_dist, int32_t vertices, int32_t source)
{
    return guard([&] {
        SSSP_REQUIRE(d_dist && vertices >= 1 && source >= 0 && source < vertices, "bad arguments");
        ref_init<WF64, unsigned long long>(reinterpret_cast<unsigned long long *>(d_dist), vertices,
                                           source);
    });
}

int sssp_bellman_ford_f32(const int32_t *d_src, const int32_t *d_dst, const float *d_w,
                          int32_t vertices, int64_t edges, int32_t source, float *d_dist,
                          int32_t *passes)
{
    return guard([&] {
        SSSP_REQUIRE(d_dist && vertices >= 1 && edges >= 0 && (edges == 0 || (d_src && d_dst && d_w)),
                     "bad arguments");
        SSSP_REQUIRE(source >= 0 && source < vertices, "source vertex out of range");
        validate_edges(d_src, d_dst, d_w, SSSP_W_F32, edges, vertices, nullptr);
        ref_bf<WF32, uint32_t>(d_src, d_dst, reinterpret_cast<const uint32_t *>(d_w), vertices, edges,
                               reinterpret_cast<uint32_t *>(d_dist), passes);
    });
}

int sssp_bellman_ford_noalloc_f32(const int32_t *d_src, const int32_t *d_dst, const float *d_w,
                                  int32_t vertices, int64_t edges, int32_t source, float *d_dist,
                                  int32_t *passes)
{
    return guard([&] {
        SSSP_REQUIRE(d_dist && vertices >= 1 && edges >= 0 && (edges == 0 || (d_src && d_dst && d_w)),
                     "bad arguments");
        SSSP_REQUIRE(source >= 0 && source < vertices, "source vertex out of range");
        ref_bf_noalloc_f32(d_src, d_dst, reinterpret_cast<const uint32_t *>(d_w), vertices, edges,
                           reinterpret_cast<uint32_t *>(d_dist), passes);
    });
}

int sssp_bellman_ford_u32(const int32_t *d_src, const int32_t *d_dst, const uint32_t *d_w,
                          int32_t vertices, int64_t edges, int32_t source, uint32_t *d_dist,
                          int32_t *passes)
{
    return guard([&] {
        SSSP_REQUIRE(d_dist && vertices >= 1 && edges >= 0 && (edges == 0 || (d_src && d_dst && d_w)),
                     "bad arguments");
        SSSP_REQUIRE(source >= 0 && source < vertices, "source vertex out of range");
        validate_edges(d_src, d_dst, d_w, SSSP_W_U32, edges, vertices, nullptr);
        ref_bf<WU32, uint32_t>(d_src, d_dst, d_w, vertices, edges, d_dist, passes);
    });
}

int sssp_bellman_ford_f64(const int32_t *d_src, const int32_t *d_dst, const double *d_w,
                          int32_t vertices, int64_t edges, int32_t source, double *d_dist,
                          int32_t *passes)
{
    return guard([&] {
        SSSP_REQUIRE(d_dist && vertices >= 1 && edges >= 0 && (edges == 0 || (d_src && d_dst && d_w)),
                     "bad arguments");
        SSSP_REQUIRE(source >= 0 && source < vertices, "source vertex out of range");
        validate_edges(d_src, d_dst, d_w, SSSP_W_F64, edges, vertices, nullptr);
        ref_bf<WF64, unsigned long long>(d_src, d_dst,
                                         reinterpret_cast<const unsigned long long *>(d_w), vertices,
                                         edges, reinterpret_cast<unsigned long long *>(d_dist),
                                         passes);
    });
}

int sssp_collect_timings(sssp_graph *g, int32_t source, int32_t cap, float *features,
                         float *times_ms, int32_t *rows)
{
    return guard([&] {
        SSSP_REQUIRE(g && rows && cap >= 0 && (cap == 0 || (features && times_ms)), "bad arguments");
        SSSP_REQUIRE(source >= 0 && source < g->V, "source vertex out of range");
        set_device(g->device);
        g->have_result = false;
        g->bfs_result = false;
        collect_timings(g, slot_of(g, source), cap, features, times_ms, rows);
        g->last_source = source;
        if (g->collect_kind == 1) g->bfs_result = true;  // the collected solve was a BFS
        else g->have_result = true;
    });
}

int sssp_features(double frontier_vertices, double frontier_edges, double unsettled_frac,
                  double round_index, double mean_degree, double vertices, double edges,
                  float *out5)
{
    return guard([&] {
        SSSP_REQUIRE(out5, "null output");
        make_features(frontier_vertices, frontier_edges, unsettled_frac, round_index, mean_degree,
                      vertices, edges, out5);
    });
}

}  // extern "C"
