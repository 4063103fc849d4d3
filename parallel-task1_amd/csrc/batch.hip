// batch.hip -- batched solves (many sources on one graph, many small graphs).
#include "sssp_graph.hpp"

using namespace sssp;

struct sssp_pack {
    int device = 0;
};

extern "C" {

int sssp_solve_batch(sssp_graph *g, const int32_t *sources, int32_t n, int32_t policy,
                     void *dist_out, sssp_stats *stats)
{
    return fail(SSSP_ERR_INTERNAL, "batched solve not built yet");
}

int sssp_pack_create(const int32_t *src, const int32_t *dst, const float *weights,
                     const int64_t *edge_offsets, const int32_t *vertex_counts, int32_t n_graphs,
                     int32_t device, sssp_pack **out)
{
    return fail(SSSP_ERR_INTERNAL, "pack not built yet");
}

int sssp_pack_destroy(sssp_pack *p) { return fail(SSSP_ERR_INTERNAL, "pack not built yet"); }

int sssp_pack_solve(sssp_pack *p, const int32_t *sources, int32_t policy, float *dist_out,
                    sssp_stats *stats)
{
    return fail(SSSP_ERR_INTERNAL, "pack not built yet");
}

int sssp_autotune_batch(sssp_graph *g, const int32_t *sources, int32_t n, const char *tuning_file,
                        double *best_ms)
{
    return fail(SSSP_ERR_INTERNAL, "autotune not built yet");
}

int sssp_autotune_pack(sssp_pack *p, const int32_t *sources, const char *tuning_file,
                       double *best_ms)
{
    return fail(SSSP_ERR_INTERNAL, "autotune not built yet");
}

}  // extern "C"
