// gen.hip -- device-side R-MAT generator, bit-identical to the host restatement
// of GraphGenerationAPI::R_MAT_parallel (graph_generation_API.hpp:404-484)
// with a fixed base seed and virtual thread count.
//
// Every edge consumes exactly n+2 rand_r() calls (n-1 quadrant draws, two coin
// flips, one weight) and every rand_r() is three steps of the LCG
// x <- 1103515245 x + 12345 (mod 2^32). So the generator state at edge j of
// virtual thread k is seed_k advanced by 3 (n+2) (j - begin_k) steps, which an
// affine jump-ahead computes in O(log) time: each lane jumps once to the start
// of its run of edges and then generates sequentially, writing straight into
// HBM.
#include "sssp_graph.hpp"

namespace sssp {

namespace {

constexpr uint32_t kLcgA = 1103515245u, kLcgC = 12345u;

struct Affine {
    uint32_t a, c;  // x -> a x + c
};

__host__ __device__ inline Affine compose(Affine f, Affine g)  // g after f
{
    return Affine{g.a * f.a, g.a * f.c + g.c};
}

__host__ __device__ inline uint32_t lcg_jump(uint32_t x, uint64_t k)
{
    Affine acc{1u, 0u}, step{kLcgA, kLcgC};
    while (k) {
        if (k & 1) acc = compose(acc, step);
        step = compose(step, step);
        k >>= 1;
    }
    return acc.a * x + acc.c;
}

__device__ __forceinline__ int32_t dev_rand_r(uint32_t &next)
{
    uint32_t result;
    next = next * kLcgA + kLcgC;
    result = (next / 65536u) % 2048u;
    next = next * kLcgA + kLcgC;
    result <<= 10;
    result ^= (next / 65536u) % 1024u;
    next = next * kLcgA + kLcgC;
    result <<= 10;
    result ^= (next / 65536u) % 1024u;
    return (int32_t)result;
}

__global__ __launch_bounds__(256) void k_rmat(int n, int32_t V, int64_t E, int T, uint32_t base_seed,
                                              int a, int b, int c, int d, int64_t per_lane,
                                              int32_t *src, int32_t *dst, float *w)
{
    const int64_t lane_id = (int64_t)blockIdx.x * 256 + threadIdx.x;
    int64_t e = lane_id * per_lane;
    const int64_t e_end = min(e + per_lane, E);
    if (e >= e_end) return;
    const int64_t q = E / T, r = E % T;
    const int a_end = a, b_end = a + b, c_end = a + b + c, d_end = a + b + c + d;
    const uint64_t steps_per_edge = 3ull * (uint64_t)(n + 2);
    while (e < e_end) {
        // virtual thread owning edge e and the end of its chunk
        int64_t k = e / (q + 1);
        if (k >= r) k = r + (e - r * (q + 1)) / (q > 0 ? q : 1);
        if (k >= T) k = T - 1;
        const int64_t kb = k * q + (k < r ? k : r);
        const int64_t ke = kb + q + (k < r ? 1 : 0);
        uint32_t st = lcg_jump((uint32_t)((int64_t)base_seed * k), steps_per_edge * (uint64_t)(e - kb));
        const int64_t stop = min(e_end, ke);
        for (; e < stop; e++) {
            int x = V / 2, y = V / 2;
            for (int i = 1; i < n; i++) {
                const int step = 1 << (n - (i + 1));
                const int p = dev_rand_r(st) % 100;
                const bool in_any = p < d_end;
                const bool xplus = p >= b_end;
                const bool yplus = (p >= a_end && p < b_end) || p >= c_end;
                x += in_any ? (xplus ? step : -step) : 0;
                y += in_any ? (yplus ? step : -step) : 0;
            }
            if (dev_rand_r(st) % 2 == 0) x--;
            if (dev_rand_r(st) % 2 == 0) y--;
            src[e] = x;
            dst[e] = y;
            // float(rand_r) / float(RAND_MAX): the divisor is 2^31, so the
            // quotient is the exact scaling of the RNE conversion
            w[e] = __int2float_rn(dev_rand_r(st)) * 0x1p-31f;
        }
    }
}

}  // namespace

void rmat_device(int scale, int ef, int a, int b, int c, int d, uint32_t base_seed, int T,
                 int32_t *src, int32_t *dst, float *w, hipStream_t s)
{
    const int32_t V = 1 << scale;
    const int64_t E = (int64_t)V * ef;
    const int64_t per_lane = 64;
    const int64_t lanes = (E + per_lane - 1) / per_lane;
    const unsigned grid = (unsigned)((lanes + 255) / 256);
    if (E == 0) return;
    k_rmat<<<grid, 256, 0, s>>>(scale, V, E, T, base_seed, a, b, c, d, per_lane, src, dst, w);
    SSSP_LAUNCH_CHECK("k_rmat");
}

}  // namespace sssp
