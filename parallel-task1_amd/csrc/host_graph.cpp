// host_graph.cpp -- host services of the SSSP library: the seeded graph
// generators the reference uses for its inputs, the binary edge-list format,
// and source selection. Pure host C++ (OpenMP), no device code.
#include <omp.h>

#include <cerrno>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "sssp_internal.hpp"

namespace {

// glibc rand_r (stdlib/rand_r.c): three steps of the LCG
// next = next*1103515245 + 12345, taking 11 + 10 + 10 bits of next/65536.
// Restated so the generated edge sequence does not depend on the host libc;
// tests compare it with the libc function.
inline int32_t lcg_rand_r(uint32_t *seed)
{
    uint32_t next = *seed;
    uint32_t result;
    next = next * 1103515245u + 12345u;
    result = (next / 65536u) % 2048u;
    next = next * 1103515245u + 12345u;
    result <<= 10;
    result ^= (next / 65536u) % 1024u;
    next = next * 1103515245u + 12345u;
    result <<= 10;
    result ^= (next / 65536u) % 1024u;
    *seed = next;
    return (int32_t)result;
}

// glibc rand()/srand() (stdlib/random_r.c, TYPE_3: degree 31, separation 3):
// r[0] = seed (0 -> 1), r[i] = 16807 * r[i-1] mod (2^31 - 1) for i < 31
// (Schrage's method), then r[k] = r[k-31] + r[k-3] mod 2^32 with the first
// 310 outputs discarded; each output is r[k] >> 1.
struct GlibcRand {
    int32_t r[34];
    int f, b;  // front / rear indices into the 31-word ring
    uint32_t ring[31];
    explicit GlibcRand(uint32_t seed)
    {
        int32_t word = (int32_t)(seed == 0 ? 1 : seed);
        ring[0] = (uint32_t)word;
        for (int i = 1; i < 31; i++) {
            int32_t hi = word / 127773;
            int32_t lo = word % 127773;
            word = 16807 * lo - 2836 * hi;
            if (word < 0) word += 2147483647;
            ring[i] = (uint32_t)word;
        }
        f = 3;
        b = 0;
        for (int i = 0; i < 310; i++) next();
    }
    int32_t next()
    {
        ring[f] += ring[b];
        int32_t result = (int32_t)(ring[f] >> 1);
        if (++f >= 31) f = 0;
        if (++b >= 31) b = 0;
        return result;
    }
};

constexpr float kRandMaxF = 2147483648.0f;  // static_cast<float>(RAND_MAX)

// One R-MAT edge (graph_generation_API.hpp:435-470): n-1 quadrant draws with
// integer percentages, two coin flips, one weight draw.
inline void rmat_edge(uint32_t *seed, int n, int32_t V, int a, int b, int c, int d,
                      int32_t *from, int32_t *to, float *w)
{
    int x = V / 2, y = V / 2;
    const int a_end = a, b_end = a + b, c_end = a + b + c, d_end = a + b + c + d;
    for (int i = 1; i < n; i++) {
        const int step = 1 << (n - (i + 1));
        const int p = lcg_rand_r(seed) % 100;
        // quadrant a: (x-, y-), b: (x-, y+), c: (x+, y-), d: (x+, y+), else no
        // move (graph_generation_API.hpp:446-461), evaluated without branches
        const bool in_any = p < d_end;
        const bool xplus = p >= b_end;
        const bool yplus = (p >= a_end && p < b_end) || p >= c_end;
        x += in_any ? (xplus ? step : -step) : 0;
        y += in_any ? (yplus ? step : -step) : 0;
    }
    if (lcg_rand_r(seed) % 2 == 0) x--;
    if (lcg_rand_r(seed) % 2 == 0) y--;
    *from = x;
    *to = y;
    *w = static_cast<float>(lcg_rand_r(seed)) / kRandMaxF;
}

void rmat_fill(int scale, int ef, int a, int b, int c, int d, uint32_t base_seed, int T,
               int32_t *src, int32_t *dst, float *w)
{
    const int32_t V = 1 << scale;
    const int64_t E = (int64_t)V * ef;
    // OpenMP schedule(static) without a chunk size gives thread k the k-th of
    // T contiguous, near-equal chunks (the first E % T get one more).
    const int64_t q = E / T, r = E % T;
#pragma omp parallel for schedule(dynamic, 1)
    for (int k = 0; k < T; k++) {
        int64_t begin = k * q + (k < r ? k : r);
        int64_t end = begin + q + (k < r ? 1 : 0);
        uint32_t seed = (uint32_t)((int64_t)base_seed * k);
        for (int64_t e = begin; e < end; e++)
            rmat_edge(&seed, scale, V, a, b, c, d, &src[e], &dst[e], &w[e]);
    }
}

}  // namespace

extern "C" {

int sssp_rmat_generate(int32_t scale, int32_t edge_factor, int32_t a, int32_t b, int32_t c,
                       int32_t d, uint32_t base_seed, int32_t virtual_threads, int32_t *src,
                       int32_t *dst, float *weights)
{
    return sssp::guard([&] {
        SSSP_REQUIRE(scale >= 1 && scale <= 30, "rmat: scale must be in [1, 30]");
        SSSP_REQUIRE(edge_factor >= 1, "rmat: edge_factor must be >= 1");
        SSSP_REQUIRE(a >= 0 && b >= 0 && c >= 0 && d >= 0 && a + b + c + d <= 100,
                     "rmat: probabilities are integer percents summing to <= 100");
        SSSP_REQUIRE(virtual_threads >= 1, "rmat: virtual_threads must be >= 1");
        SSSP_REQUIRE(src && dst && weights, "rmat: null output array");
        SSSP_REQUIRE((int64_t)(1 << scale) * edge_factor < (int64_t)1 << 31,
                     "rmat: edge count must fit the reference's int product");
        rmat_fill(scale, edge_factor, a, b, c, d, base_seed, virtual_threads, src, dst, weights);
    });
}

int sssp_rmat_generate_many(int32_t n, const int32_t *scales, int32_t edge_factor, int32_t a,
                            int32_t b, int32_t c, int32_t d, uint32_t seed0,
                            const int64_t *edge_offsets, int32_t *src, int32_t *dst,
                            float *weights)
{
    return sssp::guard([&] {
        SSSP_REQUIRE(n >= 1 && scales && edge_offsets, "rmat_many: bad arguments");
        for (int k = 0; k < n; k++) {
            SSSP_REQUIRE(scales[k] >= 1 && scales[k] <= 24, "rmat_many: scale out of range");
            SSSP_REQUIRE(edge_offsets[k + 1] - edge_offsets[k] ==
                             ((int64_t)1 << scales[k]) * edge_factor,
                         "rmat_many: edge_offsets do not match the scales");
        }
#pragma omp parallel for schedule(dynamic, 1)
        for (int k = 0; k < n; k++) {
            int64_t o = edge_offsets[k];
            // one virtual thread per graph: seed = (seed0 + k) * 0 for the
            // only chunk would make every graph identical, so graph k uses
            // the stream of virtual thread 1 of base seed (seed0 + k).
            uint32_t seed = seed0 + (uint32_t)k;
            int32_t V = 1 << scales[k];
            int64_t E = (int64_t)V * edge_factor;
            for (int64_t e = 0; e < E; e++)
                rmat_edge(&seed, scales[k], V, a, b, c, d, &src[o + e], &dst[o + e], &weights[o + e]);
        }
    });
}

int sssp_er_generate(int32_t vertices, int32_t avg_degree, uint32_t seed, int32_t weight_mode,
                     int32_t *src, int32_t *dst, void *weights)
{
    return sssp::guard([&] {
        SSSP_REQUIRE(vertices >= 1 && avg_degree >= 0, "er: bad sizes");
        SSSP_REQUIRE(avg_degree <= vertices,
                     "average connections in graph is greater than number of vertices");
        SSSP_REQUIRE(weight_mode >= 0 && weight_mode <= 2, "er: weight_mode must be 0, 1 or 2");
        GlibcRand rng(seed);
        // vertex values first (graph_generation_API.hpp:22-26): one draw each
        for (int32_t v = 0; v < vertices; v++) (void)rng.next();
        const int64_t E = (int64_t)vertices * avg_degree;
        float *wf = static_cast<float *>(weights);
        uint32_t *wu = static_cast<uint32_t *>(weights);
        for (int64_t e = 0; e < E; e++) {
            int32_t from = rng.next() % vertices;
            int32_t to = rng.next() % vertices;
            int32_t r = rng.next();
            src[e] = from;
            dst[e] = to;
            if (weight_mode == 0)
                wf[e] = (float)(r % 10000) / 10000;  // rand_uniform_val<float>(10000)
            else if (weight_mode == 1)
                wu[e] = 1u + (uint32_t)(r % 255);
            else
                wu[e] = (uint32_t)(r % 10000);
        }
    });
}

int sssp_save_bin(const char *path, int32_t directed, const int32_t *src, const int32_t *dst,
                  const void *weights, int32_t weight_bytes, int32_t vertices, int64_t edges)
{
    return sssp::guard([&] {
        SSSP_REQUIRE(path && (edges == 0 || (src && dst && weights)), "save_bin: null argument");
        SSSP_REQUIRE(weight_bytes == 4 || weight_bytes == 8, "save_bin: weight_bytes must be 4 or 8");
        FILE *f = std::fopen(path, "wb");
        if (!f) throw sssp::Error{SSSP_ERR_IO, std::string("cannot open ") + path};
        const uint8_t dir = directed ? 1 : 0;
        bool ok = std::fwrite(&dir, 1, 1, f) == 1 && std::fwrite(&vertices, 4, 1, f) == 1 &&
                  std::fwrite(&edges, 8, 1, f) == 1;
        // packed records, written in blocks
        const int64_t B = 1 << 16;
        const int rec = 8 + weight_bytes;
        std::vector<uint8_t> buf((size_t)B * rec);
        const uint8_t *wb = static_cast<const uint8_t *>(weights);
        for (int64_t e0 = 0; ok && e0 < edges; e0 += B) {
            int64_t n = edges - e0 < B ? edges - e0 : B;
            for (int64_t i = 0; i < n; i++) {
                uint8_t *p = &buf[(size_t)i * rec];
                std::memcpy(p, &src[e0 + i], 4);
                std::memcpy(p + 4, &dst[e0 + i], 4);
                std::memcpy(p + 8, wb + (size_t)(e0 + i) * weight_bytes, weight_bytes);
            }
            ok = std::fwrite(buf.data(), rec, (size_t)n, f) == (size_t)n;
        }
        if (std::fclose(f) != 0) ok = false;
        if (!ok) throw sssp::Error{SSSP_ERR_IO, std::string("short write to ") + path};
    });
}

int sssp_load_bin_header(const char *path, int32_t *directed, int32_t *vertices, int64_t *edges)
{
    return sssp::guard([&] {
        SSSP_REQUIRE(path && directed && vertices && edges, "load_bin_header: null argument");
        FILE *f = std::fopen(path, "rb");
        if (!f) throw sssp::Error{SSSP_ERR_IO, std::string("cannot open ") + path};
        uint8_t dir = 0;
        int32_t V = 0;
        int64_t E = 0;
        bool ok = std::fread(&dir, 1, 1, f) == 1 && std::fread(&V, 4, 1, f) == 1 &&
                  std::fread(&E, 8, 1, f) == 1;
        std::fclose(f);
        if (!ok) throw sssp::Error{SSSP_ERR_IO, std::string("truncated header in ") + path};
        if (V < 0 || E < 0) throw sssp::Error{SSSP_ERR_IO, std::string("corrupt header in ") + path};
        *directed = dir ? 1 : 0;
        *vertices = V;
        *edges = E;
    });
}

int sssp_load_bin(const char *path, int32_t weight_bytes, int32_t *src, int32_t *dst,
                  void *weights, int64_t edges_capacity)
{
    return sssp::guard([&] {
        SSSP_REQUIRE(path, "load_bin: null path");
        SSSP_REQUIRE(weight_bytes == 4 || weight_bytes == 8, "load_bin: weight_bytes must be 4 or 8");
        FILE *f = std::fopen(path, "rb");
        if (!f) throw sssp::Error{SSSP_ERR_IO, std::string("cannot open ") + path};
        uint8_t dir = 0;
        int32_t V = 0;
        int64_t E = 0;
        bool ok = std::fread(&dir, 1, 1, f) == 1 && std::fread(&V, 4, 1, f) == 1 &&
                  std::fread(&E, 8, 1, f) == 1;
        if (!ok) {
            std::fclose(f);
            throw sssp::Error{SSSP_ERR_IO, std::string("truncated header in ") + path};
        }
        if (E > edges_capacity || E < 0) {
            std::fclose(f);
            throw sssp::Error{SSSP_ERR_ARG, "load_bin: file holds more edges than the buffers"};
        }
        const int64_t B = 1 << 16;
        const int rec = 8 + weight_bytes;
        std::vector<uint8_t> buf((size_t)B * rec);
        uint8_t *wb = static_cast<uint8_t *>(weights);
        for (int64_t e0 = 0; ok && e0 < E; e0 += B) {
            int64_t n = E - e0 < B ? E - e0 : B;
            ok = std::fread(buf.data(), rec, (size_t)n, f) == (size_t)n;
            for (int64_t i = 0; ok && i < n; i++) {
                const uint8_t *p = &buf[(size_t)i * rec];
                std::memcpy(&src[e0 + i], p, 4);
                std::memcpy(&dst[e0 + i], p + 4, 4);
                std::memcpy(wb + (size_t)(e0 + i) * weight_bytes, p + 8, weight_bytes);
            }
        }
        std::fclose(f);
        if (!ok) throw sssp::Error{SSSP_ERR_IO, std::string("truncated edge records in ") + path};
    });
}

int sssp_pick_sources(const int32_t *src, int64_t edges, int32_t vertices, uint64_t seed,
                      int32_t n, int32_t *out)
{
    return sssp::guard([&] {
        SSSP_REQUIRE(vertices >= 1 && n >= 0 && (n == 0 || out), "pick_sources: bad arguments");
        std::vector<uint8_t> has_out((size_t)vertices, 0);
        int64_t bad = 0;
        uint8_t *ho = has_out.data();
#pragma omp parallel for schedule(static) reduction(+ : bad)
        for (int64_t e = 0; e < edges; e++) {
            const int32_t s = src[e];
            if (s < 0 || s >= vertices) {
                bad++;
                continue;
            }
            __atomic_store_n(&ho[s], (uint8_t)1, __ATOMIC_RELAXED);
        }
        SSSP_REQUIRE(bad == 0, "pick_sources: src id out of range");
        int64_t cand = 0;
        for (int32_t v = 0; v < vertices; v++) cand += has_out[v];
        SSSP_REQUIRE(cand >= n, "pick_sources: fewer vertices with out-degree > 0 than requested");
        std::mt19937_64 rng(seed);
        std::vector<uint8_t> taken((size_t)vertices, 0);
        for (int32_t i = 0; i < n; i++) {
            int32_t v;
            do {
                v = (int32_t)(rng() % (uint64_t)vertices);
            } while (!has_out[v] || taken[v]);
            taken[v] = 1;
            out[i] = v;
        }
    });
}

}  // extern "C"
