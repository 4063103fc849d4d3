// sssp_internal.hpp -- shared types, error plumbing and device helpers of the
// MI355X (gfx950) Bellman-Ford SSSP library. Not part of the public C ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/sssp_mi355x.h"

// ---------------------------------------------------------------------------
// Error plumbing: every C entry point returns a status code; the message of
// the last failure on the calling thread is kept for sssp_last_error().
// (Replaces SAFE_CALL / SAFE_KERNEL_CALL, include/common/cuda_error_hadling.h:4-24,
// which printed and threw.)
// ---------------------------------------------------------------------------
namespace sssp {

void set_error(const std::string &msg);
int fail(int code, const std::string &msg);

struct Error {
    int code;
    std::string msg;
};

#define SSSP_HIP(call)                                                              \
    do {                                                                            \
        hipError_t e_ = (call);                                                     \
        if (e_ != hipSuccess)                                                       \
            throw ::sssp::Error{e_ == hipErrorOutOfMemory ? SSSP_ERR_OOM : SSSP_ERR_HIP, \
                                std::string("HIP error '") + hipGetErrorString(e_) +  \
                                    "' at " #call " (" __FILE__ ":" + std::to_string(__LINE__) + ")"}; \
    } while (0)

#define SSSP_LAUNCH_CHECK(what)                                                     \
    do {                                                                            \
        hipError_t e_ = hipGetLastError();                                          \
        if (e_ != hipSuccess)                                                       \
            throw ::sssp::Error{SSSP_ERR_HIP, std::string("kernel launch failed: ") + \
                                                  hipGetErrorString(e_) + " at " what}; \
    } while (0)

#define SSSP_REQUIRE(cond, msg)                                                     \
    do {                                                                            \
        if (!(cond)) throw ::sssp::Error{SSSP_ERR_ARG, std::string(msg)};           \
    } while (0)

// Runs `body` and converts exceptions into a status code + message.
template <class F>
int guard(F &&body)
{
    try {
        body();
        return SSSP_OK;
    } catch (const Error &e) {
        return fail(e.code, e.msg);
    } catch (const std::bad_alloc &) {
        return fail(SSSP_ERR_OOM, "host allocation failed");
    } catch (...) {
        return fail(SSSP_ERR_INTERNAL, "unexpected exception");
    }
}

// Test hook: when non-zero, the next device allocation made through dmalloc
// fails with hipErrorOutOfMemory (exercised by the CPU-side error tests via
// sssp_debug_fail_next_alloc).
extern int g_fail_next_alloc;

template <class T>
T *dmalloc(size_t count)
{
    if (g_fail_next_alloc) {
        g_fail_next_alloc = 0;
        throw Error{SSSP_ERR_OOM, "device allocation failed (injected)"};
    }
    void *p = nullptr;
    size_t bytes = count * sizeof(T);
    if (bytes == 0) bytes = 256;
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess)
        throw Error{SSSP_ERR_OOM, std::string("hipMalloc of ") + std::to_string(bytes) +
                                      " bytes failed: " + hipGetErrorString(e)};
    return static_cast<T *>(p);
}

inline void dfree(void *p)
{
    if (p) (void)hipFree(p);
}

// ---------------------------------------------------------------------------
// Weight arithmetic. Distances live in HBM as 32-bit patterns for both weight
// types. For f32 every distance is a non-negative float (or FLT_MAX = "not
// reached", sssp.hpp:21 / bellman_ford.cu:36), and non-negative IEEE floats
// order exactly like their unsigned bit patterns, so one unsigned atomicMin
// serves both types.
// ---------------------------------------------------------------------------
struct WF32 {
    static constexpr uint32_t INF = 0x7F7FFFFFu;  // FLT_MAX
    __device__ __forceinline__ static uint32_t add(uint32_t d, uint32_t w)
    {
        // one IEEE fp32 add, round-to-nearest-even: the reference's
        // `src_distance + weight` (bellman_ford.cu:108)
        return __float_as_uint(__uint_as_float(d) + __uint_as_float(w));
    }
};

struct WU32 {
    static constexpr uint32_t INF = 0xFFFFFFFFu;
    __device__ __forceinline__ static uint32_t add(uint32_t d, uint32_t w)
    {
        uint32_t s = d + w;  // saturating: the INF guard of the integer path
        return s < d ? INF : s;
    }
};

// BFS levels: every edge is one hop and its weight is ignored; the fixpoint
// of d[v] <= d[u] + 1 is the hop distance, and d != INF is the reference's
// visited flag (bfs.cu:25-39). Callers never add to INF; levels stay < V.
struct WUnit {
    static constexpr uint32_t INF = 0xFFFFFFFFu;
    __device__ __forceinline__ static uint32_t add(uint32_t d, uint32_t) { return d + 1u; }
};

// 64-bit distances (f64 reference-shaped wrapper only).
struct WF64 {
    static constexpr unsigned long long INF = 0x7FEFFFFFFFFFFFFFull;  // DBL_MAX
    __device__ __forceinline__ static unsigned long long add(unsigned long long d,
                                                             unsigned long long w)
    {
        return (unsigned long long)__double_as_longlong(__longlong_as_double((long long)d) +
                                                        __longlong_as_double((long long)w));
    }
};

// ---------------------------------------------------------------------------
// Wavefront helpers (CDNA: 64 lanes).
// ---------------------------------------------------------------------------
__device__ __forceinline__ int lane_id() { return __lane_id(); }

__device__ __forceinline__ uint64_t lanemask_lt() { return (1ull << __lane_id()) - 1ull; }

// Wave-aggregated append of `item` (when `pred`) to queue `q` with counter
// `len`: one global atomic per wave, lanes write consecutive slots.
// Must be called by all lanes of the wave in converged control flow.
__device__ __forceinline__ void wave_append(bool pred, int32_t item, int32_t *q, int32_t *len)
{
    uint64_t m = __ballot(pred);
    if (m == 0) return;
    int leader = __ffsll((unsigned long long)m) - 1;
    int base = 0;
    if (__lane_id() == leader) base = atomicAdd(len, __popcll(m));
    base = __shfl(base, leader);
    if (pred) q[base + __popcll(m & lanemask_lt())] = item;
}

// Wave-local staging of queue appends in LDS: lanes write their items into the
// wave's own LDS buffer (no atomics), and the wave reserves space in the
// global queue with one atomicAdd per flush. `cnt` is wave-uniform.
__device__ __forceinline__ void stage_push(bool pred, int32_t item, int32_t *buf, int &cnt)
{
    const uint64_t m = __ballot(pred);
    if (pred) buf[cnt + __popcll(m & lanemask_lt())] = item;
    cnt += __popcll(m);
}

__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ void stage_flush(int32_t *buf, int &cnt, int32_t *q, int32_t *len)
{
    if (cnt == 0) return;
    wave_lds_sync();
    int base = 0;
    if (__lane_id() == 0) base = atomicAdd(len, cnt);
    base = __shfl(base, 0);
    for (int i = __lane_id(); i < cnt; i += 64) q[base + i] = buf[i];
    wave_lds_sync();
    cnt = 0;
}

// Wave sum of a 64-bit value, result valid in every lane.
__device__ __forceinline__ long long wave_sum64(long long x)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
    return x;
}

__device__ __forceinline__ int wave_sum32(int x)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
    return x;
}

__device__ __forceinline__ uint32_t wave_min_u32(uint32_t x)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        uint32_t y = __shfl_xor(x, o);
        x = y < x ? y : x;
    }
    return x;
}

// Block of 256 threads = 4 waves. XCD-aware tile order: consecutive hardware
// block ids land on different XCDs (round-robin over 8), so give each XCD a
// contiguous range of tiles to keep neighbouring tiles in one L2.
__device__ __forceinline__ int64_t xcd_tile(int64_t block, int64_t nblocks)
{
    const int64_t X = 8;
    if (nblocks % X != 0) return block;
    return (block % X) * (nblocks / X) + block / X;
}

}  // namespace sssp
