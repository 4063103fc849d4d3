// mlp.hip -- the learned round scheduler: a 5 -> 64 -> 64 -> (3 + 4) fp32 MLP
// with ReLU, trained on the device by this file's own forward, backward and
// Adam kernels (cross-entropy on the fastest mode and the fastest chunk
// bucket of each timed round). Rows are independent, so every kernel maps one
// 64-lane wave to one row (lane = hidden unit); parameter gradients are
// reduced over the batch by one thread per parameter (deterministic, no
// float atomics).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>

#include "mlp_dev.hpp"
#include "sssp_graph.hpp"

struct sssp_mlp {
    int device = 0;
    hipStream_t stream = nullptr;
    float *P = nullptr;      // parameters
    float *G = nullptr;      // gradient
    float *M = nullptr;      // Adam first moment
    float *Vv = nullptr;     // Adam second moment
    int64_t step = 0;
    // batch workspace
    int32_t cap = 0;
    float *X = nullptr;      // [cap][kIn]
    int32_t *ym = nullptr, *yc = nullptr, *idx = nullptr;
    float *H1 = nullptr, *H2 = nullptr, *O = nullptr, *dO = nullptr, *dH2 = nullptr, *dH1 = nullptr;
    float *loss_rows = nullptr;
};

namespace sssp {

const float *mlp_device_params(sssp_mlp *m) { return m ? m->P : nullptr; }

namespace {

// forward for rows idx[r] (or r when idx == nullptr); one wave per row
__global__ __launch_bounds__(64) void k_fwd(const float *P, const float *X, const int32_t *idx,
                                            int32_t n, float *H1, float *H2, float *O)
{
    __shared__ float s_x[kIn], s_h1[kHid], s_h2[kHid];
    const int r = blockIdx.x, j = threadIdx.x;
    if (r >= n) return;
    const int32_t row = idx ? idx[r] : r;
    if (j < kIn) s_x[j] = X[(int64_t)row * kIn + j];
    __syncthreads();
    float acc = P[kOffB1 + j];
    for (int i = 0; i < kIn; i++) acc += P[kOffW1 + j * kIn + i] * s_x[i];
    acc = acc > 0.f ? acc : 0.f;
    s_h1[j] = acc;
    H1[(int64_t)r * kHid + j] = acc;
    __syncthreads();
    acc = P[kOffB2 + j];
    for (int i = 0; i < kHid; i++) acc += P[kOffW2 + j * kHid + i] * s_h1[i];
    acc = acc > 0.f ? acc : 0.f;
    s_h2[j] = acc;
    H2[(int64_t)r * kHid + j] = acc;
    __syncthreads();
    if (j < kOut) {
        acc = P[kOffB3 + j];
        for (int i = 0; i < kHid; i++) acc += P[kOffW3 + j * kHid + i] * s_h2[i];
        O[(int64_t)r * kOut + j] = acc;
    }
}

// softmax cross-entropy of the two heads; dO = (softmax - onehot) / n
__global__ __launch_bounds__(64) void k_out_grad(const float *O, const int32_t *ym,
                                                 const int32_t *yc, const int32_t *idx, int32_t n,
                                                 float *dO, float *loss_rows)
{
    const int r = blockIdx.x * 64 + threadIdx.x;
    if (r >= n) return;
    const int32_t row = idx ? idx[r] : r;
    const float *o = O + (int64_t)r * kOut;
    float *d = dO + (int64_t)r * kOut;
    float loss = 0.f;
    // head 1: modes [0, kModes), head 2: chunk buckets [kModes, kOut)
    const int lo[2] = {0, kModes}, hi[2] = {kModes, kOut};
    const int y[2] = {kModes > 0 ? ym[row] : 0, yc[row]};
    for (int h = 0; h < 2; h++) {
        float mx = o[lo[h]];
        for (int i = lo[h] + 1; i < hi[h]; i++) mx = fmaxf(mx, o[i]);
        float se = 0.f;
        for (int i = lo[h]; i < hi[h]; i++) se += expf(o[i] - mx);
        const float lse = mx + logf(se);
        loss += lse - o[lo[h] + y[h]];
        for (int i = lo[h]; i < hi[h]; i++) {
            const float p = expf(o[i] - lse);
            d[i] = (p - (i == lo[h] + y[h] ? 1.f : 0.f)) / (float)n;
        }
    }
    loss_rows[r] = loss;
}

__global__ __launch_bounds__(64) void k_back_h2(const float *P, const float *dO, const float *H2,
                                                int32_t n, float *dH2)
{
    const int r = blockIdx.x, j = threadIdx.x;
    if (r >= n) return;
    float acc = 0.f;
    for (int o = 0; o < kOut; o++) acc += dO[(int64_t)r * kOut + o] * P[kOffW3 + o * kHid + j];
    dH2[(int64_t)r * kHid + j] = H2[(int64_t)r * kHid + j] > 0.f ? acc : 0.f;
}

__global__ __launch_bounds__(64) void k_back_h1(const float *P, const float *dH2, const float *H1,
                                                int32_t n, float *dH1)
{
    const int r = blockIdx.x, j = threadIdx.x;
    if (r >= n) return;
    float acc = 0.f;
    for (int k = 0; k < kHid; k++) acc += dH2[(int64_t)r * kHid + k] * P[kOffW2 + k * kHid + j];
    dH1[(int64_t)r * kHid + j] = H1[(int64_t)r * kHid + j] > 0.f ? acc : 0.f;
}

// one thread per parameter: gradient = sum over the batch rows
__global__ __launch_bounds__(256) void k_param_grad(const float *X, const int32_t *idx, int32_t n,
                                                    const float *H1, const float *H2,
                                                    const float *dO, const float *dH2,
                                                    const float *dH1, float *G)
{
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p >= kNParam) return;
    float acc = 0.f;
    if (p < kOffB1) {  // W1[j][i]
        const int j = (p - kOffW1) / kIn, i = (p - kOffW1) % kIn;
        for (int r = 0; r < n; r++) {
            const int32_t row = idx ? idx[r] : r;
            acc += dH1[(int64_t)r * kHid + j] * X[(int64_t)row * kIn + i];
        }
    } else if (p < kOffW2) {  // b1[j]
        const int j = p - kOffB1;
        for (int r = 0; r < n; r++) acc += dH1[(int64_t)r * kHid + j];
    } else if (p < kOffB2) {  // W2[j][i]
        const int j = (p - kOffW2) / kHid, i = (p - kOffW2) % kHid;
        for (int r = 0; r < n; r++) acc += dH2[(int64_t)r * kHid + j] * H1[(int64_t)r * kHid + i];
    } else if (p < kOffW3) {  // b2[j]
        const int j = p - kOffB2;
        for (int r = 0; r < n; r++) acc += dH2[(int64_t)r * kHid + j];
    } else if (p < kOffB3) {  // W3[o][i]
        const int o = (p - kOffW3) / kHid, i = (p - kOffW3) % kHid;
        for (int r = 0; r < n; r++) acc += dO[(int64_t)r * kOut + o] * H2[(int64_t)r * kHid + i];
    } else {  // b3[o]
        const int o = p - kOffB3;
        for (int r = 0; r < n; r++) acc += dO[(int64_t)r * kOut + o];
    }
    G[p] = acc;
}

__global__ __launch_bounds__(256) void k_adam(float *P, const float *G, float *M, float *Vv,
                                              float lr, float b1, float b2, float eps, float c1,
                                              float c2)
{
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p >= kNParam) return;
    const float g = G[p];
    const float m = b1 * M[p] + (1.f - b1) * g;
    const float v = b2 * Vv[p] + (1.f - b2) * g * g;
    M[p] = m;
    Vv[p] = v;
    P[p] -= lr * (m / c1) / (sqrtf(v / c2) + eps);
}

void ensure_cap(sssp_mlp *m, int32_t n)
{
    if (n <= m->cap) return;
    void *ptrs[] = {m->X, m->ym, m->yc, m->idx, m->H1, m->H2, m->O, m->dO, m->dH2, m->dH1, m->loss_rows};
    for (void *p : ptrs) dfree(p);
    m->cap = n;
    m->X = dmalloc<float>((size_t)n * kIn);
    m->ym = dmalloc<int32_t>((size_t)n);
    m->yc = dmalloc<int32_t>((size_t)n);
    m->idx = dmalloc<int32_t>((size_t)n);
    m->H1 = dmalloc<float>((size_t)n * kHid);
    m->H2 = dmalloc<float>((size_t)n * kHid);
    m->O = dmalloc<float>((size_t)n * kOut);
    m->dO = dmalloc<float>((size_t)n * kOut);
    m->dH2 = dmalloc<float>((size_t)n * kHid);
    m->dH1 = dmalloc<float>((size_t)n * kHid);
    m->loss_rows = dmalloc<float>((size_t)n);
}

// loss + gradient for rows idx[0..n) of the resident data set (idx may be null)
void loss_grad_dev(sssp_mlp *m, const int32_t *idx, int32_t n)
{
    hipStream_t s = m->stream;
    k_fwd<<<n, 64, 0, s>>>(m->P, m->X, idx, n, m->H1, m->H2, m->O);
    k_out_grad<<<(n + 63) / 64, 64, 0, s>>>(m->O, m->ym, m->yc, idx, n, m->dO, m->loss_rows);
    k_back_h2<<<n, 64, 0, s>>>(m->P, m->dO, m->H2, n, m->dH2);
    k_back_h1<<<n, 64, 0, s>>>(m->P, m->dH2, m->H1, n, m->dH1);
    k_param_grad<<<(kNParam + 255) / 256, 256, 0, s>>>(m->X, idx, n, m->H1, m->H2, m->dO, m->dH2,
                                                       m->dH1, m->G);
    SSSP_LAUNCH_CHECK("mlp loss_grad");
}

void upload_set(sssp_mlp *m, const float *x, const int32_t *y_mode, const int32_t *y_chunk, int32_t n)
{
    ensure_cap(m, n);
    hipStream_t s = m->stream;
    SSSP_HIP(hipMemcpyAsync(m->X, x, sizeof(float) * kIn * (size_t)n, hipMemcpyHostToDevice, s));
    if (y_mode)
        SSSP_HIP(hipMemcpyAsync(m->ym, y_mode, sizeof(int32_t) * (size_t)n, hipMemcpyHostToDevice, s));
    if (y_chunk)
        SSSP_HIP(hipMemcpyAsync(m->yc, y_chunk, sizeof(int32_t) * (size_t)n, hipMemcpyHostToDevice, s));
}

}  // namespace
}  // namespace sssp

using namespace sssp;

extern "C" {

int sssp_mlp_create(int32_t device, uint64_t seed, sssp_mlp **out)
{
    sssp_mlp *m = nullptr;
    int rc = guard([&] {
        SSSP_REQUIRE(out, "null output handle");
        *out = nullptr;
        int n = 0;
        SSSP_HIP(hipGetDeviceCount(&n));
        SSSP_REQUIRE(device >= 0 && device < n, "device index out of range");
        SSSP_HIP(hipSetDevice(device));
        m = new sssp_mlp();
        m->device = device;
        SSSP_HIP(hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking));
        m->P = dmalloc<float>(kNParam);
        m->G = dmalloc<float>(kNParam);
        m->M = dmalloc<float>(kNParam);
        m->Vv = dmalloc<float>(kNParam);
        // He-uniform init from a seeded host generator, zero biases
        std::vector<float> p((size_t)kNParam, 0.f);
        std::mt19937_64 rng(seed);
        std::uniform_real_distribution<float> u(-1.f, 1.f);
        auto fill = [&](int off, int fan_in, int count) {
            const float lim = std::sqrt(6.f / (float)fan_in);
            for (int i = 0; i < count; i++) p[(size_t)off + i] = u(rng) * lim;
        };
        fill(kOffW1, kIn, kHid * kIn);
        fill(kOffW2, kHid, kHid * kHid);
        fill(kOffW3, kHid, kOut * kHid);
        SSSP_HIP(hipMemcpy(m->P, p.data(), sizeof(float) * kNParam, hipMemcpyHostToDevice));
        SSSP_HIP(hipMemset(m->M, 0, sizeof(float) * kNParam));
        SSSP_HIP(hipMemset(m->Vv, 0, sizeof(float) * kNParam));
        *out = m;
    });
    if (rc != SSSP_OK && m) sssp_mlp_destroy(m);
    return rc;
}

int sssp_mlp_destroy(sssp_mlp *m)
{
    return guard([&] {
        SSSP_REQUIRE(m, "null mlp");
        (void)hipSetDevice(m->device);
        if (m->stream) (void)hipStreamSynchronize(m->stream);
        void *ptrs[] = {m->P, m->G, m->M, m->Vv, m->X, m->ym, m->yc, m->idx, m->H1, m->H2,
                        m->O, m->dO, m->dH2, m->dH1, m->loss_rows};
        for (void *p : ptrs) dfree(p);
        if (m->stream) (void)hipStreamDestroy(m->stream);
        delete m;
    });
}

int sssp_mlp_get_params(const sssp_mlp *m, float *params)
{
    return guard([&] {
        SSSP_REQUIRE(m && params, "null argument");
        SSSP_HIP(hipSetDevice(m->device));
        SSSP_HIP(hipMemcpy(params, m->P, sizeof(float) * kNParam, hipMemcpyDeviceToHost));
    });
}

int sssp_mlp_set_params(sssp_mlp *m, const float *params)
{
    return guard([&] {
        SSSP_REQUIRE(m && params, "null argument");
        SSSP_HIP(hipSetDevice(m->device));
        for (int i = 0; i < kNParam; i++)
            SSSP_REQUIRE(std::isfinite(params[i]), "parameters must be finite");
        SSSP_HIP(hipMemcpy(m->P, params, sizeof(float) * kNParam, hipMemcpyHostToDevice));
        SSSP_HIP(hipMemset(m->M, 0, sizeof(float) * kNParam));
        SSSP_HIP(hipMemset(m->Vv, 0, sizeof(float) * kNParam));
        m->step = 0;
    });
}

int sssp_mlp_forward(sssp_mlp *m, const float *x, int32_t n, float *logits)
{
    return guard([&] {
        SSSP_REQUIRE(m && x && logits && n >= 1, "bad arguments");
        SSSP_HIP(hipSetDevice(m->device));
        upload_set(m, x, nullptr, nullptr, n);
        k_fwd<<<n, 64, 0, m->stream>>>(m->P, m->X, nullptr, n, m->H1, m->H2, m->O);
        SSSP_LAUNCH_CHECK("k_fwd");
        SSSP_HIP(hipMemcpyAsync(logits, m->O, sizeof(float) * kOut * (size_t)n,
                                hipMemcpyDeviceToHost, m->stream));
        SSSP_HIP(hipStreamSynchronize(m->stream));
    });
}

int sssp_mlp_loss_grad(sssp_mlp *m, const float *x, const int32_t *y_mode, const int32_t *y_chunk,
                       int32_t n, float *loss, float *grad)
{
    return guard([&] {
        SSSP_REQUIRE(m && x && y_mode && y_chunk && n >= 1, "bad arguments");
        for (int32_t i = 0; i < n; i++) {
            SSSP_REQUIRE(y_mode[i] >= 0 && y_mode[i] < kModes, "mode label out of range");
            SSSP_REQUIRE(y_chunk[i] >= 0 && y_chunk[i] < SSSP_MLP_CHUNKS, "chunk label out of range");
        }
        SSSP_HIP(hipSetDevice(m->device));
        upload_set(m, x, y_mode, y_chunk, n);
        loss_grad_dev(m, nullptr, n);
        std::vector<float> lr((size_t)n);
        SSSP_HIP(hipMemcpyAsync(lr.data(), m->loss_rows, sizeof(float) * (size_t)n,
                                hipMemcpyDeviceToHost, m->stream));
        if (grad)
            SSSP_HIP(hipMemcpyAsync(grad, m->G, sizeof(float) * kNParam, hipMemcpyDeviceToHost,
                                    m->stream));
        SSSP_HIP(hipStreamSynchronize(m->stream));
        double sum = 0;
        for (float v : lr) sum += v;
        if (loss) *loss = (float)(sum / n);
    });
}

int sssp_mlp_train(sssp_mlp *m, const float *x, const int32_t *y_mode, const int32_t *y_chunk,
                   int32_t n, int32_t epochs, int32_t batch, float lr, uint64_t seed, float *losses)
{
    return guard([&] {
        SSSP_REQUIRE(m && x && y_mode && y_chunk && n >= 1 && epochs >= 0 && batch >= 1,
                     "bad arguments");
        SSSP_REQUIRE(lr > 0.f && std::isfinite(lr), "learning rate must be positive");
        for (int32_t i = 0; i < n; i++) {
            SSSP_REQUIRE(y_mode[i] >= 0 && y_mode[i] < kModes, "mode label out of range");
            SSSP_REQUIRE(y_chunk[i] >= 0 && y_chunk[i] < SSSP_MLP_CHUNKS, "chunk label out of range");
        }
        SSSP_HIP(hipSetDevice(m->device));
        upload_set(m, x, y_mode, y_chunk, n);
        const int32_t B = batch < n ? batch : n;
        std::vector<int32_t> perm((size_t)n);
        for (int32_t i = 0; i < n; i++) perm[(size_t)i] = i;
        std::mt19937_64 rng(seed);
        int32_t *h_idx = nullptr;
        SSSP_HIP(hipHostMalloc((void **)&h_idx, sizeof(int32_t) * (size_t)n, hipHostMallocDefault));
        struct FreeHost {
            void *p;
            ~FreeHost() { (void)hipHostFree(p); }
        } fh{h_idx};
        std::vector<float> lrow((size_t)B);
        const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
        for (int32_t ep = 0; ep < epochs; ep++) {
            std::shuffle(perm.begin(), perm.end(), rng);
            std::memcpy(h_idx, perm.data(), sizeof(int32_t) * (size_t)n);
            SSSP_HIP(hipMemcpyAsync(m->idx, h_idx, sizeof(int32_t) * (size_t)n, hipMemcpyHostToDevice,
                                    m->stream));
            double ep_loss = 0;
            for (int32_t b0 = 0; b0 + B <= n; b0 += B) {
                loss_grad_dev(m, m->idx + b0, B);
                m->step++;
                const float c1 = 1.f - std::pow(b1, (float)m->step);
                const float c2 = 1.f - std::pow(b2, (float)m->step);
                k_adam<<<(kNParam + 255) / 256, 256, 0, m->stream>>>(m->P, m->G, m->M, m->Vv, lr, b1,
                                                                     b2, eps, c1, c2);
                SSSP_LAUNCH_CHECK("k_adam");
                if (losses) {
                    SSSP_HIP(hipMemcpyAsync(lrow.data(), m->loss_rows, sizeof(float) * (size_t)B,
                                            hipMemcpyDeviceToHost, m->stream));
                    SSSP_HIP(hipStreamSynchronize(m->stream));
                    for (float v : lrow) ep_loss += v;
                }
            }
            if (losses) losses[ep] = (float)(ep_loss / ((n / B) * (double)B));
        }
        SSSP_HIP(hipStreamSynchronize(m->stream));
    });
}

int sssp_graph_set_mlp(sssp_graph *g, sssp_mlp *m)
{
    return guard([&] {
        SSSP_REQUIRE(g, "null graph");
        SSSP_REQUIRE(!m || m->device == g->device, "mlp and graph live on different devices");
        g->mlp = m;
    });
}

int sssp_graph_set_bfs_mlp(sssp_graph *g, sssp_mlp *m)
{
    return guard([&] {
        SSSP_REQUIRE(g, "null graph");
        SSSP_REQUIRE(!m || m->device == g->device, "mlp and graph live on different devices");
        g->bfs_mlp = m;
    });
}

}  // extern "C"
