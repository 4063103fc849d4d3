// mlp_dev.hpp -- the scheduling network's parameter layout, feature map and a
// single-row forward pass usable inside the per-round decision kernel.
#pragma once

#include "sssp_internal.hpp"

namespace sssp {

constexpr int kIn = SSSP_MLP_IN;
constexpr int kHid = SSSP_MLP_HID;
constexpr int kOut = SSSP_MLP_OUT;
constexpr int kModes = SSSP_MLP_MODES;
// parameter offsets (row-major weights: W[out][in])
constexpr int kOffW1 = 0;
constexpr int kOffB1 = kOffW1 + kHid * kIn;
constexpr int kOffW2 = kOffB1 + kHid;
constexpr int kOffB2 = kOffW2 + kHid * kHid;
constexpr int kOffW3 = kOffB2 + kHid;
constexpr int kOffB3 = kOffW3 + kOut * kHid;
constexpr int kNParam = kOffB3 + kOut;
static_assert(kNParam == SSSP_MLP_NPARAM, "parameter count mismatch");

// Scale-free features of one round: log-scaled frontier vertex and edge
// fractions, unsettled fraction, squashed round index and mean degree.
__host__ __device__ inline void make_features(double fv, double fe, double unsettled,
                                              double round, double mean_deg, double V,
                                              double E, float *out)
{
    auto lg = [](double x) {
        const double lo = 1e-9;
        double y = x < lo ? lo : (x > 1.0 ? 1.0 : x);
        return (float)(1.0 + log10(y) / 9.0);  // [1e-9, 1] -> [0, 1]
    };
    out[0] = lg(V > 0 ? fv / V : 0.0);
    out[1] = lg(E > 0 ? fe / E : 0.0);
    out[2] = lg(unsettled);
    out[3] = (float)(round / (round + 16.0));
    out[4] = (float)(mean_deg / (mean_deg + 16.0));
}

}  // namespace sssp
