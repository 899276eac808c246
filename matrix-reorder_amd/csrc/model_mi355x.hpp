// model_mi355x.hpp -- RD headroom models fitted on MI355X timings.
// Written by mrpy/refit.py (fit) from mi355x_rd_train.csv; do not edit by hand.
// speedup = exp(b + sum_i w[i] * r[i]) over the ten sampled RD bin
// probabilities computed with kEdges; target t(order) / t(pGOrder order).
// Feature set: reference_edges.
#pragma once
#include <cstdint>

namespace mr {
namespace mi355x {

constexpr int32_t kEdges[9] = {1, 112, 512, 18432, 36864, 73728, 147456, 294912, 589824};
constexpr int32_t kRdCap = 589824;
constexpr double kMerbitB = 1.02709067864766593e-01;
constexpr double kMerbitW[10] = {
    -1.41880941063794092e-01,
    2.80477331437968026e-02,
    -1.91018779882737827e-01,
    -2.48961938036907464e-02,
    -5.35003155512363687e-02,
    4.53352948573334802e-01,
    1.58684085747984516e+00,
    -5.53706116564178674e-01,
    -4.36297817434215918e-01,
    0.00000000000000000e+00};
constexpr double kCsrB = 1.13100383976560873e-01;
constexpr double kCsrW[10] = {
    -1.54832190828243899e-01,
    -1.12784980234101356e-01,
    -7.77577934278313315e-02,
    -1.61796419048432123e-02,
    -5.39829590154392719e-02,
    7.22634863165507291e-01,
    1.37711454653633258e+00,
    -8.92118312339352515e-02,
    -7.68543171192074048e-01,
    0.00000000000000000e+00};

}  // namespace mi355x
}  // namespace mr
