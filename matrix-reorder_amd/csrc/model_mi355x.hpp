// model_mi355x.hpp -- RD headroom models fitted on MI355X timings.
// Written by mrpy/refit.py (fit) from mi355x_rd_train.csv; do not edit by hand.
// speedup = exp(b + sum_i w[i] * r[i]) over the ten sampled RD bin
// probabilities computed with kEdges; target t(order) / t(pGOrder order).
// Feature set: mi355x_edges.
#pragma once
#include <cstdint>

namespace mr {
namespace mi355x {

constexpr int32_t kEdges[9] = {1, 32, 112, 256, 1024, 8192, 32768, 131072, 262144};
constexpr int32_t kRdCap = 262144;
constexpr double kMerbitB = 1.73328011948709754e-01;
constexpr double kMerbitW[10] = {
    -3.02546653968843504e-01,
    4.29554919397135970e-01,
    -1.27908865587784082e-01,
    -8.37731448790604016e-01,
    2.39636422739685245e-01,
    -1.49872799344546836e-01,
    -6.02353580382591694e-02,
    9.55997949159451355e-01,
    4.50158192790066547e+00,
    0.00000000000000000e+00};
constexpr double kCsrB = 1.32175979119135034e-01;
constexpr double kCsrW[10] = {
    -1.54358492973378864e-01,
    -2.89846662499796629e-03,
    -1.48951882984405981e-01,
    -2.35116544816686479e-01,
    -1.43034179613825557e-01,
    -8.61775869231759528e-02,
    -4.08774935846230175e-02,
    7.93306515858193451e-01,
    1.06863912464851332e+01,
    0.00000000000000000e+00};

}  // namespace mi355x
}  // namespace mr
