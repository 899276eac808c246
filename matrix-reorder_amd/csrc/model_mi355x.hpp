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
constexpr double kMerbitB = 1.36764000769223737e-01;
constexpr double kMerbitW[10] = {
    -2.26696665280884113e-01,
    7.15538935963671522e-03,
    -2.32716085786676879e-01,
    -5.44776732387246276e-02,
    -3.38495412491703410e-02,
    4.35910635450482087e-01,
    1.66809806143071393e+00,
    1.05577629253658389e+00,
    -1.14244831952275305e+00,
    0.00000000000000000e+00};
constexpr double kCsrB = 1.39387591968371971e-01;
constexpr double kCsrW[10] = {
    -2.35667924735178530e-01,
    -1.15214917283845295e-01,
    -1.05927855596343642e-01,
    -4.61277525715407394e-02,
    -6.93526136654715464e-02,
    7.02141198338476991e-01,
    1.31886358531133285e+00,
    8.15877043726177087e-01,
    -5.14789086288595099e-01,
    0.00000000000000000e+00};

}  // namespace mi355x
}  // namespace mr
