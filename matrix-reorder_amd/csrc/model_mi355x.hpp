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
constexpr double kMerbitB = 1.63398337184711667e-01;
constexpr double kMerbitW[10] = {
    -3.51794074582339744e-01,
    5.31324813545729802e-01,
    -1.66389724261768374e-01,
    -7.98905254006095578e-01,
    2.80750310336807640e-01,
    -1.61563603057311711e-01,
    -6.25217327073647466e-02,
    9.76358860299390119e-01,
    -4.03639884710281227e-02,
    -6.99778603294693546e-01};
constexpr double kCsrB = 1.24099095973469756e-01;
constexpr double kCsrW[10] = {
    -2.03048208790278378e-01,
    9.88761711003832283e-02,
    -1.91238423157624932e-01,
    -2.01382571619076994e-01,
    -1.01575576950520288e-01,
    -9.74694858397428116e-02,
    -4.82297503596397087e-02,
    8.30507451848058409e-01,
    6.56543641088430385e-02,
    -6.38010089589619467e-01};

}  // namespace mi355x
}  // namespace mr
