// rd.cpp -- reuse-distance (RD) features and the locality predictor:
// exact sector RD histogram on the host (Fenwick tree), sampled HLL RD on the
// device, and the exported ridge models.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

#include "common.hpp"
#include "kernels/launch_rd.hpp"
#include "model_mi355x.hpp"

namespace mr {
namespace {

// ridge/out/merbit/sampled/*.json "ridge_raw_finite_features" (== include/model.cuh:74-87)
const double kMerbitB = 0.26649956401604796;
const double kMerbitW[10] = {-0.3362573716950983, -0.2557862191260704, -0.24248079087236038, -0.0603012191232085,
                             0.28053687310764397, -0.016399984212930544, -0.04634379957053418,
                             -0.30165791361557115, 2.754071221226108, 2.3042879761742348};
// ridge/out/csr/sampled/*.json "ridge_raw_finite_features"
const double kCsrB = 0.2565235315979766;
const double kCsrW[10] = {-0.3238047246104749, -0.3004866395371231, 0.00597989718749251, -0.06044308050866558,
                          0.45229148690047805, -0.1288758680132092, 0.2973684717736622, -0.45425853466807625,
                          1.7676276675955795, 2.4597865184371206};

int bin_of(double rd, const int32_t* edges, int32_t n_edges)
{
    return (int)(std::upper_bound(edges, edges + n_edges, rd,
                                  [](double v, int32_t e) { return v < (double)e; }) - edges);
}

// reference hllrd_estimate (src/class.cu:982-1012) from the serial register sum
float hll_estimate(float sum, int32_t zeros, int32_t m)
{
    const float fm = (float)m;
    float alpha;
    if (m == 16) alpha = 0.673f;
    else if (m == 32) alpha = 0.697f;
    else if (m == 64) alpha = 0.709f;
    else alpha = 0.7213f / (1.0f + 1.079f / fm);
    float est = alpha * fm * fm / sum;
    if (est <= 2.5f * fm && zeros > 0) est = fm * logf(fm / (float)zeros);
    return est;
}

int validate(const mr_rd_params* p)
{
    if (!p) return set_error(MR_ERR_INVALID_ARG, "null RD parameters");
    if (p->sector_size <= 0) return set_error(MR_ERR_INVALID_ARG, "sector_size must be positive");
    if (p->n_edges <= 0 || p->n_edges > 32) return set_error(MR_ERR_INVALID_ARG, "bin_edges must not be empty");
    for (int i = 1; i < p->n_edges; ++i)
        if (p->bin_edges[i] < p->bin_edges[i - 1]) return set_error(MR_ERR_INVALID_ARG, "bin_edges must be sorted");
    if (p->rd_cap > 0 && p->rd_cap < p->bin_edges[p->n_edges - 1])
        return set_error(MR_ERR_INVALID_ARG,
                         "rd_cap must be at least the last RD bin edge; otherwise status=2 cannot be "
                         "assigned to the final bin safely");
    return MR_OK;
}

}  // namespace
}  // namespace mr

using namespace mr;

extern "C" {

int mr_rd_params_default(mr_rd_params* p)
{
    if (!p) return set_error(MR_ERR_INVALID_ARG, "null parameters");
    std::memset(p, 0, sizeof(*p));
    // GRAPH::locality_predict (src/class.cu:1265-1277)
    p->sector_size = 32;
    const int32_t edges[9] = {1, 112, 512, 18432, 36864, 73728, 147456, 294912, 589824};
    p->n_edges = 9;
    std::memcpy(p->bin_edges, edges, sizeof(edges));
    p->samples = 30960;
    p->rd_cap = 589824;
    p->hll_p = 12;
    p->max_scan = 0;
    p->stream_type = 0;
    p->graph_id = 0;
    return MR_OK;
}

int mr_rd_exact_stream(const int32_t* stream, int64_t n, int64_t nnz, const mr_rd_params* p, int64_t* counts,
                       double* feat, mr_rd_stats* stats)
{
    MR_TRY(validate(p));
    if (nnz > 0 && !stream) return set_error(MR_ERR_INVALID_ARG, "null stream");
    const auto t0 = std::chrono::steady_clock::now();
    // compute_sector_rd_histogram (src/class.cu:895-936): sectors sized by n
    const int64_t nsec = (n + p->sector_size - 1) / p->sector_size;
    std::vector<int64_t> last((size_t)std::max<int64_t>(nsec, 1), -1);
    std::vector<int32_t> bit((size_t)nnz + 2, 0);
    const int nb = p->n_edges + 1;
    std::vector<int64_t> cnt((size_t)nb, 0);
    int64_t valid = 0, cold = 0;
    for (int64_t k2 = 0; k2 < nnz; ++k2) {
        const int64_t s = stream[k2] / p->sector_size;
        if (s < 0 || s >= nsec) return set_error(MR_ERR_INVALID_ARG, "stream value outside [0, n)");
        const int64_t old = last[(size_t)s];
        if (old < 0) {
            for (int64_t i = k2 + 1; i <= nnz; i += i & -i) bit[(size_t)i] += 1;
            last[(size_t)s] = k2;
            ++cold;
            continue;
        }
        for (int64_t i = old + 1; i <= nnz; i += i & -i) bit[(size_t)i] -= 1;
        int64_t rd = 0;
        for (int64_t i = k2; i > 0; i -= i & -i) rd += bit[(size_t)i];
        for (int64_t i = old + 1; i > 0; i -= i & -i) rd -= bit[(size_t)i];
        cnt[(size_t)bin_of((double)rd, p->bin_edges, p->n_edges)]++;
        ++valid;
        for (int64_t i = k2 + 1; i <= nnz; i += i & -i) bit[(size_t)i] += 1;
        last[(size_t)s] = k2;
    }
    for (int b = 0; b < nb; ++b) {
        if (counts) counts[b] = cnt[(size_t)b];
        if (feat) feat[b] = valid > 0 ? (double)cnt[(size_t)b] / (double)valid : 0.0;
    }
    if (stats) {
        std::memset(stats, 0, sizeof(*stats));
        stats->reuses = valid;
        stats->cold = cold;
        stats->ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
    return MR_OK;
}

int mr_rd_sampled_raw(mr_mat m, const mr_rd_params* p, float* est, int32_t* status, int32_t* pos)
{
    if (!m) return set_error(MR_ERR_INVALID_ARG, "null matrix handle");
    MR_TRY(validate(p));
    const int32_t S = p->samples;
    if (S <= 0 || m->nnz <= 1) return MR_OK;
    MR_HIP_TRY(hipSetDevice(m->ctx->device));
    int hp = p->hll_p <= 0 ? 12 : p->hll_p;
    hp = std::max(4, std::min(hp, 14));
    // 2^p int registers live in the workgroup's LDS next to a few scalars;
    // p = 14 (64 KB) does not fit the 64 KB static+dynamic budget
    if (hp > 13) return set_error(MR_ERR_UNSUPPORTED, "hll_p above 13 is not supported on this device");
    hipStream_t s = m->ctx->stream;
    DevBuf d_sum, d_zero, d_est, d_st, d_pos;
    MR_TRY(d_sum.alloc(sizeof(float) * (size_t)S));
    MR_TRY(d_zero.alloc(sizeof(int32_t) * (size_t)S));
    MR_TRY(d_est.alloc(sizeof(float) * (size_t)S));
    MR_TRY(d_st.alloc(sizeof(int32_t) * (size_t)S));
    MR_TRY(d_pos.alloc(sizeof(int32_t) * (size_t)S));
    const uint32_t seed = 246813579U + (uint32_t)p->graph_id * 1009U + (uint32_t)p->stream_type * 9176U;
    // the access stream is the column index array in CSR order (class.cu:1205-1206)
    MR_HIP_TRY(k::hll_sample(m->ci.as<int32_t>(), m->nnz, S, p->sector_size, seed, p->rd_cap, p->max_scan, hp,
                             d_sum.as<float>(), d_zero.as<int32_t>(), d_est.as<float>(), d_st.as<int32_t>(),
                             d_pos.as<int32_t>(), s));
    std::vector<float> hs((size_t)S), he((size_t)S);
    std::vector<int32_t> hz((size_t)S), hst((size_t)S), hpos((size_t)S);
    MR_HIP_TRY(hipMemcpyAsync(hs.data(), d_sum.p, sizeof(float) * (size_t)S, hipMemcpyDeviceToHost, s));
    MR_HIP_TRY(hipMemcpyAsync(hz.data(), d_zero.p, sizeof(int32_t) * (size_t)S, hipMemcpyDeviceToHost, s));
    MR_HIP_TRY(hipMemcpyAsync(he.data(), d_est.p, sizeof(float) * (size_t)S, hipMemcpyDeviceToHost, s));
    MR_HIP_TRY(hipMemcpyAsync(hst.data(), d_st.p, sizeof(int32_t) * (size_t)S, hipMemcpyDeviceToHost, s));
    MR_HIP_TRY(hipMemcpyAsync(hpos.data(), d_pos.p, sizeof(int32_t) * (size_t)S, hipMemcpyDeviceToHost, s));
    MR_HIP_TRY(hipStreamSynchronize(s));
    for (int32_t i = 0; i < S; ++i) {
        // the estimate is formed on the host from the serial register sum for
        // every status (the device value only drives the early-stop check)
        const float e = hll_estimate(hs[(size_t)i], hz[(size_t)i], 1 << hp);
        (void)he;
        if (est) est[i] = e;
        if (status) status[i] = hst[(size_t)i];
        if (pos) pos[i] = hpos[(size_t)i];
    }
    return MR_OK;
}

int mr_rd_features(mr_mat m, int mode, const mr_rd_params* p, double* feat, mr_rd_stats* stats)
{
    if (!m || !feat) return set_error(MR_ERR_INVALID_ARG, "null argument");
    MR_TRY(validate(p));
    const int nb = p->n_edges + 1;
    if (mode == MR_RD_EXACT) {
        std::vector<int32_t> stream((size_t)std::max<int64_t>(m->nnz, 1));
        MR_TRY(mr_mat_export(m, nullptr, stream.data(), nullptr));
        return mr_rd_exact_stream(stream.data(), m->n_rows, m->nnz, p, nullptr, feat, stats);
    }
    if (mode != MR_RD_SAMPLED) return set_error(MR_ERR_INVALID_ARG, "unknown RD mode %d", mode);
    const auto t0 = std::chrono::steady_clock::now();
    for (int b = 0; b < nb; ++b) feat[b] = 0.0;
    if (stats) std::memset(stats, 0, sizeof(*stats));
    if (p->samples <= 0 || m->nnz <= 1) return MR_OK;
    std::vector<float> est((size_t)p->samples);
    std::vector<int32_t> st((size_t)p->samples);
    MR_TRY(mr_rd_sampled_raw(m, p, est.data(), st.data(), nullptr));
    // host binning (src/class.cu:1213-1256): status 2 goes to the last bin;
    // normalise over found + large
    std::vector<int64_t> cnt((size_t)nb, 0);
    int32_t found = 0, cold = 0, large = 0, unresolved = 0;
    for (int32_t i = 0; i < p->samples; ++i) {
        switch (st[(size_t)i]) {
        case 0: ++found; cnt[(size_t)bin_of((double)est[(size_t)i], p->bin_edges, p->n_edges)]++; break;
        case 1: ++cold; break;
        case 2: ++large; cnt[(size_t)nb - 1]++; break;
        default: ++unresolved; break;
        }
    }
    const int32_t valid = found + large;
    if (valid > 0)
        for (int b = 0; b < nb; ++b) feat[b] = (double)cnt[(size_t)b] / (double)valid;
    if (stats) {
        stats->found = found;
        stats->cold_samples = cold;
        stats->large = large;
        stats->unresolved = unresolved;
        stats->ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
    return MR_OK;
}

int mr_rd_params_model(int model, mr_rd_params* p)
{
    MR_TRY(mr_rd_params_default(p));
    switch (model) {
    case MR_MODEL_MERBIT:
    case MR_MODEL_CSR: return MR_OK;
    case MR_MODEL_MERBIT_MI355X:
    case MR_MODEL_CSR_MI355X:
        p->n_edges = 9;
        std::memcpy(p->bin_edges, mi355x::kEdges, sizeof(mi355x::kEdges));
        p->rd_cap = mi355x::kRdCap;
        return MR_OK;
    default: return set_error(MR_ERR_INVALID_ARG, "unknown model %d", model);
    }
}

int mr_predict(const double* feat, int model, double* speedup)
{
    if (!feat || !speedup) return set_error(MR_ERR_INVALID_ARG, "null argument");
    const double* w;
    double v;
    switch (model) {
    case MR_MODEL_MERBIT: w = kMerbitW; v = kMerbitB; break;
    case MR_MODEL_CSR: w = kCsrW; v = kCsrB; break;
    case MR_MODEL_MERBIT_MI355X: w = mi355x::kMerbitW; v = mi355x::kMerbitB; break;
    case MR_MODEL_CSR_MI355X: w = mi355x::kCsrW; v = mi355x::kCsrB; break;
    default: return set_error(MR_ERR_INVALID_ARG, "unknown model");
    }
    for (int i = 0; i < 10; ++i)
        if (!std::isfinite(feat[i])) return set_error(MR_ERR_NUMERIC, "feature %d is not finite", i);
    for (int i = 0; i < 10; ++i) v += w[i] * feat[i];
    const double sp = std::exp(v);
    if (!std::isfinite(sp) || sp <= 0.0) return set_error(MR_ERR_NUMERIC, "Predictor returned an invalid speedup");
    *speedup = sp;
    return MR_OK;
}

int mr_locality_predict(mr_mat m, int model, double* speedup, double* feat_out, mr_rd_stats* stats)
{
    if (!m || !speedup) return set_error(MR_ERR_INVALID_ARG, "null argument");
    mr_rd_params p;
    MR_TRY(mr_rd_params_model(model, &p));
    double feat[10];
    MR_TRY(mr_rd_features(m, MR_RD_SAMPLED, &p, feat, stats));
    if (feat_out) std::memcpy(feat_out, feat, sizeof(feat));
    return mr_predict(feat, model, speedup);
}

}  // extern "C"
