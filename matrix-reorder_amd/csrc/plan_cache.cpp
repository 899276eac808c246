// plan_cache.cpp -- persisted reorderings.  A pGOrder permutation, and the
// MERBIT options tuned for the reordered matrix, are stored on disk under the
// fingerprint of the matrix pattern, the width and the window, so that a later
// process (or the serving loop's next upload of the same matrix) relabels
// without running pGOrder again.
//
// File <dir>/<fingerprint:016x>-pgorder-w<width>-k<window>.mrplan, little
// endian: a PlanHeader, then perm[n_rows] (int32, perm[old] = new).  Files are
// written to a temporary name and renamed, so readers never see a partial
// file.  A file whose header, size, permutation fingerprint or permutation
// check does not match is rejected and replaced.
#include <hip/hip_runtime_api.h>

#include <chrono>
#include <cinttypes>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <unistd.h>
#include <vector>

#include "common.hpp"
#include "kernels/launch_cache.hpp"
#include "parallel.hpp"
#include "reorder.hpp"

namespace mr {
namespace {

using clk = std::chrono::steady_clock;

constexpr char kMagic[8] = {'M', 'R', 'P', 'L', 'A', 'N', '0', '1'};
constexpr int32_t kMethodPgorder = 1;

struct PlanHeader {
    char magic[8];
    uint64_t pattern_hash;    // fingerprint of the matrix before reordering
    int64_t n_rows, n_cols, nnz;
    int32_t method, width, window;
    int32_t has_options;      // 1: the MERBIT options below were saved
    int32_t policy, sigma, xs, xw, xc_req;
    int32_t reserved;
    uint64_t perm_hash;       // fingerprint of perm[]
};
static_assert(sizeof(PlanHeader) == 88, "plan header layout");

uint64_t combine(uint64_t n_rows, uint64_t n_cols, uint64_t nnz, uint64_t s_rp, uint64_t s_ci)
{
    uint64_t h = k::fp_mix64(0x4D52504C414E3031ull ^ n_rows);
    h = k::fp_mix64(h ^ n_cols);
    h = k::fp_mix64(h ^ nnz);
    h = k::fp_mix64(h ^ s_rp);
    return k::fp_mix64(h ^ s_ci);
}

uint64_t host_sum(const int32_t* a, int64_t n, uint64_t salt)
{
    if (n <= 0) return 0;
    std::vector<uint64_t> part((size_t)std::max(1, host_threads()), 0);
    parallel_chunks(n, [&](int t, int64_t b, int64_t e) {
        uint64_t acc = 0;
        for (int64_t i = b; i < e; ++i) acc += k::fingerprint_term((uint64_t)i, (uint32_t)a[i], salt);
        part[(size_t)t] += acc;
    });
    uint64_t s = 0;
    for (uint64_t v : part) s += v;
    return s;
}

int device_hash(mr_mat m, uint64_t* out)
{
    MR_HIP_TRY(hipSetDevice(m->ctx->device));
    hipStream_t s = m->ctx->stream;
    DevBuf acc;
    MR_TRY(acc.alloc(2 * sizeof(uint64_t)));
    MR_HIP_TRY(hipMemsetAsync(acc.p, 0, 2 * sizeof(uint64_t), s));
    uint64_t* d = acc.as<uint64_t>();
    MR_HIP_TRY(k::fingerprint_sum(m->rp.as<int32_t>(), m->n_rows + 1, k::kSaltRowPtr, d, m->ctx->cu_count, s));
    MR_HIP_TRY(k::fingerprint_sum(m->ci.as<int32_t>(), m->nnz, k::kSaltColIdx, d + 1, m->ctx->cu_count, s));
    uint64_t h[2];
    MR_HIP_TRY(hipMemcpyAsync(h, d, sizeof(h), hipMemcpyDeviceToHost, s));
    MR_HIP_TRY(hipStreamSynchronize(s));
    *out = combine((uint64_t)m->n_rows, (uint64_t)m->n_cols, (uint64_t)m->nnz, h[0], h[1]);
    return MR_OK;
}

std::string plan_path(const std::string& dir, uint64_t h, int width, int window)
{
    char name[96];
    std::snprintf(name, sizeof(name), "/%016" PRIx64 "-pgorder-w%d-k%d.mrplan", h, width, window);
    return dir + name;
}

// 1 = valid entry read, 0 = no file, 2 = a file that does not match
int read_plan(const std::string& path, mr_mat m, uint64_t h, int width, int window, PlanHeader& hd,
              std::vector<int32_t>& perm)
{
    FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) return 0;
    int state = 2;
    const int64_t n = m->n_rows;
    if (std::fread(&hd, sizeof(hd), 1, f) == 1 && std::memcmp(hd.magic, kMagic, 8) == 0 &&
        hd.pattern_hash == h && hd.n_rows == n && hd.n_cols == m->n_cols && hd.nnz == m->nnz &&
        hd.method == kMethodPgorder && hd.width == width && hd.window == window) {
        perm.resize((size_t)n);
        if (std::fread(perm.data(), sizeof(int32_t), (size_t)n, f) == (size_t)n && std::fgetc(f) == EOF &&
            host_sum(perm.data(), n, k::kSaltPerm) == hd.perm_hash && mr_check_perm(perm.data(), n) == MR_OK)
            state = 1;
    }
    std::fclose(f);
    clear_error();       // a rejected permutation is not the caller's error
    return state;
}

int write_plan(const std::string& path, const PlanHeader& hd, const std::vector<int32_t>& perm)
{
    char suffix[64];
    std::snprintf(suffix, sizeof(suffix), ".tmp.%ld.%zu", (long)getpid(),
                  std::hash<std::thread::id>{}(std::this_thread::get_id()));
    const std::string tmp = path + suffix;
    FILE* f = std::fopen(tmp.c_str(), "wb");
    if (!f) return set_error(MR_ERR_IO, "plan cache: cannot create %s", tmp.c_str());
    bool ok = std::fwrite(&hd, sizeof(hd), 1, f) == 1 &&
              std::fwrite(perm.data(), sizeof(int32_t), perm.size(), f) == perm.size();
    ok = (std::fclose(f) == 0) && ok;
    if (!ok || std::rename(tmp.c_str(), path.c_str()) != 0) {
        std::remove(tmp.c_str());
        return set_error(MR_ERR_IO, "plan cache: cannot write %s", path.c_str());
    }
    return MR_OK;
}

void fill_options(mr_mat m, PlanHeader& hd, bool with_options)
{
    hd.has_options = with_options ? 1 : 0;
    hd.policy = m->merbit.policy;
    hd.sigma = m->merbit.sigma;
    hd.xs = m->merbit.xs;
    hd.xw = m->merbit.xw;
    hd.xc_req = m->merbit.xc_req;
}

int apply_options(mr_mat m, const PlanHeader& hd)
{
    if (!hd.has_options) return MR_OK;
    MR_TRY(mr_merbit_options(m, hd.policy));
    MR_TRY(mr_merbit_configure(m, hd.sigma));
    if (hd.policy & 4) MR_TRY(mr_merbit_xc_configure(m, hd.xs, hd.xw, hd.xc_req));
    return MR_OK;
}

}  // namespace
}  // namespace mr

using namespace mr;

extern "C" {

int mr_csr_hash(int64_t n_rows, int64_t n_cols, int64_t nnz, const int32_t* row_ptr, const int32_t* col_idx,
                uint64_t* hash)
{
    if (!row_ptr || !hash || (nnz > 0 && !col_idx)) return set_error(MR_ERR_INVALID_ARG, "null argument");
    if (n_rows < 0 || n_cols < 0 || nnz < 0) return set_error(MR_ERR_INVALID_ARG, "negative size");
    *hash = combine((uint64_t)n_rows, (uint64_t)n_cols, (uint64_t)nnz,
                    host_sum(row_ptr, n_rows + 1, k::kSaltRowPtr), host_sum(col_idx, nnz, k::kSaltColIdx));
    return MR_OK;
}

int mr_mat_hash(mr_mat m, uint64_t* hash)
{
    if (!m || !hash) return set_error(MR_ERR_INVALID_ARG, "null argument");
    return device_hash(m, hash);
}

int mr_reorder_pgorder_cached(mr_mat m, int width, int window, const char* cache_dir, int32_t* perm_out,
                              mr_reorder_stats* stats, int* hit)
{
    if (!m) return set_error(MR_ERR_INVALID_ARG, "null matrix handle");
    if (!cache_dir || !*cache_dir) return set_error(MR_ERR_INVALID_ARG, "empty plan cache directory");
    if (m->n_rows != m->n_cols) return set_error(MR_ERR_UNSUPPORTED, "pGOrder needs a square matrix");
    const auto t0 = clk::now();
    if (width == 0) MR_TRY(mr_get_width(m->n_rows, m->nnz, &width));
    const int64_t n = m->n_rows;
    uint64_t h = 0;
    MR_TRY(device_hash(m, &h));
    const std::string dir(cache_dir);
    const std::string path = plan_path(dir, h, width, window);
    PlanHeader hd;
    std::vector<int32_t> perm;
    const int state = read_plan(path, m, h, width, window, hd, perm);
    int result;
    mr_reorder_stats st{};
    if (state == 1) {
        if (n > 0) MR_TRY(relabel(m, perm, &st.relabel_ms));
        MR_TRY(apply_options(m, hd));
        result = 1;
    } else {
        perm.resize((size_t)n);
        MR_TRY(mr_reorder_pgorder(m, width, window, perm.data(), &st));
        std::memset(&hd, 0, sizeof(hd));
        std::memcpy(hd.magic, kMagic, 8);
        hd.pattern_hash = h;
        hd.n_rows = n;
        hd.n_cols = m->n_cols;
        hd.nnz = m->nnz;
        hd.method = kMethodPgorder;
        hd.width = width;
        hd.window = window;
        fill_options(m, hd, false);
        hd.perm_hash = host_sum(perm.data(), n, k::kSaltPerm);
        if (write_plan(path, hd, perm) == MR_OK) {
            result = state == 2 ? 2 : 0;
        } else {
            clear_error();      // the reorder itself succeeded
            result = -1;
        }
    }
    m->plan_path = path;
    m->plan_state = host_sum(m->perm.data(), (int64_t)m->perm.size(), k::kSaltPerm);
    if (perm_out && n > 0) std::memcpy(perm_out, perm.data(), sizeof(int32_t) * (size_t)n);
    if (stats) {
        *stats = st;
        stats->total_ms = std::chrono::duration<double, std::milli>(clk::now() - t0).count();
    }
    if (hit) *hit = result;
    return MR_OK;
}

int mr_plan_save(mr_mat m)
{
    if (!m) return set_error(MR_ERR_INVALID_ARG, "null matrix handle");
    if (m->plan_path.empty())
        return set_error(MR_ERR_INVALID_ARG, "the matrix was not reordered through the plan cache");
    if (host_sum(m->perm.data(), (int64_t)m->perm.size(), k::kSaltPerm) != m->plan_state)
        return set_error(MR_ERR_INVALID_ARG, "the matrix was relabelled after its cached reorder");
    PlanHeader hd;
    std::vector<int32_t> perm;
    FILE* f = std::fopen(m->plan_path.c_str(), "rb");
    if (!f) return set_error(MR_ERR_IO, "plan cache: %s is gone", m->plan_path.c_str());
    const bool ok = std::fread(&hd, sizeof(hd), 1, f) == 1 && std::memcmp(hd.magic, kMagic, 8) == 0 &&
                    hd.n_rows == m->n_rows;
    if (ok) {
        perm.resize((size_t)hd.n_rows);
        if (std::fread(perm.data(), sizeof(int32_t), perm.size(), f) != perm.size()) perm.clear();
    }
    std::fclose(f);
    if (!ok || (int64_t)perm.size() != m->n_rows || host_sum(perm.data(), m->n_rows, k::kSaltPerm) != hd.perm_hash)
        return set_error(MR_ERR_IO, "plan cache: %s does not hold this matrix's plan", m->plan_path.c_str());
    fill_options(m, hd, true);
    return write_plan(m->plan_path, hd, perm);
}

}  // extern "C"
