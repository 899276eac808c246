// gen.cpp -- host-side input generators and the MatrixMarket loader.
//
// R-MAT (Chakrabarti et al. 2004): every edge draws `scale` quadrant choices
// with probabilities (a, b, c, d = 1 - a - b - c), most significant bit first.
// Randomness is counter based: edge e uses a SplitMix64 stream seeded by
// (seed, e), value i uses (seed, i), so the output does not depend on the
// number of host threads.  With scramble != 0 vertex ids pass through a seeded
// bijection of [0, 2^scale) (two odd multiplications and two xor-shifts modulo
// 2^scale), as Graph500 generators do, so the hubs are not all at low ids.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>

#include "common.hpp"
#include "parallel.hpp"

using namespace mr;

namespace {

inline uint64_t scramble_id(uint64_t v, int scale, uint64_t seed)
{
    const uint64_t mask = (scale >= 64) ? ~0ull : ((1ull << scale) - 1);
    uint64_t s = seed ^ 0x5851F42D4C957F2Dull;
    const uint64_t k0 = splitmix64_next(s);
    const uint64_t k1 = splitmix64_next(s) | 1ull;
    const uint64_t k2 = splitmix64_next(s) | 1ull;
    v = (v * k1 + k0) & mask;
    v ^= v >> (scale / 2 + 1);
    v = (v * k2) & mask;
    v ^= v >> (scale / 3 + 1);
    return v & mask;
}

}  // namespace

extern "C" {

int mr_gen_rmat_edges(int scale, int edge_factor, double a, double b, double c, uint64_t seed,
                      int scramble, int32_t* src, int32_t* dst)
{
    if (scale < 1 || scale > 30 || edge_factor < 1 || !src || !dst)
        return set_error(MR_ERR_INVALID_ARG, "bad R-MAT arguments");
    if (a < 0 || b < 0 || c < 0 || a + b + c > 1.0)
        return set_error(MR_ERR_INVALID_ARG, "R-MAT probabilities must be >= 0 with a+b+c <= 1");
    const int64_t m = (int64_t)edge_factor << scale;
    const double ab = a + b, abc = a + b + c;
    parallel_chunks(m, [&](int, int64_t lo, int64_t hi) {
        for (int64_t e = lo; e < hi; ++e) {
            uint64_t st = stream_state(seed, (uint64_t)e);
            uint64_t u = 0, v = 0;
            for (int l = 0; l < scale; ++l) {
                const double r = u01(splitmix64_next(st));
                const uint64_t rb = (r >= ab) ? 1u : 0u;          // c or d: lower half
                const uint64_t cb = ((r >= a && r < ab) || r >= abc) ? 1u : 0u;  // b or d
                u = (u << 1) | rb;
                v = (v << 1) | cb;
            }
            if (scramble) {
                u = scramble_id(u, scale, seed);
                v = scramble_id(v, scale, seed);
            }
            src[e] = (int32_t)u;
            dst[e] = (int32_t)v;
        }
    });
    return MR_OK;
}

int mr_coo_to_csr(int64_t n_rows, int64_t m, const int32_t* src, const int32_t* dst, int dedup,
                  int32_t* row_ptr, int32_t* col_idx, int64_t* nnz_out)
{
    if (n_rows < 0 || m < 0 || !row_ptr || (m > 0 && (!src || !dst || !col_idx)))
        return set_error(MR_ERR_INVALID_ARG, "bad arguments");
    for (int64_t e = 0; e < m; ++e)
        if (src[e] < 0 || src[e] >= n_rows)
            return set_error(MR_ERR_INVALID_ARG, "src[%lld]=%d out of range", (long long)e, src[e]);
    // counting sort by row
    std::vector<std::atomic<int32_t>> cnt((size_t)n_rows + 1);
    parallel_chunks(n_rows + 1, [&](int, int64_t lo, int64_t hi) {
        for (int64_t i = lo; i < hi; ++i) cnt[(size_t)i].store(0, std::memory_order_relaxed);
    });
    parallel_chunks(m, [&](int, int64_t lo, int64_t hi) {
        for (int64_t e = lo; e < hi; ++e) cnt[(size_t)src[e]].fetch_add(1, std::memory_order_relaxed);
    });
    std::vector<int64_t> start((size_t)n_rows + 1, 0);
    for (int64_t i = 0; i < n_rows; ++i) start[(size_t)i + 1] = start[(size_t)i] + cnt[(size_t)i].load();
    std::vector<int32_t> tmp((size_t)(m > 0 ? m : 1));
    parallel_chunks(n_rows, [&](int, int64_t lo, int64_t hi) {
        for (int64_t i = lo; i < hi; ++i) cnt[(size_t)i].store((int32_t)0, std::memory_order_relaxed);
    });
    parallel_chunks(m, [&](int, int64_t lo, int64_t hi) {
        for (int64_t e = lo; e < hi; ++e) {
            const int32_t r = src[e];
            const int32_t k = cnt[(size_t)r].fetch_add(1, std::memory_order_relaxed);
            tmp[(size_t)(start[(size_t)r] + k)] = dst[e];
        }
    });
    // sort + dedup each row, record unique counts
    std::vector<int32_t> ulen((size_t)n_rows + 1, 0);
    parallel_chunks(n_rows, [&](int, int64_t lo, int64_t hi) {
        for (int64_t i = lo; i < hi; ++i) {
            int32_t* b = tmp.data() + start[(size_t)i];
            int32_t* e = tmp.data() + start[(size_t)i + 1];
            std::sort(b, e);
            ulen[(size_t)i] = dedup ? (int32_t)(std::unique(b, e) - b) : (int32_t)(e - b);
        }
    });
    row_ptr[0] = 0;
    for (int64_t i = 0; i < n_rows; ++i) {
        const int64_t nxt = (int64_t)row_ptr[i] + ulen[(size_t)i];
        if (nxt > INT32_MAX) return set_error(MR_ERR_UNSUPPORTED, "nnz exceeds int32");
        row_ptr[i + 1] = (int32_t)nxt;
    }
    parallel_chunks(n_rows, [&](int, int64_t lo, int64_t hi) {
        for (int64_t i = lo; i < hi; ++i)
            std::memcpy(col_idx + row_ptr[i], tmp.data() + start[(size_t)i],
                        sizeof(int32_t) * (size_t)ulen[(size_t)i]);
    });
    if (nnz_out) *nnz_out = row_ptr[n_rows];
    return MR_OK;
}

int mr_gen_values(uint64_t seed, int kind, double lo, double hi, int64_t count, double* out)
{
    if (count < 0 || (count > 0 && !out) || (kind != 0 && kind != 1))
        return set_error(MR_ERR_INVALID_ARG, "bad arguments");
    if (kind == 1 && (lo <= 0 || hi < lo))
        return set_error(MR_ERR_INVALID_ARG, "log-uniform needs 0 < lo <= hi");
    const double llo = kind == 1 ? std::log(lo) : 0.0, lhi = kind == 1 ? std::log(hi) : 0.0;
    parallel_chunks(count, [&](int, int64_t b, int64_t e) {
        for (int64_t i = b; i < e; ++i) {
            uint64_t st = stream_state(seed, (uint64_t)i);
            const double u = u01(splitmix64_next(st));
            if (kind == 0) {
                out[i] = lo + (hi - lo) * u;
            } else {
                const double sgn = (splitmix64_next(st) >> 63) ? -1.0 : 1.0;
                out[i] = sgn * std::exp(llo + (lhi - llo) * u);
            }
        }
    });
    return MR_OK;
}

int mr_gen_stencil27(int nx, int ny, int nz, int32_t* row_ptr, int32_t* col_idx, int64_t* n_out,
                     int64_t* nnz_out)
{
    if (nx < 1 || ny < 1 || nz < 1) return set_error(MR_ERR_INVALID_ARG, "bad grid");
    const int64_t n = (int64_t)nx * ny * nz;
    const int64_t nnz = (int64_t)(3LL * nx - 2) * (3LL * ny - 2) * (3LL * nz - 2);
    if (n_out) *n_out = n;
    if (nnz_out) *nnz_out = nnz;
    if (nnz > INT32_MAX) return set_error(MR_ERR_UNSUPPORTED, "stencil nnz exceeds int32");
    if (!row_ptr || !col_idx) return MR_OK;
    auto deg = [&](int v, int nv) { return (v > 0) + 1 + (v < nv - 1); };
    // row pointer: degrees are separable
    row_ptr[0] = 0;
    for (int64_t i = 0; i < n; ++i) {
        const int x = (int)(i % nx), y = (int)((i / nx) % ny), z = (int)(i / ((int64_t)nx * ny));
        row_ptr[i + 1] = row_ptr[i] + deg(x, nx) * deg(y, ny) * deg(z, nz);
    }
    parallel_chunks(n, [&](int, int64_t b, int64_t e) {
        for (int64_t i = b; i < e; ++i) {
            const int x = (int)(i % nx), y = (int)((i / nx) % ny), z = (int)(i / ((int64_t)nx * ny));
            int64_t k = row_ptr[i];
            for (int dz = -1; dz <= 1; ++dz) {
                if (z + dz < 0 || z + dz >= nz) continue;
                for (int dy = -1; dy <= 1; ++dy) {
                    if (y + dy < 0 || y + dy >= ny) continue;
                    for (int dx = -1; dx <= 1; ++dx) {
                        if (x + dx < 0 || x + dx >= nx) continue;
                        col_idx[k++] = (int32_t)(i + dx + (int64_t)nx * (dy + (int64_t)ny * dz));
                    }
                }
            }
        }
    });
    return MR_OK;
}

int mr_gen_banded(int64_t n, int bandwidth, int nnz_per_row, uint64_t seed, int32_t* row_ptr,
                  int32_t* col_idx, int64_t* nnz_out)
{
    if (n < 1 || bandwidth < 0 || nnz_per_row < 1) return set_error(MR_ERR_INVALID_ARG, "bad arguments");
    auto row_len = [&](int64_t i) {
        const int64_t lo = std::max<int64_t>(0, i - bandwidth), hi = std::min<int64_t>(n - 1, i + bandwidth);
        return (int32_t)std::min<int64_t>(nnz_per_row, hi - lo + 1);
    };
    int64_t total = 0;
    for (int64_t i = 0; i < n; ++i) total += row_len(i);
    if (nnz_out) *nnz_out = total;
    if (total > INT32_MAX) return set_error(MR_ERR_UNSUPPORTED, "nnz exceeds int32");
    if (!row_ptr || !col_idx) return MR_OK;
    row_ptr[0] = 0;
    for (int64_t i = 0; i < n; ++i) row_ptr[i + 1] = row_ptr[i] + row_len(i);
    parallel_chunks(n, [&](int, int64_t b, int64_t e) {
        std::vector<int32_t> cols;
        for (int64_t i = b; i < e; ++i) {
            const int64_t lo = std::max<int64_t>(0, i - bandwidth), hi = std::min<int64_t>(n - 1, i + bandwidth);
            const int32_t want = row_len(i);
            cols.clear();
            cols.push_back((int32_t)i);
            uint64_t st = stream_state(seed, (uint64_t)i);
            if (want == hi - lo + 1) {
                cols.clear();
                for (int64_t c = lo; c <= hi; ++c) cols.push_back((int32_t)c);
            } else {
                while ((int32_t)cols.size() < want) {
                    const int64_t c = lo + (int64_t)(u01(splitmix64_next(st)) * (double)(hi - lo + 1));
                    const int32_t cc = (int32_t)std::min(c, hi);
                    if (std::find(cols.begin(), cols.end(), cc) == cols.end()) cols.push_back(cc);
                }
                std::sort(cols.begin(), cols.end());
            }
            std::memcpy(col_idx + row_ptr[i], cols.data(), sizeof(int32_t) * cols.size());
        }
    });
    return MR_OK;
}

// MatrixMarket loader with the reference's semantics (BASEGRAPH ctor,
// src/class.cu:5-73): the first line decides "symmetric" (every entry is
// doubled, diagonal included) and "basezero"; values are forced to 1 when
// force_ones != 0 (the reference always forces).  The multiplied matrix is the
// CSR whose rows are the file's column index (class.cu:180,192).
int mr_mtx_load(mr_ctx ctx, const char* path, int store_dtype, int force_ones, mr_mat* out)
{
    if (!ctx || !path || !out) return set_error(MR_ERR_INVALID_ARG, "null argument");
    std::ifstream f(path);
    if (!f) return set_error(MR_ERR_IO, "cannot open %s", path);
    std::string line;
    bool symmetric = false, zero_base = false, first = true, pattern = false;
    do {
        if (!std::getline(f, line)) return set_error(MR_ERR_IO, "%s: no size line", path);
        if (first) {
            symmetric = line.find("symmetric") != std::string::npos;
            zero_base = line.find("basezero") != std::string::npos;
            pattern = line.find("pattern") != std::string::npos;
            first = false;
        }
    } while (!line.empty() && line[0] == '%');
    int64_t R = 0, C = 0, Z = 0;
    {
        std::istringstream ss(line);
        if (!(ss >> R >> C >> Z)) return set_error(MR_ERR_IO, "%s: bad size line", path);
    }
    const int64_t m = symmetric ? 2 * Z : Z;
    std::vector<int32_t> src((size_t)m), dst((size_t)m);
    std::vector<double> val((size_t)m, 1.0);
    int64_t i = 0;
    while (i < Z && std::getline(f, line)) {
        if (line.empty() || line[0] == '%') continue;
        std::istringstream ss(line);
        int64_t r = 0, c = 0;
        double v = 1.0;
        if (!(ss >> r >> c)) return set_error(MR_ERR_IO, "%s: bad entry line %lld", path, (long long)i);
        if (!pattern) ss >> v;
        if (!zero_base) { --r; --c; }
        if (r < 0 || c < 0 || r >= R || c >= C) return set_error(MR_ERR_IO, "%s: entry out of range", path);
        // multiplied matrix row = file column (reference CSR over col_ind)
        src[(size_t)i] = (int32_t)c;
        dst[(size_t)i] = (int32_t)r;
        val[(size_t)i] = force_ones ? 1.0 : v;
        if (symmetric) {
            src[(size_t)(Z + i)] = (int32_t)r;
            dst[(size_t)(Z + i)] = (int32_t)c;
            val[(size_t)(Z + i)] = force_ones ? 1.0 : v;
        }
        ++i;
    }
    if (i != Z) return set_error(MR_ERR_IO, "%s: expected %lld entries, read %lld", path, (long long)Z, (long long)i);
    // stable row bucketing keeps file order inside a row, then sort by column
    // (stable, as the reference's merge sort keeps duplicate order)
    const int64_t n = std::max(R, C);
    std::vector<int32_t> rp((size_t)n + 1, 0);
    for (int64_t e = 0; e < m; ++e) rp[(size_t)src[(size_t)e] + 1]++;
    for (int64_t r = 0; r < n; ++r) rp[(size_t)r + 1] += rp[(size_t)r];
    std::vector<int32_t> fill(rp.begin(), rp.end() - 1);
    std::vector<std::pair<int32_t, double>> ent((size_t)m);
    for (int64_t e = 0; e < m; ++e) ent[(size_t)fill[(size_t)src[(size_t)e]]++] = {dst[(size_t)e], val[(size_t)e]};
    for (int64_t r = 0; r < n; ++r)
        std::stable_sort(ent.begin() + rp[(size_t)r], ent.begin() + rp[(size_t)r + 1],
                         [](const std::pair<int32_t, double>& a, const std::pair<int32_t, double>& b) {
                             return a.first < b.first;
                         });
    std::vector<int32_t> ci((size_t)m);
    for (int64_t e = 0; e < m; ++e) { ci[(size_t)e] = ent[(size_t)e].first; val[(size_t)e] = ent[(size_t)e].second; }
    return mr_csr_create(ctx, n, n, m, rp.data(), ci.data(), val.data(), MR_F64, store_dtype, out);
}

}  // extern "C"
