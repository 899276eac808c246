/*
 * mr_graph.hpp -- C++ mirror of the reference's operator surface over the
 * C-ABI of mr.h: ENV (include/class.cuh:13-36), the loader of BASEGRAPH /
 * GRAPH (src/class.cu:5-216, 330-421) and the GRAPH<V> methods on the SpMV
 * locality path (include/class.cuh:59-99), with the reference's names,
 * argument meaning, log records and error behaviour:
 *
 *  - a failing library call prints "file:line: message" to stderr and the
 *    method returns, as the reference's CUDA error macros print and continue
 *    (include/inc.cuh:101-117);
 *  - locality_predict throws std::runtime_error for an invalid prediction,
 *    as GRAPH::locality_predict does (src/class.cu:1317-1320);
 *  - each method appends the fields the reference writes to its log stream
 *    (FSTR): "&degree&&t1&t2", "&gorder5&vsize&t1&t2", "&t1&t2&speedup", ...
 *
 * Differences kept on purpose: the device id is honoured (the reference
 * always selects device 1, class.cuh:23); the ground truth that spmv checks
 * against is computed at load with the CSR vector kernel instead of
 * cuSPARSE; relabel_random (thrust::shuffle) and the gorder2-4 / slashburn
 * experiments are not part of this path.
 *
 * Header-only, C++17; link with libmr.
 */
#pragma once

#include <chrono>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <iostream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "mr.h"

namespace mr {

namespace detail {

inline bool ok(int status, const char* file, int line)
{
    if (status == MR_OK) return true;
    std::fprintf(stderr, "%s:%d: %s\n", file, line, mr_last_error());
    return false;
}

class Timer {
public:
    void Start() { t0_ = std::chrono::steady_clock::now(); }
    void Stop() { t1_ = std::chrono::steady_clock::now(); }
    float Millisecs() const { return std::chrono::duration<float, std::milli>(t1_ - t0_).count(); }

private:
    std::chrono::steady_clock::time_point t0_, t1_;
};

}  // namespace detail

#define MR_GRAPH_OK(expr) ::mr::detail::ok((expr), __FILE__, __LINE__)

// ENV (include/class.cuh:13-36): the device, the config file of matrix ids and
// paths, and the log file the methods append their records to
struct ENV {
    const int device_id = 0;
    const std::string config;
    const std::string log;
    mr_ctx ctx = nullptr;
    mr_stream stream = nullptr;
    mutable std::ofstream FSTR;

    ENV(const int id, const std::string& config_path, const std::string log_path)
        : device_id(id), config(config_path), log(log_path)
    {
        if (!MR_GRAPH_OK(mr_ctx_create(id, &ctx))) throw std::runtime_error("no device context");
        MR_GRAPH_OK(mr_ctx_stream(ctx, &stream));
        FSTR.open(log, std::ios_base::app);
    }
    ~ENV()
    {
        FSTR.close();
        mr_ctx_destroy(ctx);
    }
    ENV(const ENV&) = delete;
    ENV& operator=(const ENV&) = delete;
};

// the path of matrix `gid` in a config file: lines "id path [rows cols nnz]"
// (BASEGRAPH ctor, src/class.cu:18-25)
inline std::string config_path(const std::string& config, const int gid)
{
    std::ifstream file(config);
    std::string line;
    while (std::getline(file, line)) {
        int id = -1;
        std::string path;
        std::stringstream(line) >> id >> path;
        if (id == gid) return path;
    }
    return std::string();
}

template <typename V>
struct GRAPH {
    static_assert(std::is_same<V, float>::value || std::is_same<V, double>::value, "GRAPH<float> or GRAPH<double>");
    static constexpr int kDtype = std::is_same<V, double>::value ? MR_F64 : MR_F32;

    int graph_id = -1;
    int nrow = 0;
    int nnz = 0;
    mr_mat mat = nullptr;
    std::vector<V> y;          // A x with x = 1 on the loaded order (the check of spmv)

    GRAPH() {}
    // GRAPH(env, gid): matrix `gid` of env.config, a MatrixMarket file read as
    // the reference does (multiplied transpose, symmetric doubling, values 1)
    GRAPH(const ENV& _env, const int _gid) : graph_id(_gid)
    {
        const std::string path = config_path(_env.config, _gid);
        if (path.empty()) {
            std::fprintf(stderr, "graph id %d not found in %s\n", _gid, _env.config.c_str());
            return;
        }
        load(_env, path);
    }
    // a MatrixMarket file directly
    GRAPH(const ENV& _env, const std::string& _path, const int _gid = -1) : graph_id(_gid) { load(_env, _path); }
    ~GRAPH() { mr_mat_destroy(mat); }
    GRAPH(const GRAPH&) = delete;
    GRAPH& operator=(const GRAPH&) = delete;

    int get_width()
    {
        int w = 0;
        MR_GRAPH_OK(mr_get_width(nrow, nnz, &w));
        return w;
    }

    // reorderings: each relabels the matrix in place and logs its record
    void relabel_degree() { plain("degree", mr_reorder_degree(mat, nullptr, &st_), st_); }
    void dbg(const ENV& _env) { (void)_env; plain("dbg", mr_reorder_dbg(mat, nullptr, &st_), st_); }
    void relabel_rcm() { plain("rcm", mr_reorder_rcm(mat, nullptr, &st_), st_); }
    void rabbitorder() { plain("rabbitorder", mr_reorder_rabbit(mat, 1, nullptr, &st_), st_); }
    void gorder1(const ENV& _env, const int _width)
    {
        (void)_env;
        plain("gorder1", mr_reorder_gorder_cpu(mat, _width, nullptr, &st_), st_);
    }
    void gorder5(const ENV& _env, const int _vsize, const int _width1, const int _width2)
    {
        if (!MR_GRAPH_OK(mr_reorder_gorder5(mat, _vsize, _width1, _width2, nullptr, &st_))) return;
        log(_env) << "&gorder5&" << _vsize << "&" << reorder_ms(st_) << "&" << st_.relabel_ms;
    }
    void gorder6(const ENV& _env, const int _width)
    {
        if (!MR_GRAPH_OK(mr_reorder_pgorder(mat, _width, 5, nullptr, &st_))) return;
        log(_env) << "&" << reorder_ms(st_) << "&" << st_.relabel_ms;
    }
    void order(const ENV& _env, const int _width1, const int _width2)
    {
        if (!MR_GRAPH_OK(mr_reorder_tile_order(mat, _width1, _width2, nullptr, &st_))) return;
        log(_env) << "&order&" << _width1 << "&" << _width2 << "&" << reorder_ms(st_) << "&" << st_.relabel_ms;
    }

    // GRAPH::locality_predict (src/class.cu:1263-1326)
    void locality_predict(const ENV& _env, float& speedup)
    {
        double sp = 0.0;
        mr_rd_stats st{};
        detail::Timer timer;
        timer.Start();
        const int rc = mr_locality_predict(mat, MR_MODEL_MERBIT, &sp, nullptr, &st);
        timer.Stop();
        if (rc != MR_OK || !std::isfinite(sp) || sp <= 0.0)
            throw std::runtime_error(std::string("Predictor returned an invalid speedup: ") + mr_last_error());
        speedup = (float)sp;
        log(_env) << "&" << st.ms << "&" << (timer.Millisecs() - (float)st.ms) << "&" << speedup;
    }

    // GRAPH::spmv (src/class.cu:1331-1454): x = 1, 400 products, y back to the
    // loaded order, element check against the load-time product
    void spmv(const ENV& _env, std::string _method)
    {
        const int kernel = _method == "csr" ? MR_KERNEL_CSR_VECTOR : MR_KERNEL_MERBIT;
        if (_method != "csr" && _method != "merbit") {
            std::fprintf(stderr, "spmv: unknown method %s\n", _method.c_str());
            return;
        }
        void *x = nullptr, *yd = nullptr, *z = nullptr;
        const size_t bytes = sizeof(V) * (size_t)(nrow > 0 ? nrow : 1);
        if (!MR_GRAPH_OK(mr_malloc(_env.ctx, bytes, &x)) || !MR_GRAPH_OK(mr_malloc(_env.ctx, bytes, &yd)) ||
            !MR_GRAPH_OK(mr_malloc(_env.ctx, bytes, &z))) {
            mr_free(_env.ctx, x);
            mr_free(_env.ctx, yd);
            return;
        }
        std::vector<V> ones((size_t)nrow, (V)1);
        MR_GRAPH_OK(mr_memcpy(_env.ctx, x, ones.data(), bytes, 0));
        detail::Timer timer;
        timer.Start();
        MR_GRAPH_OK(mr_spmv(mat, kernel, x, yd, _env.stream));   // first call builds the kernel's metadata
        MR_GRAPH_OK(mr_stream_sync(_env.ctx, _env.stream));
        timer.Stop();
        const float tim1 = timer.Millisecs();
        timer.Start();
        for (int rnd = 0; rnd < 400; ++rnd) MR_GRAPH_OK(mr_spmv(mat, kernel, x, yd, _env.stream));
        MR_GRAPH_OK(mr_stream_sync(_env.ctx, _env.stream));
        timer.Stop();
        log(_env) << "&" << tim1 << "&" << timer.Millisecs() / 400 << std::endl;
        // y in the loaded order: z[i] = y[perm[i]]
        MR_GRAPH_OK(mr_permute_vec(mat, 1, yd, z, _env.stream));
        MR_GRAPH_OK(mr_stream_sync(_env.ctx, _env.stream));
        std::vector<V> h((size_t)nrow);
        MR_GRAPH_OK(mr_memcpy(_env.ctx, h.data(), z, bytes, 1));
        Fun_Check(y.data(), h.data(), nrow);
        mr_free(_env.ctx, x);
        mr_free(_env.ctx, yd);
        mr_free(_env.ctx, z);
    }

    // Fun_Check (include/utils.cuh:113-129)
    static void Fun_Check(const V* X, const V* Y, const int n)
    {
        V err(0.0);
        for (int i = 0; i < n; ++i) {
            const V d = std::abs(X[i] - Y[i]);
            if (d > std::abs(X[i]) * (V)0.0001) {
                err += d;
                std::cout << "X[" << i << "] = " << X[i] << ", Y[" << i << "] = " << Y[i] << "\n";
            }
        }
        std::cout << ", total error = " << err << std::endl;
    }

private:
    mr_reorder_stats st_{};

    std::ofstream& log(const ENV& _env) { return _env.FSTR; }
    static double reorder_ms(const mr_reorder_stats& s) { return s.total_ms - s.relabel_ms; }

    void plain(const char* name, int rc, const mr_reorder_stats& s)
    {
        if (!MR_GRAPH_OK(rc)) return;
        if (env_) *env_ << "&" << name << "&&" << reorder_ms(s) << "&" << s.relabel_ms;
    }
    std::ofstream* env_ = nullptr;

    // GAP-benchmark suffixes go through mr_gap_load (their weights kept, 1 for
    // unweighted formats); everything else is MatrixMarket, values forced to 1
    static bool gap_suffix(const std::string& p)
    {
        const std::string::size_type dot = p.rfind('.');
        if (dot == std::string::npos) return false;
        const std::string s = p.substr(dot);
        return s == ".el" || s == ".net" || s == ".wel" || s == ".gr" || s == ".graph" || s == ".sg" || s == ".wsg";
    }

    void load(const ENV& _env, const std::string& path)
    {
        env_ = &_env.FSTR;
        const int rc0 = gap_suffix(path) ? mr_gap_load(_env.ctx, path.c_str(), 0, kDtype, &mat)
                                         : mr_mtx_load(_env.ctx, path.c_str(), kDtype, 1, &mat);
        if (!MR_GRAPH_OK(rc0)) return;
        int64_t r = 0, c = 0, z = 0;
        int dt = 0;
        MR_GRAPH_OK(mr_mat_info(mat, &r, &c, &z, &dt));
        nrow = (int)r;
        nnz = (int)z;
        // ground truth y = A 1 on the loaded order (class.cu:182-207)
        void *x = nullptr, *yd = nullptr;
        const size_t bytes = sizeof(V) * (size_t)(nrow > 0 ? nrow : 1);
        if (!MR_GRAPH_OK(mr_malloc(_env.ctx, bytes, &x)) || !MR_GRAPH_OK(mr_malloc(_env.ctx, bytes, &yd))) return;
        std::vector<V> ones((size_t)nrow, (V)1);
        MR_GRAPH_OK(mr_memcpy(_env.ctx, x, ones.data(), bytes, 0));
        MR_GRAPH_OK(mr_spmv(mat, MR_KERNEL_CSR_VECTOR, x, yd, _env.stream));
        MR_GRAPH_OK(mr_stream_sync(_env.ctx, _env.stream));
        y.assign((size_t)nrow, (V)0);
        MR_GRAPH_OK(mr_memcpy(_env.ctx, y.data(), yd, bytes, 1));
        mr_free(_env.ctx, x);
        mr_free(_env.ctx, yd);
    }
};

}  // namespace mr
