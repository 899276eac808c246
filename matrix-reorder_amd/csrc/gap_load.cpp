// gap_load.cpp -- graph files in the formats of the reference's GAP-benchmark
// reader (parsb/include/reader.h:50-160, 334-400) into a device CSR.
//
// Text formats are built the way its builder builds them
// (parsb/include/builder.h:216-415): n = largest vertex id + 1, the edges
// sorted by (u, v, weight), one edge kept per (u, v) (the smallest weight),
// self-loops dropped, and, when asked, every missing reverse edge added with
// the same weight.  Serialized graphs (.sg / .wsg) are read as stored.  The
// CSR's row u holds the out-neighbours v of u, values = the weights (1 for
// unweighted files).  Unlike the reference reader, a malformed file is an
// error instead of a silently shorter edge list.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "common.hpp"

using namespace mr;

namespace {

struct GapEdge {
    int64_t u, v;
    double w;
};

int read_whole(const char* path, std::string& buf)
{
    FILE* f = std::fopen(path, "rb");
    if (!f) return set_error(MR_ERR_IO, "cannot open %s", path);
    std::fseek(f, 0, SEEK_END);
    const long sz = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    if (sz < 0) {
        std::fclose(f);
        return set_error(MR_ERR_IO, "cannot size %s", path);
    }
    buf.assign((size_t)sz, '\0');
    const size_t got = sz > 0 ? std::fread(&buf[0], 1, (size_t)sz, f) : 0;
    std::fclose(f);
    if (got != (size_t)sz) return set_error(MR_ERR_IO, "short read of %s", path);
    return MR_OK;
}

// whitespace-separated numbers of one text region [p, e)
struct Tokens {
    const char* p;
    const char* e;
    void skip()
    {
        while (p < e && (*p == ' ' || *p == '\t' || *p == '\r' || *p == '\n')) ++p;
    }
    bool at_end()
    {
        skip();
        return p >= e;
    }
    // false at the end of the region or on a token that is not a number
    bool integer(int64_t& v)
    {
        skip();
        if (p >= e) return false;
        char* end = nullptr;
        const long long x = std::strtoll(p, &end, 10);
        if (end == p || end > e) return false;
        p = end;
        v = x;
        return true;
    }
    bool real(double& v)
    {
        skip();
        if (p >= e) return false;
        char* end = nullptr;
        const double x = std::strtod(p, &end);
        if (end == p || end > e) return false;
        p = end;
        v = x;
        return true;
    }
};

std::string suffix_of(const char* path)
{
    const char* dot = std::strrchr(path, '.');
    return dot ? std::string(dot) : std::string();
}

// reader.h:52-60 (.el / .net) and :63-71 (.wel): pairs or triples to the end
int parse_pairs(const std::string& buf, bool weighted, const char* path, std::vector<GapEdge>& el)
{
    Tokens t{buf.data(), buf.data() + buf.size()};
    while (!t.at_end()) {
        GapEdge g{0, 0, 1.0};
        if (!t.integer(g.u) || !t.integer(g.v) || (weighted && !t.real(g.w)))
            return set_error(MR_ERR_IO, "%s: malformed edge %zu", path, el.size());
        el.push_back(g);
    }
    return MR_OK;
}

// reader.h:74-94 (.gr, DIMACS): "a u v w" lines, 1-based; other lines ignored
int parse_dimacs(const std::string& buf, const char* path, std::vector<GapEdge>& el)
{
    const char* p = buf.data();
    const char* end = p + buf.size();
    while (p < end) {
        const char* nl = static_cast<const char*>(std::memchr(p, '\n', (size_t)(end - p)));
        const char* le = nl ? nl : end;
        if (*p == 'a') {
            Tokens t{p + 1, le};
            GapEdge g{0, 0, 1.0};
            if (!t.integer(g.u) || !t.integer(g.v) || !t.real(g.w))
                return set_error(MR_ERR_IO, "%s: malformed arc %zu", path, el.size());
            g.u -= 1;
            g.v -= 1;
            el.push_back(g);
        }
        p = nl ? nl + 1 : end;
    }
    return MR_OK;
}

// reader.h:97-162 (.graph, METIS): '%' comment lines, a header "n m [fmt]"
// (fmt 1: a weight after every neighbour; 0 or 100: none), then one line of
// 1-based neighbours per vertex (an empty line is a vertex without any)
int parse_metis(const std::string& buf, const char* path, std::vector<GapEdge>& el)
{
    const char* p = buf.data();
    const char* end = p + buf.size();
    auto next_line = [&](const char*& ls, const char*& le) -> bool {
        while (p < end) {
            const char* nl = static_cast<const char*>(std::memchr(p, '\n', (size_t)(end - p)));
            ls = p;
            le = nl ? nl : end;
            p = nl ? nl + 1 : end;
            if (*ls != '%') return true;
        }
        return false;
    };
    const char *ls = nullptr, *le = nullptr;
    if (!next_line(ls, le)) return set_error(MR_ERR_IO, "%s: no METIS header", path);
    Tokens h{ls, le};
    int64_t nv = 0, ne = 0, fmt = 0;
    if (!h.integer(nv) || !h.integer(ne)) return set_error(MR_ERR_IO, "%s: bad METIS header", path);
    if (!h.at_end() && !h.integer(fmt)) return set_error(MR_ERR_IO, "%s: bad METIS format field", path);
    if (fmt != 0 && fmt != 1 && fmt != 100)
        return set_error(MR_ERR_UNSUPPORTED, "%s: METIS format %lld is not supported (0, 1, 100)", path,
                         (long long)fmt);
    const bool weighted = fmt == 1;
    (void)ne;
    for (int64_t u = 0; u < nv; ++u) {
        if (!next_line(ls, le)) break;   // missing trailing lines: vertices without neighbours
        Tokens t{ls, le};
        while (!t.at_end()) {
            GapEdge g{u, 0, 1.0};
            if (!t.integer(g.v) || (weighted && !t.real(g.w)))
                return set_error(MR_ERR_IO, "%s: malformed adjacency of vertex %lld", path, (long long)u + 1);
            g.v -= 1;
            el.push_back(g);
        }
    }
    return MR_OK;
}

// builder.h:216-415: sort, one edge per (u, v), no self-loops, optional
// reverse edges; then the CSR of out-neighbours
int build_csr(std::vector<GapEdge>& el, bool symmetrize, const char* path, int64_t& n, std::vector<int32_t>& rp,
              std::vector<int32_t>& ci, std::vector<double>& val)
{
    int64_t maxid = -1;
    for (const GapEdge& g : el) {
        if (g.u < 0 || g.v < 0) return set_error(MR_ERR_IO, "%s: negative vertex id", path);
        maxid = std::max(maxid, std::max(g.u, g.v));
    }
    if (maxid >= INT32_MAX) return set_error(MR_ERR_UNSUPPORTED, "%s: vertex ids beyond int32", path);
    n = maxid + 1;
    auto less = [](const GapEdge& a, const GapEdge& b) {
        return a.u != b.u ? a.u < b.u : (a.v != b.v ? a.v < b.v : a.w < b.w);
    };
    std::sort(el.begin(), el.end(), less);
    el.erase(std::unique(el.begin(), el.end(), [](const GapEdge& a, const GapEdge& b) {
                 return a.u == b.u && a.v == b.v;
             }),
             el.end());
    el.erase(std::remove_if(el.begin(), el.end(), [](const GapEdge& g) { return g.u == g.v; }), el.end());
    if (symmetrize) {
        const size_t m = el.size();
        auto has = [&](int64_t u, int64_t v) {
            const GapEdge key{u, v, -1e308};
            auto it = std::lower_bound(el.begin(), el.begin() + (std::ptrdiff_t)m, key, less);
            return it != el.begin() + (std::ptrdiff_t)m && it->u == u && it->v == v;
        };
        std::vector<GapEdge> add;
        for (size_t i = 0; i < m; ++i)
            if (!has(el[i].v, el[i].u)) add.push_back(GapEdge{el[i].v, el[i].u, el[i].w});
        el.insert(el.end(), add.begin(), add.end());
        std::sort(el.begin(), el.end(), less);
    }
    if ((int64_t)el.size() >= INT32_MAX) return set_error(MR_ERR_UNSUPPORTED, "%s: more than 2^31 edges", path);
    rp.assign((size_t)n + 1, 0);
    ci.resize(el.size());
    val.resize(el.size());
    for (size_t i = 0; i < el.size(); ++i) {
        rp[(size_t)el[i].u + 1]++;
        ci[i] = (int32_t)el[i].v;
        val[i] = el[i].w;
    }
    for (int64_t r = 0; r < n; ++r) rp[(size_t)r + 1] += rp[(size_t)r];
    return MR_OK;
}

// reader.h:334-400 (.sg / .wsg): bool directed, int64 edges, int64 nodes,
// int64 offsets[nodes + 1], then int32 neighbours (.sg) or int32 (v, w) pairs
// (.wsg); a directed graph's incoming lists follow and are not needed
int parse_serialized(const std::string& buf, bool weighted, const char* path, int64_t& n, std::vector<int32_t>& rp,
                     std::vector<int32_t>& ci, std::vector<double>& val)
{
    const size_t head = 1 + 2 * sizeof(int64_t);
    if (buf.size() < head) return set_error(MR_ERR_IO, "%s: truncated header", path);
    int64_t ne = 0, nv = 0;
    std::memcpy(&ne, buf.data() + 1, sizeof(int64_t));
    std::memcpy(&nv, buf.data() + 1 + sizeof(int64_t), sizeof(int64_t));
    if (nv < 0 || ne < 0 || nv >= INT32_MAX || ne >= INT32_MAX)
        return set_error(MR_ERR_UNSUPPORTED, "%s: %lld nodes / %lld edges do not fit int32", path, (long long)nv,
                         (long long)ne);
    const size_t ent = weighted ? 2 * sizeof(int32_t) : sizeof(int32_t);
    const size_t need = head + sizeof(int64_t) * ((size_t)nv + 1) + ent * (size_t)ne;
    if (buf.size() < need) return set_error(MR_ERR_IO, "%s: truncated (%zu of %zu bytes)", path, buf.size(), need);
    const char* offs = buf.data() + head;
    const char* nb = offs + sizeof(int64_t) * ((size_t)nv + 1);
    n = nv;
    rp.resize((size_t)nv + 1);
    for (int64_t i = 0; i <= nv; ++i) {
        int64_t o = 0;
        std::memcpy(&o, offs + sizeof(int64_t) * (size_t)i, sizeof(int64_t));
        if (o < 0 || o > ne || (i > 0 && o < rp[(size_t)i - 1]) || (i == 0 && o != 0) || (i == nv && o != ne))
            return set_error(MR_ERR_MALFORMED_CSR, "%s: bad offset %lld at %lld", path, (long long)o, (long long)i);
        rp[(size_t)i] = (int32_t)o;
    }
    ci.resize((size_t)ne);
    val.assign((size_t)ne, 1.0);
    for (int64_t k = 0; k < ne; ++k) {
        int32_t v = 0;
        std::memcpy(&v, nb + ent * (size_t)k, sizeof(int32_t));
        ci[(size_t)k] = v;
        if (weighted) {
            int32_t w = 0;
            std::memcpy(&w, nb + ent * (size_t)k + sizeof(int32_t), sizeof(int32_t));
            val[(size_t)k] = (double)w;
        }
    }
    return MR_OK;
}

}  // namespace

extern "C" {

int mr_gap_load(mr_ctx ctx, const char* path, int symmetrize, int store_dtype, mr_mat* out)
{
    if (!ctx || !path || !out) return set_error(MR_ERR_INVALID_ARG, "null argument");
    const std::string suf = suffix_of(path);
    const bool text = suf == ".el" || suf == ".net" || suf == ".wel" || suf == ".gr" || suf == ".graph";
    const bool serial = suf == ".sg" || suf == ".wsg";
    if (!text && !serial)
        return set_error(MR_ERR_UNSUPPORTED, "%s: suffix '%s' is not a GAP graph (.el .net .wel .gr .graph .sg .wsg;"
                         " MatrixMarket files go through mr_mtx_load)", path, suf.c_str());
    std::string buf;
    MR_TRY(read_whole(path, buf));
    int64_t n = 0;
    std::vector<int32_t> rp, ci;
    std::vector<double> val;
    if (serial) {
        MR_TRY(parse_serialized(buf, suf == ".wsg", path, n, rp, ci, val));
    } else {
        std::vector<GapEdge> el;
        if (suf == ".el" || suf == ".net") MR_TRY(parse_pairs(buf, false, path, el));
        else if (suf == ".wel") MR_TRY(parse_pairs(buf, true, path, el));
        else if (suf == ".gr") MR_TRY(parse_dimacs(buf, path, el));
        else MR_TRY(parse_metis(buf, path, el));
        buf.clear();
        buf.shrink_to_fit();
        MR_TRY(build_csr(el, symmetrize != 0, path, n, rp, ci, val));
    }
    return mr_csr_create(ctx, n, n, (int64_t)ci.size(), rp.data(), ci.data(), val.data(), MR_F64, store_dtype, out);
}

}  // extern "C"
