// rabbit.cpp -- Rabbit Order CPU baseline (rabbitorder/src/rabbit_order.cpp in
// the reference, after Arai et al., "Rabbit Order: Just-in-Time Parallel
// Reordering for Fast Graph Analysis", IPDPS 2016).
//
// The reference runs the incremental aggregation on OpenMP threads with
// compare-and-swap merges, so its result depends on thread timing
// (rabbit_order.cpp:260-311).  This is the same algorithm executed in one
// well-defined order -- what the reference computes with one thread:
//
//  1. the pattern of M is symmetrised, self loops dropped, parallel edges
//     merged with summed weight (make_adj_list, :12-81);
//  2. vertices are visited by ascending degree (ties: lower id) (merge_order,
//     :103-115); each visit first gathers the edges of the vertex and of every
//     vertex merged into it, mapped to their current communities (unite,
//     :202-258), then joins the neighbour community u with the largest
//     modularity gain  w(v,u) - str(v) * str(u) / total  when it is positive
//     (find_best, :137-151; merge, :260-311);
//  3. the ids are a depth-first numbering of the merge dendrogram, community
//     by community in the order the top-level communities arose
//     (compute_perm, :380-420).
//
// Result: perm[v] = new id of v.
#include <algorithm>
#include <cstdint>
#include <numeric>
#include <utility>
#include <vector>

#include "parallel.hpp"
#include "reorder.hpp"

namespace mr {

namespace {

constexpr uint32_t kNone = 0xffffffffu;
using Edge = std::pair<uint32_t, float>;

struct Aggregation {
    std::vector<std::vector<Edge>> es;   // vertex -> (neighbour community, weight)
    std::vector<float> str;              // community strength; -1 while merged
    std::vector<uint32_t> coms, child, sibling, united;
    double total = 0.0;
    std::vector<uint32_t> tops;

    uint32_t trace(uint32_t v)
    {
        uint32_t c = v;
        while (coms[c] != c) c = coms[c];
        if (v != c && coms[v] != c) coms[v] = c;
        return c;
    }

    // sort by neighbour id and add up parallel edges
    static void compact(std::vector<Edge>& in, std::vector<Edge>& out)
    {
        out.clear();
        std::sort(in.begin(), in.end(), [](const Edge& a, const Edge& b) { return a.first < b.first; });
        for (const Edge& e : in) {
            if (!out.empty() && out.back().first == e.first) out.back().second += e.second;
            else out.push_back(e);
        }
    }

    void unite(uint32_t v, std::vector<Edge>& nbrs, std::vector<Edge>& tmp)
    {
        nbrs.clear();
        auto push = [&](uint32_t u) {
            for (const Edge& e : es[u]) {
                const uint32_t c = trace(e.first);
                if (c != v) nbrs.push_back({c, e.second});
            }
        };
        push(v);
        while (united[v] != child[v]) {
            const uint32_t c = child[v];
            for (uint32_t w = c; w != kNone && w != united[v]; w = sibling[w]) push(w);
            united[v] = c;
        }
        compact(nbrs, tmp);
        es[v].assign(tmp.begin(), tmp.end());
    }

    uint32_t merge(uint32_t v, std::vector<Edge>& nbrs, std::vector<Edge>& tmp)
    {
        unite(v, nbrs, tmp);
        const float vstr = str[v];
        str[v] = -1.0f;
        double dmax = 0.0;
        uint32_t best = v;
        for (const Edge& e : es[v]) {
            const double d = (double)e.second - (double)(vstr * str[e.first]) / total;
            if (dmax < d) {
                dmax = d;
                best = e.first;
            }
        }
        if (best == v) {
            str[v] = vstr;
            return v;
        }
        sibling[v] = child[best];
        str[best] += vstr;
        child[best] = v;
        coms[v] = best;
        return best;
    }
};

}  // namespace

void rabbit_order(int32_t n, const int32_t* rp, const int32_t* ci, uint32_t* new_id)
{
    if (n <= 0) return;
    // 1. symmetric pattern without self loops, weights = multiplicities
    std::vector<int64_t> deg((size_t)n + 1, 0);
    for (int32_t i = 0; i < n; ++i)
        for (int32_t q = rp[i]; q < rp[i + 1]; ++q) {
            const int32_t j = ci[q];
            if (j == i) continue;
            ++deg[(size_t)i];
            ++deg[(size_t)j];
        }
    std::vector<int64_t> off((size_t)n + 1, 0);
    for (int32_t i = 0; i < n; ++i) off[(size_t)i + 1] = off[(size_t)i] + deg[(size_t)i];
    std::vector<uint32_t> adj((size_t)std::max<int64_t>(off[(size_t)n], 1));
    std::vector<int64_t> cur(off.begin(), off.end() - 1);
    for (int32_t i = 0; i < n; ++i)
        for (int32_t q = rp[i]; q < rp[i + 1]; ++q) {
            const int32_t j = ci[q];
            if (j == i) continue;
            adj[(size_t)cur[(size_t)i]++] = (uint32_t)j;
            adj[(size_t)cur[(size_t)j]++] = (uint32_t)i;
        }
    Aggregation g;
    g.es.resize((size_t)n);
    g.str.assign((size_t)n, 0.0f);
    parallel_chunks(n, [&](int, int64_t b, int64_t e) {
        for (int64_t v = b; v < e; ++v) {
            uint32_t* first = adj.data() + off[(size_t)v];
            uint32_t* last = adj.data() + off[(size_t)v + 1];
            std::sort(first, last);
            auto& ev = g.es[(size_t)v];
            for (uint32_t* p = first; p != last; ++p) {
                if (!ev.empty() && ev.back().first == *p) ev.back().second += 1.0f;
                else ev.push_back({*p, 1.0f});
            }
            float s = 0.0f;
            for (const Edge& x : ev) s += x.second;
            g.str[(size_t)v] = s;
        }
    });
    adj.clear();
    adj.shrink_to_fit();
    double w = 0.0;
    for (int32_t v = 0; v < n; ++v) w += g.str[(size_t)v];
    g.total = w;
    g.coms.resize((size_t)n);
    std::iota(g.coms.begin(), g.coms.end(), 0u);
    g.child.assign((size_t)n, kNone);
    g.sibling.assign((size_t)n, kNone);
    g.united.assign((size_t)n, kNone);

    // 2. incremental aggregation by ascending degree
    std::vector<uint32_t> ord((size_t)n);
    std::iota(ord.begin(), ord.end(), 0u);
    std::stable_sort(ord.begin(), ord.end(),
                     [&](uint32_t a, uint32_t b) { return g.es[a].size() < g.es[b].size(); });
    std::vector<Edge> nbrs, tmp;
    for (uint32_t v : ord)
        if (g.merge(v, nbrs, tmp) == v) g.tops.push_back(v);

    // 3. depth-first numbering of every top-level community
    std::vector<uint32_t> stack;
    uint32_t next = 0;
    auto descend = [&](uint32_t v) {
        stack.push_back(v);
        while ((v = g.child[v]) != kNone) stack.push_back(v);
    };
    for (uint32_t top : g.tops) {
        descend(top);
        while (!stack.empty()) {
            const uint32_t v = stack.back();
            stack.pop_back();
            new_id[v] = next++;
            if (g.sibling[v] != kNone) descend(g.sibling[v]);
        }
    }
}

}  // namespace mr
