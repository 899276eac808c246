// cpu_order.cpp -- the host-side reorder baselines of the reference:
// sequential Gorder (GRAPH::gorder1, src/class.cu:1825-2013), the tile-aware
// order (GRAPH::order, src/class.cu:3158-3365) and Reverse Cuthill-McKee
// (rcm/src/rcm.cpp:1-154, rcm/src/queue.cpp:39-55).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "bucket_heap.hpp"
#include "reorder.hpp"

namespace mr {

void gorder_cpu(int32_t n, const int32_t* rp, const int32_t* ci, const int32_t* tp, const int32_t* ti,
                int32_t window, uint32_t* order)
{
    const int32_t L = (int32_t)(std::sqrt((double)n) / 4);
    // start vertex: first maximum of the SpMV row length
    int32_t start = 0;
    for (int32_t v = 1; v < n; ++v)
        if (rp[v + 1] - rp[v] > rp[start + 1] - rp[start]) start = v;
    BucketHeap heap;
    heap.reset(n);
    std::vector<uint8_t> sts((size_t)n, 1);
    for (int32_t v = 0; v < n; ++v) heap.push(v);
    heap.modify(start, 1);
    // one window step of node with sign d: neighbours through the out-list
    // (when short), in-neighbours, and siblings through short out-lists
    auto apply = [&](int32_t node, int32_t d) {
        if (tp[node + 1] - tp[node] < L)
            for (int32_t c = tp[node]; c < tp[node + 1]; ++c)
                if (sts[(size_t)ti[c]]) heap.modify(ti[c], d);
        for (int32_t a = rp[node]; a < rp[node + 1]; ++a) {
            const int32_t y = ci[a];
            if (sts[(size_t)y]) heap.modify(y, d);
            if (tp[y + 1] - tp[y] < L)
                for (int32_t c = tp[y]; c < tp[y + 1]; ++c)
                    if (sts[(size_t)ti[c]]) heap.modify(ti[c], d);
        }
    };
    // with MR_PROGRESS set, a line on stderr at most every 20 s (a run at
    // R-MAT s22 takes minutes)
    const bool progress = std::getenv("MR_PROGRESS") != nullptr;
    auto last = std::chrono::steady_clock::now();
    for (int32_t id = 0; id < n; ++id) {
        const int32_t node = heap.pop();
        sts[(size_t)node] = 0;
        order[id] = (uint32_t)node;
        apply(node, 1);
        if (id >= window) apply((int32_t)order[id - window], -1);
        if (progress && (id & 0xffff) == 0) {
            const auto now = std::chrono::steady_clock::now();
            if (now - last >= std::chrono::seconds(20)) {
                std::fprintf(stderr, "[gorder_cpu] %d of %d vertices placed\n", id, n);
                last = now;
            }
        }
    }
}

// GRAPH::order (src/class.cu:3158-3365): greedy order that keeps rows sharing
// x entries inside one tile of width1 x entries.  curr counts the x entries
// laid out so far and [start, start + width1) is the current tile.  A placed
// row whose entries fit raises every unplaced row sharing one of its columns
// (columns of width2 or more entries skipped); a row that crosses the tile
// boundary first drops every row the old tile raised back to key 0
// (refresh), resets the top bucket, moves the tile to the one holding its
// last entry and raises rows only for its entries inside it.
void tile_order(int32_t n, const int32_t* rp, const int32_t* ci, const int32_t* tp, const int32_t* ti,
                int32_t width1, int32_t width2, uint32_t* order)
{
    int32_t first = 0;
    for (int32_t v = 1; v < n; ++v)
        if (rp[v + 1] - rp[v] > rp[first + 1] - rp[first]) first = v;
    BucketHeap heap;
    heap.reset(n);
    std::vector<uint8_t> sts((size_t)n, 1);
    std::vector<int32_t> tile((size_t)width1);
    for (int32_t v = 0; v < n; ++v) heap.push(v);
    heap.modify(first, 1);
    auto raise = [&](int32_t col) {
        if (tp[col + 1] - tp[col] >= width2) return;
        for (int32_t j = tp[col]; j < tp[col + 1]; ++j)
            if (sts[(size_t)ti[j]]) heap.modify(ti[j], 1);
    };
    int64_t start = 0, curr = 0;
    for (int32_t id = 0; id < n; ++id) {
        const int32_t node = heap.pop();
        sts[(size_t)node] = 0;
        order[id] = (uint32_t)node;
        const int32_t deg = rp[node + 1] - rp[node];
        if (curr - start + deg < width1) {
            for (int32_t i = 0; i < deg; ++i) {
                const int32_t col = ci[rp[node] + i];
                tile[(size_t)(curr - start)] = col;
                raise(col);
                ++curr;
            }
            continue;
        }
        for (int64_t i = start; i < curr; ++i) {
            const int32_t col = tile[(size_t)(i - start)];
            if (tp[col + 1] - tp[col] >= width2) continue;
            for (int32_t j = tp[col]; j < tp[col + 1]; ++j)
                if (sts[(size_t)ti[j]]) heap.refresh(ti[j]);
        }
        heap.set_pmax(0);
        start = (curr + deg) / width1 * width1;
        for (int32_t i = 0; i < deg; ++i, ++curr) {
            if (curr < start) continue;
            const int32_t col = ci[rp[node] + i];
            tile[(size_t)(curr - start)] = col;
            raise(col);
        }
    }
}

void rcm_order(int32_t n, int64_t nz, const int32_t* rows, const int32_t* cols, uint32_t* order)
{
    // degree = off-diagonal entries touching the vertex, both directions
    std::vector<int32_t> degree((size_t)n, 0);
    for (int64_t i = 0; i < nz; ++i)
        if (rows[i] != cols[i]) { degree[(size_t)rows[i]]++; degree[(size_t)cols[i]]++; }
    std::vector<int64_t> start((size_t)n + 1, 0);
    for (int32_t v = 0; v < n; ++v) start[(size_t)v + 1] = start[(size_t)v] + degree[(size_t)v];
    std::vector<int32_t> nb((size_t)std::max<int64_t>(start[(size_t)n], 1));
    std::vector<int64_t> fill(start.begin(), start.end() - 1);
    for (int64_t i = 0; i < nz; ++i)
        if (rows[i] != cols[i]) {
            nb[(size_t)fill[(size_t)rows[i]]++] = cols[i];
            nb[(size_t)fill[(size_t)cols[i]]++] = rows[i];
        }
    std::vector<uint8_t> in_perm((size_t)n, 0), in_q((size_t)n, 0);
    std::vector<int32_t> Q;
    Q.reserve((size_t)n);
    std::vector<int32_t> tmp;
    size_t qh = 0;
    int32_t added = 0, index = 0;
    // lowest degree vertex not yet placed, first index on ties: the first
    // unplaced entry of the (degree, id) order, found with a forward cursor
    std::vector<int32_t> by_deg((size_t)n);
    for (int32_t v = 0; v < n; ++v) by_deg[(size_t)v] = v;
    std::stable_sort(by_deg.begin(), by_deg.end(),
                     [&](int32_t a, int32_t b) { return degree[(size_t)a] < degree[(size_t)b]; });
    size_t cursor = 0;
    auto min_node = [&]() {
        while (cursor < by_deg.size() && in_perm[(size_t)by_deg[cursor]]) ++cursor;
        const int32_t mn = by_deg[cursor];
        in_q[(size_t)mn] = 1;
        Q.push_back(mn);
    };
    min_node();
    while (qh < Q.size()) {
        const int32_t e = Q[qh++];
        in_perm[(size_t)e] = 1;
        order[index++] = (uint32_t)e;
        ++added;
        tmp.clear();
        for (int64_t k = start[(size_t)e]; k < start[(size_t)e + 1]; ++k) {
            const int32_t w = nb[(size_t)k];
            if (!in_perm[(size_t)w] && !in_q[(size_t)w]) { in_q[(size_t)w] = 1; tmp.push_back(w); }
        }
        // queueSort: exchange sort swapping on >= (its tie order is part of the result)
        const size_t t = tmp.size();
        for (size_t i = 0; i < t; ++i)
            for (size_t j = i + 1; j < t; ++j)
                if (degree[(size_t)tmp[i]] >= degree[(size_t)tmp[j]]) std::swap(tmp[i], tmp[j]);
        for (int32_t w : tmp) Q.push_back(w);
        if (qh == Q.size() && added != n) min_node();
    }
    std::reverse(order, order + n);
}

}  // namespace mr
