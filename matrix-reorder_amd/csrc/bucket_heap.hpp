// bucket_heap.hpp -- Gorder's lazy integer bucket priority queue over dense
// node ids, with the reference's exact tie behaviour
// (include/integer_heap.cuh:1-155): a node enters a non-empty bucket right
// after its head, removing the head promotes head->right, a positive pending
// delta re-buckets at once, a non-positive one is applied when the node
// reaches the head of the top bucket.  Lists are index arrays instead of the
// reference's heap-allocated Node objects; the bucket table grows on demand.
#pragma once
#include <algorithm>
#include <cstddef>
#include <cstdint>
#include <vector>

namespace mr {

class BucketHeap {
public:
    void reset(int32_t n_nodes)
    {
        key1_.assign((size_t)n_nodes, 0);
        key2_.assign((size_t)n_nodes, 0);
        left_.resize((size_t)n_nodes);
        right_.resize((size_t)n_nodes);
        for (int32_t i = 0; i < n_nodes; ++i) { left_[(size_t)i] = i; right_[(size_t)i] = i; }
        if (htab_.empty()) htab_.assign(64, -1);
        else std::fill(htab_.begin(), htab_.begin() + (std::ptrdiff_t)std::min<size_t>(htab_.size(), (size_t)hmax_used_ + 1), -1);
        pmax_ = 0;
        cnt_ = 0;
        hmax_used_ = 0;
    }
    int64_t size() const { return cnt_; }

    void push(int32_t v)
    {
        insert(v);
        ++cnt_;
        if (key1_[(size_t)v] > pmax_) pmax_ = key1_[(size_t)v];
    }
    void modify(int32_t v, int32_t delta)
    {
        int32_t& k2 = key2_[(size_t)v];
        k2 += delta;
        if (k2 > 0) {
            remove(v);
            insert(v);
            if (key1_[(size_t)v] > pmax_) pmax_ = key1_[(size_t)v];
        }
    }
    // Integer_Heap::refresh (integer_heap.cuh:112-120): a node with a positive
    // key goes back to key 0 (plus its pending delta), entering that bucket
    // after its head
    void refresh(int32_t v)
    {
        if (key1_[(size_t)v] > 0) {
            remove(v);
            key1_[(size_t)v] = 0;
            insert(v);
        }
    }
    // the caller's direct write of PMAX (class.cu:3310)
    void set_pmax(int32_t p) { pmax_ = p; }

    int32_t pop()
    {
        while (cnt_ > 0) {
            while (pmax_ > 0 && htab_[(size_t)pmax_] < 0) --pmax_;
            int32_t t = htab_[(size_t)pmax_];
            while (t >= 0) {
                if (key2_[(size_t)t] == 0) {
                    remove(t);
                    --cnt_;
                    return t;
                }
                remove(t);
                insert(t);
                t = htab_[(size_t)pmax_];
            }
            if (pmax_ == 0) break;
        }
        return -1;
    }

private:
    void remove(int32_t v)
    {
        int32_t& slot = htab_[(size_t)key1_[(size_t)v]];
        const int32_t l = left_[(size_t)v], r = right_[(size_t)v];
        if (slot == v) {
            if (r == v) { slot = -1; return; }
            slot = r;
        }
        right_[(size_t)l] = r;
        left_[(size_t)r] = l;
        left_[(size_t)v] = v;
        right_[(size_t)v] = v;
    }
    void insert(int32_t v)
    {
        int32_t& k1 = key1_[(size_t)v];
        k1 += key2_[(size_t)v];
        key2_[(size_t)v] = 0;
        if ((size_t)k1 >= htab_.size()) htab_.resize(std::max<size_t>((size_t)k1 + 1, htab_.size() * 2), -1);
        if (k1 > hmax_used_) hmax_used_ = k1;
        int32_t& slot = htab_[(size_t)k1];
        if (slot < 0) {
            slot = v;
        } else {
            const int32_t h = slot;
            left_[(size_t)v] = h;
            right_[(size_t)v] = right_[(size_t)h];
            left_[(size_t)right_[(size_t)h]] = v;
            right_[(size_t)h] = v;
        }
    }

    std::vector<int32_t> key1_, key2_, left_, right_, htab_;
    int32_t pmax_ = 0;
    int32_t hmax_used_ = 0;
    int64_t cnt_ = 0;
};

}  // namespace mr
