/*
 * degree.c -- oracle restatement of the two degree-based relabels.
 * TEST INFRASTRUCTURE ONLY.
 *
 *   oracle_order_degree  GRAPH::relabel_degree (src/class.cu:1496-1549): the
 *                        degree of v is the number of COO entries with
 *                        row_ind == v (desc_kernel1), i.e. the column count
 *                        of v in M; vertices sorted by degree descending
 *                        (thrust::sort_by_key with greater<int>, whose radix
 *                        path for int keys is stable, so ties keep ascending
 *                        id); perm[sequence[k]] = k.
 *   oracle_order_dbg     GRAPH::dbg (src/class.cu:1683-1756): group =
 *                        ceilf(log2(column count + 1.0)), sorted by
 *                        (group desc, id asc) (DegIdCompare); perm as above.
 *
 * Both run as a counting sort by key, scanning ids in ascending order.
 * Pinned by hand-worked matrices whose row and column counts differ and by
 * the group boundaries 0, 1, 2-3, 4-7, ... (tests/test_oracle_basic.py).
 */
#include "oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

static void order_by_key(int32_t n, const uint32_t* key, int32_t* perm)
{
    uint32_t kmax = 0;
    for (int32_t v = 0; v < n; ++v)
        if (key[v] > kmax) kmax = key[v];
    int64_t* start = (int64_t*)calloc((size_t)kmax + 2, sizeof(int64_t));
    for (int32_t v = 0; v < n; ++v) start[key[v]]++;
    /* descending key: the largest key starts at 0 */
    int64_t acc = 0;
    for (int64_t k = (int64_t)kmax; k >= 0; --k) {
        const int64_t c = start[k];
        start[k] = acc;
        acc += c;
    }
    for (int32_t v = 0; v < n; ++v) perm[v] = (int32_t)start[key[v]]++;
    free(start);
}

static uint32_t* column_counts(int32_t n, const int32_t* rp, const int32_t* ci)
{
    uint32_t* cnt = (uint32_t*)calloc((size_t)(n > 0 ? n : 1), sizeof(uint32_t));
    for (int64_t q = 0; q < rp[n]; ++q) cnt[ci[q]]++;
    return cnt;
}

void oracle_order_degree(int32_t n, const int32_t* rp, const int32_t* ci, int32_t* perm)
{
    uint32_t* cnt = column_counts(n, rp, ci);
    order_by_key(n, cnt, perm);
    free(cnt);
}

void oracle_order_dbg(int32_t n, const int32_t* rp, const int32_t* ci, int32_t* perm)
{
    uint32_t* cnt = column_counts(n, rp, ci);
    for (int32_t v = 0; v < n; ++v) cnt[v] = (uint32_t)ceilf((float)log2((double)cnt[v] + 1.0));
    order_by_key(n, cnt, perm);
    free(cnt);
}
