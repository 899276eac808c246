/*
 * gorder.c -- oracle restatements of the Gorder family.  TEST INFRASTRUCTURE ONLY.
 *
 *   oracle_get_width         src/class.cu:733-750  (float arithmetic kept)
 *   oracle_pgorder_schedule  src/class.cu:2995-3016 (Pstart/Pleng/Ostart/Oleng)
 *   oracle_pgorder_gpu_loop  src/class.cu:3041-3059 with gorder_gpu_cpu_kernel1
 *   (_ex: hub cap and a batch cap as parameters)
 *                            (:2357-2369), gorder6_kernel2 (:2897-2913),
 *                            gorder6_kernel3 (:2917-2953) and the stable
 *                            descending CUB radix sort (:3046,3052), restated
 *                            as "score descending, vertex id ascending".
 *   oracle_order_on_cpu      src/class.cu:2220-2314
 *   oracle_pgorder           src/class.cu:2955-3140 (CPU phase from batch log2(W),
 *                            :3077-3093; perm scatter :3101-3102)
 *   oracle_gorder1           src/class.cu:1825-2013
 *   oracle_gorder5           src/class.cu:2727-2873 (fixed batches, column-count
 *                            start scores, no hub caps, CPU phase on every batch)
 *   oracle_tile_order        src/class.cu:3158-3365 (GRAPH::order)
 *
 * Divergence kept on purpose: for n < 32 the reference's first window reads
 * d_permutation past n (Oleng[0] = 32 is never clipped, class.cu:3003); here the
 * window is clipped to n.
 *
 * Pinned by: heap golden traces; hand-derived schedule for W=1024 and small W;
 * brute-force top-|P| selection per batch on small graphs (selection equals the
 * first |P| of all unplaced vertices by naive 2-hop score, id ascending); the
 * max-score invariant of every gorder1/order_on_cpu pop against a naive score
 * recomputation; hand-worked tiny graphs (tests/test_oracle_gorder.py).
 */
#include "oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

int32_t oracle_get_width(int64_t n, int64_t nnz)
{
    /* float bound = round(log2(sqrt((float)nrow / (nnz / nrow)))) */
    int32_t nrow = (int32_t)n, nz = (int32_t)nnz;
    float q = (float)nrow / (float)(nz / nrow);
    float bound = roundf(log2f(sqrtf(q)));
    if (bound < 10.0f) return 1024;
    if (bound >= 10.0f && bound < 12.0f) return 16384;
    return 262144;
}

void oracle_transpose(int32_t n_rows, int32_t n_cols, const int32_t* rp,
                      const int32_t* ci, int32_t* tp, int32_t* ti)
{
    memset(tp, 0, sizeof(int32_t) * ((size_t)n_cols + 1));
    for (int32_t i = 0; i < n_rows; ++i)
        for (int32_t k = rp[i]; k < rp[i + 1]; ++k) tp[ci[k] + 1]++;
    for (int32_t j = 0; j < n_cols; ++j) tp[j + 1] += tp[j];
    int32_t* fill = (int32_t*)malloc(sizeof(int32_t) * ((size_t)n_cols + 1));
    memcpy(fill, tp, sizeof(int32_t) * ((size_t)n_cols + 1));
    for (int32_t i = 0; i < n_rows; ++i)              /* rows ascending -> sorted */
        for (int32_t k = rp[i]; k < rp[i + 1]; ++k) ti[fill[ci[k]]++] = i;
    free(fill);
}

/* ------------------------------------------------------------------------ */
static int32_t ilog2_floor(double v)
{
    int32_t r = 0;
    while (v >= 2.0) { v /= 2.0; ++r; }
    return r;
}

int32_t oracle_pgorder_schedule(int32_t n, int32_t W, int32_t cap, int32_t* Ps,
                                int32_t* Pl, int32_t* Os, int32_t* Ol)
{
    /* base = 5; batch_size = ceil(n/W) + log2(W / 2^5) (class.cu:2996-2998) */
    const int32_t base = 5;
    const int32_t k = (int32_t)log2((double)W / pow(2.0, base));
    const int32_t B = (n + W - 1) / W + k;
    for (int32_t b = 0; b < B && b < cap; ++b) {
        if (b == 0) { Ps[0] = 0; Pl[0] = 32; Os[0] = 0; Ol[0] = 32; continue; }
        const double p1 = pow(2.0, b + base - 1), p2 = pow(2.0, b + base);
        int32_t start = (p1 < W) ? (int32_t)p1 : (b - k) * W;
        int32_t leng = (p1 < W) ? (int32_t)p1 : W;
        Ps[b] = start;
        Pl[b] = (start + leng < n) ? leng : n - start;
        start = (p1 < W) ? 0 : (b - k) * W;
        leng = (p2 < W) ? (int32_t)p2 : W;
        Os[b] = start;
        Ol[b] = (start + leng < n) ? leng : n - start;
    }
    (void)ilog2_floor;
    return B;
}

/* ------------------------------------------------------------------------ */
/* The reference orders all n vertices each batch with a stable descending
 * radix sort of the scores over identity keys (class.cu:3046,3052), i.e. by
 * (score descending, vertex id ascending), and takes the first |P|.  That
 * order is total, so the first |P| are the |P| largest of the composite key
 * (score << 32) | ~id, found here by a selection and a sort of the selected. */
typedef struct { uint32_t score; uint32_t id; } skey;

/* the stable descending order written as a comparison (gorder5 sorts with it) */
static int cmp_desc_stable(const void* a, const void* b)
{
    const skey* x = (const skey*)a;
    const skey* y = (const skey*)b;
    if (x->score != y->score) return (x->score > y->score) ? -1 : 1;
    return (x->id < y->id) ? -1 : (x->id > y->id);
}

static void swap64(uint64_t* a, uint64_t* b) { uint64_t t = *a; *a = *b; *b = t; }

/* move the k largest of a[0..m) to a[0..k) (any order) */
static void select_top(uint64_t* a, int64_t m, int64_t k)
{
    int64_t lo = 0, hi = m - 1;
    while (lo < hi) {
        /* median of three as pivot; partition descending */
        const int64_t mid = lo + (hi - lo) / 2;
        if (a[mid] > a[lo]) swap64(&a[mid], &a[lo]);
        if (a[hi] > a[lo]) swap64(&a[hi], &a[lo]);
        if (a[hi] > a[mid]) swap64(&a[hi], &a[mid]);
        const uint64_t piv = a[mid];
        int64_t i = lo, j = hi;
        while (i <= j) {
            while (a[i] > piv) ++i;
            while (a[j] < piv) --j;
            if (i <= j) { swap64(&a[i], &a[j]); ++i; --j; }
        }
        if (k - 1 <= j) hi = j;
        else if (k - 1 >= i) lo = i;
        else break;
    }
}

static int cmp_u64_desc(const void* a, const void* b)
{
    const uint64_t x = *(const uint64_t*)a, y = *(const uint64_t*)b;
    return x > y ? -1 : (x < y);
}

int oracle_pgorder_gpu_loop_ex(int32_t n, const int32_t* rp, const int32_t* ci,
                               const int32_t* tp, const int32_t* ti, int32_t W, int32_t hub_cap,
                               int32_t max_batches, uint32_t* placement)
{
    const int32_t cap = (n + W - 1) / W + 64;
    int32_t *Ps = malloc(sizeof(int32_t) * cap), *Pl = malloc(sizeof(int32_t) * cap);
    int32_t *Os = malloc(sizeof(int32_t) * cap), *Ol = malloc(sizeof(int32_t) * cap);
    int32_t B = oracle_pgorder_schedule(n, W, cap, Ps, Pl, Os, Ol);
    if (max_batches >= 0 && max_batches < B) B = max_batches;
    uint32_t* score = (uint32_t*)malloc(sizeof(uint32_t) * ((size_t)n + 1));
    uint64_t* keys = (uint64_t*)malloc(sizeof(uint64_t) * ((size_t)n + 1));
    int32_t* hcnt = (int32_t*)calloc((size_t)n + 1, sizeof(int32_t));
    int32_t* touched = (int32_t*)malloc(sizeof(int32_t) * ((size_t)n + 1));
    /* kernel1: score = rowlen | 1<<31 (offset2 = SpMV row pointer) */
    for (int32_t v = 0; v < n; ++v) score[v] = (uint32_t)(rp[v + 1] - rp[v]) | 0x80000000u;
    for (int32_t b = 0; b < B; ++b) {
        /* the first Pleng vertices by (score desc, id asc) */
        const int64_t take = Pl[b] < n ? Pl[b] : n;
        int64_t m = 0;
        for (int32_t v = 0; v < n; ++v)
            if (score[v]) keys[m++] = ((uint64_t)score[v] << 32) | (uint32_t)~(uint32_t)v;
        if (m < take)   /* placed vertices (score 0) follow in id order */
            for (int32_t v = 0; v < n && m < take; ++v)
                if (!score[v]) keys[m++] = (uint32_t)~(uint32_t)v;
        if (take > 0) {
            select_top(keys, m, take);
            qsort(keys, (size_t)take, sizeof(uint64_t), cmp_u64_desc);
        }
        /* kernel2: keep only the status bit; place the first Pleng */
        for (int32_t v = 0; v < n; ++v) score[v] &= 0x80000000u;
        for (int64_t i = 0; i < take; ++i) {
            const uint32_t v = ~(uint32_t)(keys[i] & 0xffffffffu);
            score[v] = 0u;
            placement[Ps[b] + i] = v;
        }
        /* kernel3: window 2-hop sibling counts, hubs (>= hub_cap) skipped.
         * The reference adds 1 to score[z] once per (x, y, z) path; the same
         * sum is taken here as h[y] per (y, z), h[y] = number of window x
         * reaching y (integer addition, so the order does not matter).  It
         * keeps hub-heavy first windows from walking a column list once per
         * window vertex. */
        int32_t olen = Ol[b];
        if (Os[b] + olen > n) olen = n - Os[b];
        int64_t nt = 0;
        for (int32_t w = 0; w < olen; ++w) {
            const int32_t x = (int32_t)placement[Os[b] + w];
            if (rp[x + 1] - rp[x] >= hub_cap) continue;
            for (int32_t a = rp[x]; a < rp[x + 1]; ++a) {
                const int32_t y = ci[a];
                if (tp[y + 1] - tp[y] >= hub_cap) continue;
                if (hcnt[y]++ == 0) touched[nt++] = y;
            }
        }
        for (int64_t q = 0; q < nt; ++q) {
            const int32_t y = touched[q];
            for (int32_t c = tp[y]; c < tp[y + 1]; ++c) {
                const int32_t z = ti[c];
                if (score[z] >> 31) score[z] += (uint32_t)hcnt[y];
            }
            hcnt[y] = 0;
        }
    }
    free(Ps); free(Pl); free(Os); free(Ol); free(score); free(keys); free(hcnt); free(touched);
    return B;
}

int oracle_pgorder_gpu_loop(int32_t n, const int32_t* rp, const int32_t* ci,
                            const int32_t* tp, const int32_t* ti, int32_t W,
                            uint32_t* placement)
{
    /* the 10,000 hub caps of gorder6_kernel3 (class.cu:2930,2938) */
    return oracle_pgorder_gpu_loop_ex(n, rp, ci, tp, ti, W, 10000, -1, placement);
}

/* ------------------------------------------------------------------------ */
static int cmp_pair_first(const void* a, const void* b)
{
    const int32_t x = *(const int32_t*)a, y = *(const int32_t*)b;
    return (x > y) - (x < y);
}

void oracle_order_on_cpu(int32_t n, const int32_t* rp, const int32_t* ci,
                         const int32_t* tp, const int32_t* ti, int32_t* batch_of,
                         uint32_t* slice, int32_t batch_id, int32_t len,
                         int32_t window, int32_t degree)
{
    (void)n;
    if (len <= 0) return;
    /* start: first vertex of the slice with strictly maximal out-list length
     * (offset1 = transpose of M) */
    int32_t start = (int32_t)slice[0];
    int32_t deg = tp[start + 1] - tp[start];
    for (int32_t i = 0; i < len; ++i) {
        const int32_t v = (int32_t)slice[i];
        if (tp[v + 1] - tp[v] > deg) { deg = tp[v + 1] - tp[v]; start = v; }
    }
    /* local node ids = slice positions; heap over local ids */
    int32_t* local = (int32_t*)malloc(sizeof(int32_t) * (size_t)len);
    uint8_t* sts = (uint8_t*)malloc((size_t)len);
    for (int32_t i = 0; i < len; ++i) local[i] = (int32_t)slice[i];
    oheap h;
    oheap_init(&h, len, (int64_t)window * len * 8 + 64);
    /* position lookup through batch_of: batch_of[v] = batch id, and a map
     * vertex -> local index kept in a side table */
    int32_t* pos = (int32_t*)malloc(sizeof(int32_t) * (size_t)len);
    for (int32_t i = 0; i < len; ++i) {
        batch_of[local[i]] = batch_id;
        sts[i] = 1;
        pos[i] = i;
        oheap_push(&h, i);
    }
    (void)pos;
    /* map vertex -> local index: a sorted copy for bsearch */
    int32_t* sorted = (int32_t*)malloc(sizeof(int32_t) * (size_t)len * 2);
    for (int32_t i = 0; i < len; ++i) { sorted[2 * i] = local[i]; sorted[2 * i + 1] = i; }
    /* by vertex (the vertices of a slice are distinct) */
    qsort(sorted, (size_t)len, 2 * sizeof(int32_t), cmp_pair_first);
#define LOCAL_OF(v, out) do { int32_t lo_ = 0, hi_ = len - 1; out = -1;          \
        while (lo_ <= hi_) { int32_t mid_ = (lo_ + hi_) >> 1;                    \
            if (sorted[2 * mid_] == (v)) { out = sorted[2 * mid_ + 1]; break; } \
            if (sorted[2 * mid_] < (v)) lo_ = mid_ + 1; else hi_ = mid_ - 1; }  \
    } while (0)
    {
        int32_t ls; LOCAL_OF(start, ls);
        oheap_modify(&h, ls, 1);
    }
    for (int32_t id = 0; h.cnt > 0; ++id) {
        const int32_t li = oheap_pop(&h);
        sts[li] = 0;
        const int32_t node = local[li];
        slice[id] = (uint32_t)node;
        /* increase: siblings reached through the node's in-list */
        if (rp[node + 1] - rp[node] < degree) {
            for (int32_t a = rp[node]; a < rp[node + 1]; ++a) {
                const int32_t y = ci[a];
                if (tp[y + 1] - tp[y] < degree) {
                    for (int32_t c = tp[y]; c < tp[y + 1]; ++c) {
                        const int32_t z = ti[c];
                        if (batch_of[z] == batch_id) {
                            int32_t lz; LOCAL_OF(z, lz);
                            if (sts[lz]) oheap_modify(&h, lz, 1);
                        }
                    }
                }
            }
        }
        /* decrease: the vertex leaving the window */
        if (id >= window) {
            const int32_t old = (int32_t)slice[id - window];
            if (rp[old + 1] - rp[old] < degree) {
                for (int32_t a = rp[old]; a < rp[old + 1]; ++a) {
                    const int32_t y = ci[a];
                    if (tp[y + 1] - tp[y] < degree) {
                        for (int32_t c = tp[y]; c < tp[y + 1]; ++c) {
                            const int32_t z = ti[c];
                            if (batch_of[z] == batch_id) {
                                int32_t lz; LOCAL_OF(z, lz);
                                if (sts[lz]) oheap_modify(&h, lz, -1);
                            }
                        }
                    }
                }
            }
        }
    }
#undef LOCAL_OF
    oheap_free(&h);
    free(local); free(sts); free(pos); free(sorted);
}

int oracle_pgorder(int32_t n, const int32_t* rp, const int32_t* ci, int32_t W,
                   int32_t window, uint32_t* placement_gpu, uint32_t* placement_final,
                   int32_t* perm)
{
    int32_t* tp = (int32_t*)malloc(sizeof(int32_t) * ((size_t)n + 1));
    int32_t* ti = (int32_t*)malloc(sizeof(int32_t) * ((size_t)rp[n] + 1));
    oracle_transpose(n, n, rp, ci, tp, ti);
    const int32_t B = oracle_pgorder_gpu_loop(n, rp, ci, tp, ti, W, placement_gpu);
    memcpy(placement_final, placement_gpu, sizeof(uint32_t) * (size_t)n);
    const int32_t cap = B + 1;
    int32_t *Ps = malloc(sizeof(int32_t) * cap), *Pl = malloc(sizeof(int32_t) * cap);
    int32_t *Os = malloc(sizeof(int32_t) * cap), *Ol = malloc(sizeof(int32_t) * cap);
    oracle_pgorder_schedule(n, W, cap, Ps, Pl, Os, Ol);
    int32_t* batch_of = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
    for (int32_t v = 0; v < n; ++v) batch_of[v] = -1;
    const int32_t L = (int32_t)(sqrt((double)n) / 4);
    const int32_t first = (int32_t)log2((double)W);
    for (int32_t b = first; b < B; ++b)
        oracle_order_on_cpu(n, rp, ci, tp, ti, batch_of, &placement_final[Os[b]], b,
                            Ol[b], window, L);
    for (int32_t i = 0; i < n; ++i) perm[placement_final[i]] = i;
    free(tp); free(ti); free(Ps); free(Pl); free(Os); free(Ol); free(batch_of);
    return B;
}

/* ------------------------------------------------------------------------ */
int oracle_gorder1(int32_t n, const int32_t* rp, const int32_t* ci, int32_t window,
                   uint32_t* order, int32_t* perm)
{
    if (n <= 0) return 0;
    int32_t* tp = (int32_t*)malloc(sizeof(int32_t) * ((size_t)n + 1));
    int32_t* ti = (int32_t*)malloc(sizeof(int32_t) * ((size_t)rp[n] + 1));
    oracle_transpose(n, n, rp, ci, tp, ti);
    const int32_t L = (int32_t)(sqrt((double)n) / 4);
    /* start vertex: first maximum of the SpMV row length (offset2) */
    int32_t start = 0;
    for (int32_t v = 1; v < n; ++v)
        if (rp[v + 1] - rp[v] > rp[start + 1] - rp[start]) start = v;
    uint8_t* sts = (uint8_t*)malloc((size_t)n);
    oheap h;
    oheap_init(&h, n, (int64_t)rp[n] + 64);
    for (int32_t v = 0; v < n; ++v) { sts[v] = 1; oheap_push(&h, v); }
    oheap_modify(&h, start, 1);
    for (int32_t id = 0; h.cnt > 0; ++id) {
        int32_t node = oheap_pop(&h);
        sts[node] = 0;
        order[id] = (uint32_t)node;
        for (int pass = 0; pass < 2; ++pass) {
            int32_t delta;
            if (pass == 0) { delta = 1; }
            else {
                if (id < window) break;
                node = (int32_t)order[id - window];
                delta = -1;
            }
            /* neighbours through the out-list of node */
            if (tp[node + 1] - tp[node] < L)
                for (int32_t c = tp[node]; c < tp[node + 1]; ++c)
                    if (sts[ti[c]]) oheap_modify(&h, ti[c], delta);
            /* in-neighbours y, and siblings through y's out-list */
            for (int32_t a = rp[node]; a < rp[node + 1]; ++a) {
                const int32_t y = ci[a];
                if (sts[y]) oheap_modify(&h, y, delta);
                if (tp[y + 1] - tp[y] < L)
                    for (int32_t c = tp[y]; c < tp[y + 1]; ++c)
                        if (sts[ti[c]]) oheap_modify(&h, ti[c], delta);
            }
        }
    }
    for (int32_t i = 0; i < n; ++i) perm[order[i]] = i;
    oheap_free(&h);
    free(tp); free(ti); free(sts);
    return 0;
}

int64_t oracle_gorder1_score(int32_t n, const int32_t* rp, const int32_t* ci,
                             const int32_t* tp, const int32_t* ti, const uint32_t* order,
                             int32_t id, int32_t window, int32_t z)
{
    /* score(z) after `id` pops = sum over window members u = order[id-window..id)
     * of: [collen(u) < L] * #(u -> z out-edges) + #(y in cols(u) with y == z)
     *     + sum_{y in cols(u), collen(y) < L} #(z in out-list of y);
     * plus 1 for the start vertex (never undone). */
    const int32_t L = (int32_t)(sqrt((double)n) / 4);
    int64_t s = 0;
    int32_t lo = id - window; if (lo < 0) lo = 0;
    for (int32_t w = lo; w < id; ++w) {
        const int32_t u = (int32_t)order[w];
        if (tp[u + 1] - tp[u] < L)
            for (int32_t c = tp[u]; c < tp[u + 1]; ++c) s += (ti[c] == z);
        for (int32_t a = rp[u]; a < rp[u + 1]; ++a) {
            const int32_t y = ci[a];
            s += (y == z);
            if (tp[y + 1] - tp[y] < L)
                for (int32_t c = tp[y]; c < tp[y + 1]; ++c) s += (ti[c] == z);
        }
    }
    return s;
}

/* ------------------------------------------------------------------------ */
/* GRAPH::gorder5 (src/class.cu:2727-2873): fixed batches of vsize; the first
 * scores are the column counts of M (gorder_gpu_cpu_kernel1 over offset1,
 * :2789); each batch places the first len vertices by (score desc, id asc)
 * (gorder_gpu_cpu_kernel2, :2371-2389), then counts 2-hop paths from the
 * window of the last width1 batches with no hub caps (gorder_gpu_cpu_kernel3,
 * :2392-2420); the CPU phase runs order_on_cpu with width2 on every batch
 * (:2833-2850). */
int oracle_gorder5(int32_t n, const int32_t* rp, const int32_t* ci, int32_t vsize, int32_t width1,
                   int32_t width2, uint32_t* placement_gpu, uint32_t* placement_final, int32_t* perm)
{
    if (n <= 0) return 0;
    int32_t* tp = (int32_t*)malloc(sizeof(int32_t) * ((size_t)n + 1));
    int32_t* ti = (int32_t*)malloc(sizeof(int32_t) * ((size_t)rp[n] + 1));
    oracle_transpose(n, n, rp, ci, tp, ti);
    const int32_t B = (n + vsize - 1) / vsize;
    uint32_t* score = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)n);
    skey* keys = (skey*)malloc(sizeof(skey) * (size_t)n);
    for (int32_t v = 0; v < n; ++v) score[v] = (uint32_t)(tp[v + 1] - tp[v]) | 0x80000000u;
    for (int32_t b = 0; b < B; ++b) {
        const int32_t start = b * vsize;
        const int32_t len = (start + vsize < n) ? vsize : n - start;
        const int32_t os = (b < width1) ? 0 : (b - width1 + 1) * vsize;
        int32_t ol = (b < width1) ? (b + 1) * vsize : width1 * vsize;
        if (os + ol > n) ol = n - os;
        for (int32_t v = 0; v < n; ++v) { keys[v].score = score[v]; keys[v].id = (uint32_t)v; }
        qsort(keys, (size_t)n, sizeof(skey), cmp_desc_stable);
        for (int32_t v = 0; v < n; ++v) score[v] &= 0x80000000u;
        for (int32_t i = 0; i < len; ++i) {
            score[keys[i].id] = 0u;
            placement_gpu[start + i] = keys[i].id;
        }
        for (int32_t w = 0; w < ol; ++w) {
            const int32_t x = (int32_t)placement_gpu[os + w];
            for (int32_t a = rp[x]; a < rp[x + 1]; ++a) {
                const int32_t y = ci[a];
                for (int32_t c = tp[y]; c < tp[y + 1]; ++c) {
                    const int32_t z = ti[c];
                    if (score[z] >> 31) score[z] += 1u;
                }
            }
        }
    }
    memcpy(placement_final, placement_gpu, sizeof(uint32_t) * (size_t)n);
    int32_t* batch_of = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
    for (int32_t v = 0; v < n; ++v) batch_of[v] = -1;
    const int32_t L = (int32_t)(sqrt((double)n) / 4);
    for (int32_t b = 0; b < B; ++b) {
        const int32_t start = b * vsize;
        const int32_t len = (start + vsize < n) ? vsize : n - start;
        oracle_order_on_cpu(n, rp, ci, tp, ti, batch_of, &placement_final[start], b, len, width2, L);
    }
    for (int32_t i = 0; i < n; ++i) perm[placement_final[i]] = i;
    free(tp); free(ti); free(score); free(keys); free(batch_of);
    return B;
}

/* ------------------------------------------------------------------------ */
/* GRAPH::order (src/class.cu:3158-3365): the tile-aware sequential order.
 * curr counts the x entries (nonzeros of placed rows) laid out so far; a
 * tile holds width1 of them.  Placing a row whose entries fit in the current
 * tile adds one to every unplaced row sharing a column with it (columns of
 * width2 or more entries skipped); a row that crosses into the next tile first
 * resets (refresh) every row the old tile had raised, sets PMAX = 0, moves the
 * tile start to the tile holding its last entry, and raises rows only for its
 * entries inside that tile.  Start: first vertex with the longest row, raised
 * by one. */
int oracle_tile_order(int32_t n, const int32_t* rp, const int32_t* ci, int32_t width1, int32_t width2,
                      uint32_t* order, int32_t* perm)
{
    if (n <= 0) return 0;
    int32_t* tp = (int32_t*)malloc(sizeof(int32_t) * ((size_t)n + 1));
    int32_t* ti = (int32_t*)malloc(sizeof(int32_t) * ((size_t)rp[n] + 1));
    oracle_transpose(n, n, rp, ci, tp, ti);
    int32_t id0 = 0;
    for (int32_t v = 1; v < n; ++v)
        if (rp[v + 1] - rp[v] > rp[id0 + 1] - rp[id0]) id0 = v;
    uint8_t* sts = (uint8_t*)malloc((size_t)n);
    int32_t* tile = (int32_t*)malloc(sizeof(int32_t) * (size_t)(width1 > 0 ? width1 : 1));
    oheap h;
    oheap_init(&h, n, (int64_t)width1 + 64);
    for (int32_t v = 0; v < n; ++v) { sts[v] = 1; oheap_push(&h, v); }
    oheap_modify(&h, id0, 1);
    int64_t start = 0, curr = 0;
    for (int32_t id = 0; h.cnt > 0; ++id) {
        const int32_t node = oheap_pop(&h);
        sts[node] = 0;
        order[id] = (uint32_t)node;
        const int32_t deg = rp[node + 1] - rp[node];
        if (curr - start + deg < width1) {
            for (int32_t i = 0; i < deg; ++i) {
                const int32_t idx = ci[rp[node] + i];
                tile[curr - start] = idx;
                if (tp[idx + 1] - tp[idx] < width2)
                    for (int32_t j = tp[idx]; j < tp[idx + 1]; ++j)
                        if (sts[ti[j]]) oheap_modify(&h, ti[j], 1);
                ++curr;
            }
        } else {
            for (int64_t i = start; i < curr; ++i) {
                const int32_t idx = tile[i - start];
                if (tp[idx + 1] - tp[idx] < width2)
                    for (int32_t j = tp[idx]; j < tp[idx + 1]; ++j)
                        if (sts[ti[j]]) oheap_refresh(&h, ti[j]);
            }
            h.pmax = 0;
            start = (curr + deg) / width1 * width1;
            for (int32_t i = 0; i < deg; ++i) {
                if (curr >= start) {
                    const int32_t idx = ci[rp[node] + i];
                    tile[curr - start] = idx;
                    if (tp[idx + 1] - tp[idx] < width2)
                        for (int32_t j = tp[idx]; j < tp[idx + 1]; ++j)
                            if (sts[ti[j]]) oheap_modify(&h, ti[j], 1);
                }
                ++curr;
            }
        }
    }
    for (int32_t i = 0; i < n; ++i) perm[order[i]] = i;
    oheap_free(&h);
    free(tp); free(ti); free(sts); free(tile);
    return 0;
}
