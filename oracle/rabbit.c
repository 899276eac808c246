/*
 * rabbit.c -- oracle restatement of Rabbit Order (rabbitorder/src/rabbit_order.cpp,
 * rabbitorder/include/utils.h) executed in one deterministic order.
 * TEST INFRASTRUCTURE ONLY.
 *
 *   graph      make_adj_list (rabbit_order.cpp:12-81): symmetrised pattern of M,
 *              self loops dropped, parallel edges merged with summed weight;
 *              strength = sum of edge weights; total = sum of strengths
 *              (utils.h graph ctor)
 *   order      merge_order (:103-115): ascending degree; ties by lower id (the
 *              reference's parallel sort leaves ties unspecified)
 *   aggregate  merge / unite / find_best (:137-311) with one thread: gain
 *              w(v,u) - (float)(str v * str u) / total, first strict maximum
 *              over the neighbour communities in id order, merge when > 0
 *   number     compute_perm + descendants (:380-420, utils.h:347-352):
 *              depth-first over the child / sibling links, top-level
 *              communities in the order they arose
 *
 * Parity unpinned against the reference itself: its aggregation runs on
 * OpenMP threads with compare-and-swap and its result depends on timing, and
 * it needs Boost (absent here), so it cannot be built or run.  Pinned instead
 * by hand-worked graphs (path of 3, star, two cliques joined by one edge,
 * disjoint components stay contiguous) and by the invariant that every merge
 * had a positive modularity gain (tests/test_oracle_rabbit.py).
 */
#include "oracle.h"
#include <stdlib.h>
#include <string.h>

#define RB_NONE 0xffffffffu

typedef struct { uint32_t c; float w; } rb_edge;

typedef struct {
    rb_edge** es;   int64_t* ne;   int64_t* cap;
    float* str;     uint32_t* coms; uint32_t* child; uint32_t* sib; uint32_t* united;
    double total;
} rb_graph;

static void rb_push(rb_edge** a, int64_t* n, int64_t* cap, uint32_t c, float w)
{
    if (*n == *cap) {
        *cap = *cap ? *cap * 2 : 8;
        *a = (rb_edge*)realloc(*a, sizeof(rb_edge) * (size_t)*cap);
    }
    (*a)[*n].c = c;
    (*a)[*n].w = w;
    ++*n;
}

static int rb_cmp(const void* x, const void* y)
{
    const rb_edge* a = (const rb_edge*)x;
    const rb_edge* b = (const rb_edge*)y;
    return (a->c > b->c) - (a->c < b->c);
}

/* sort by community id, add up the weights of equal ids (in sorted order) */
static int64_t rb_compact(rb_edge* a, int64_t n)
{
    int64_t m = 0;
    if (n == 0) return 0;
    qsort(a, (size_t)n, sizeof(rb_edge), rb_cmp);
    for (int64_t i = 0; i < n; ++i) {
        if (m > 0 && a[m - 1].c == a[i].c) a[m - 1].w += a[i].w;
        else a[m++] = a[i];
    }
    return m;
}

static uint32_t rb_trace(rb_graph* g, uint32_t v)
{
    uint32_t c = v;
    while (g->coms[c] != c) c = g->coms[c];
    if (v != c && g->coms[v] != c) g->coms[v] = c;
    return c;
}

static uint32_t rb_merge(rb_graph* g, uint32_t v, rb_edge** nb, int64_t* nn, int64_t* ncap)
{
    *nn = 0;
    /* own edges, then those of every member merged since the last unite */
    for (int64_t i = 0; i < g->ne[v]; ++i) {
        uint32_t c = rb_trace(g, g->es[v][i].c);
        if (c != v) rb_push(nb, nn, ncap, c, g->es[v][i].w);
    }
    while (g->united[v] != g->child[v]) {
        uint32_t c = g->child[v], w;
        for (w = c; w != RB_NONE && w != g->united[v]; w = g->sib[w])
            for (int64_t i = 0; i < g->ne[w]; ++i) {
                uint32_t cc = rb_trace(g, g->es[w][i].c);
                if (cc != v) rb_push(nb, nn, ncap, cc, g->es[w][i].w);
            }
        g->united[v] = c;
    }
    *nn = rb_compact(*nb, *nn);
    g->ne[v] = 0;
    for (int64_t i = 0; i < *nn; ++i) rb_push(&g->es[v], &g->ne[v], &g->cap[v], (*nb)[i].c, (*nb)[i].w);

    float vstr = g->str[v];
    g->str[v] = -1.0f;
    double dmax = 0.0;
    uint32_t best = v;
    for (int64_t i = 0; i < g->ne[v]; ++i) {
        const rb_edge e = g->es[v][i];
        float prod = vstr * g->str[e.c];
        double d = (double)e.w - (double)prod / g->total;
        if (dmax < d) { dmax = d; best = e.c; }
    }
    if (best == v) { g->str[v] = vstr; return v; }
    g->sib[v] = g->child[best];
    g->str[best] += vstr;
    g->child[best] = v;
    g->coms[v] = best;
    return best;
}

static int64_t* g_deg_for_sort;
static int rb_ord_cmp(const void* x, const void* y)
{
    uint32_t a = *(const uint32_t*)x, b = *(const uint32_t*)y;
    if (g_deg_for_sort[a] != g_deg_for_sort[b]) return g_deg_for_sort[a] < g_deg_for_sort[b] ? -1 : 1;
    return (a > b) - (a < b);
}

int oracle_rabbit(int32_t n, const int32_t* rp, const int32_t* ci, uint32_t* new_id)
{
    rb_graph g;
    uint32_t *ord, *tops, *stack;
    int64_t ntops = 0, sp = 0;
    rb_edge* nb = NULL;
    int64_t nn = 0, ncap = 0;
    uint32_t next = 0;
    if (n <= 0) return 0;
    memset(&g, 0, sizeof(g));
    g.es = (rb_edge**)calloc((size_t)n, sizeof(rb_edge*));
    g.ne = (int64_t*)calloc((size_t)n, sizeof(int64_t));
    g.cap = (int64_t*)calloc((size_t)n, sizeof(int64_t));
    g.str = (float*)calloc((size_t)n, sizeof(float));
    g.coms = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)n);
    g.child = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)n);
    g.sib = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)n);
    g.united = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)n);
    for (int32_t i = 0; i < n; ++i)
        for (int32_t q = rp[i]; q < rp[i + 1]; ++q) {
            int32_t j = ci[q];
            if (j == i) continue;
            rb_push(&g.es[i], &g.ne[i], &g.cap[i], (uint32_t)j, 1.0f);
            rb_push(&g.es[j], &g.ne[j], &g.cap[j], (uint32_t)i, 1.0f);
        }
    g.total = 0.0;
    for (int32_t v = 0; v < n; ++v) {
        float s = 0.0f;
        g.ne[v] = rb_compact(g.es[v], g.ne[v]);
        for (int64_t i = 0; i < g.ne[v]; ++i) s += g.es[v][i].w;
        g.str[v] = s;
        g.total += s;
        g.coms[v] = (uint32_t)v;
        g.child[v] = g.sib[v] = g.united[v] = RB_NONE;
    }
    ord = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)n);
    g_deg_for_sort = (int64_t*)malloc(sizeof(int64_t) * (size_t)n);
    for (int32_t v = 0; v < n; ++v) { ord[v] = (uint32_t)v; g_deg_for_sort[v] = g.ne[v]; }
    qsort(ord, (size_t)n, sizeof(uint32_t), rb_ord_cmp);
    free(g_deg_for_sort);
    tops = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)n);
    for (int32_t i = 0; i < n; ++i)
        if (rb_merge(&g, ord[i], &nb, &nn, &ncap) == ord[i]) tops[ntops++] = ord[i];
    stack = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)n + 4);
    for (int64_t t = 0; t < ntops; ++t) {
        uint32_t v = tops[t];
        stack[sp++] = v;
        while ((v = g.child[v]) != RB_NONE) stack[sp++] = v;
        while (sp > 0) {
            uint32_t u = stack[--sp];
            new_id[u] = next++;
            if (g.sib[u] != RB_NONE) {
                uint32_t w = g.sib[u];
                stack[sp++] = w;
                while ((w = g.child[w]) != RB_NONE) stack[sp++] = w;
            }
        }
    }
    for (int32_t v = 0; v < n; ++v) free(g.es[v]);
    free(g.es); free(g.ne); free(g.cap); free(g.str); free(g.coms); free(g.child); free(g.sib); free(g.united);
    free(ord); free(tops); free(stack); free(nb);
    return (int)ntops;
}
