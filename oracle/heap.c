/*
 * heap.c -- oracle restatement of the Gorder integer bucket heap
 * (include/integer_heap.cuh:1-155).  TEST INFRASTRUCTURE ONLY.
 *
 * Nodes are integers; the intrusive doubly linked circular bucket lists of the
 * reference are kept as left/right index arrays and the bucket table as head
 * indices (-1 = nullptr).  Semantics kept exactly:
 *   _insert_  (integer_heap.cuh:56-75): key1 += key2, key2 = 0; an empty bucket
 *             takes the node as head, otherwise the node goes right after head.
 *   _remove_  (integer_heap.cuh:31-54): removing the head promotes head->right.
 *   push      (:81-89), modify (:91-102: positive pending delta re-buckets at
 *             once, non-positive stays lazy), pop (:114-145: a head with a
 *             pending delta is re-inserted lower and the scan continues).
 * The reference table has a fixed size; this one grows on demand, which is the
 * same behaviour wherever the reference stays in bounds.
 *
 * Pinned by: golden push/modify/pop traces written by the reference header
 * itself compiled into oracle/_ref (tests/golden/heap_traces.json).
 */
#include "oracle.h"
#include <stdlib.h>
#include <string.h>

int oheap_init(oheap* h, int32_t n_nodes, int64_t buckets)
{
    memset(h, 0, sizeof(*h));
    if (buckets < 16) buckets = 16;
    h->htab = (int32_t*)malloc(sizeof(int32_t) * (size_t)buckets);
    h->key1 = (int32_t*)calloc((size_t)n_nodes + 1, sizeof(int32_t));
    h->key2 = (int32_t*)calloc((size_t)n_nodes + 1, sizeof(int32_t));
    h->left = (int32_t*)malloc(sizeof(int32_t) * ((size_t)n_nodes + 1));
    h->right = (int32_t*)malloc(sizeof(int32_t) * ((size_t)n_nodes + 1));
    if (!h->htab || !h->key1 || !h->key2 || !h->left || !h->right) return -1;
    for (int64_t i = 0; i < buckets; ++i) h->htab[i] = -1;
    for (int32_t i = 0; i <= n_nodes; ++i) { h->left[i] = i; h->right[i] = i; }
    h->htab_size = buckets;
    h->n_nodes = n_nodes;
    return 0;
}

void oheap_free(oheap* h)
{
    free(h->htab); free(h->key1); free(h->key2); free(h->left); free(h->right);
    memset(h, 0, sizeof(*h));
}

static void grow(oheap* h, int64_t need)
{
    if (need < h->htab_size) return;
    int64_t ns = h->htab_size;
    while (ns <= need) ns *= 2;
    h->htab = (int32_t*)realloc(h->htab, sizeof(int32_t) * (size_t)ns);
    for (int64_t i = h->htab_size; i < ns; ++i) h->htab[i] = -1;
    h->htab_size = ns;
}

static void unlink_node(oheap* h, int32_t v)
{
    int32_t l = h->left[v], r = h->right[v];
    h->right[l] = r;
    h->left[r] = l;
    h->left[v] = v;
    h->right[v] = v;
}

static void hremove(oheap* h, int32_t v)
{
    int32_t* slot = &h->htab[h->key1[v]];
    if (*slot == v) {
        if (h->right[v] == v) {
            *slot = -1;
        } else {
            *slot = h->right[v];
            unlink_node(h, v);
        }
    } else {
        unlink_node(h, v);
    }
}

static void hinsert(oheap* h, int32_t v)
{
    h->key1[v] += h->key2[v];
    h->key2[v] = 0;
    grow(h, h->key1[v]);
    int32_t* slot = &h->htab[h->key1[v]];
    if (*slot < 0) {
        *slot = v;
    } else {
        int32_t hd = *slot;
        h->left[v] = hd;
        h->right[v] = h->right[hd];
        h->left[h->right[hd]] = v;
        h->right[hd] = v;
    }
}

void oheap_push(oheap* h, int32_t v)
{
    hinsert(h, v);
    ++h->cnt;
    if (h->key1[v] > h->pmax) h->pmax = h->key1[v];
}

void oheap_modify(oheap* h, int32_t v, int32_t val)
{
    h->key2[v] += val;
    if (h->key2[v] > 0) {
        hremove(h, v);
        hinsert(h, v);
        if (h->key1[v] > h->pmax) h->pmax = h->key1[v];
    }
}

/* Integer_Heap::refresh (integer_heap.cuh:112-120): a node with key1 > 0 is
 * taken out and re-inserted with key1 = 0 (+ its pending key2) */
void oheap_refresh(oheap* h, int32_t v)
{
    if (h->key1[v] > 0) {
        hremove(h, v);
        h->key1[v] = 0;
        hinsert(h, v);
    }
}

int32_t oheap_pop(oheap* h)
{
    while (h->cnt > 0) {
        while (h->pmax > 0 && h->htab[h->pmax] < 0) --h->pmax;
        int32_t t = h->htab[h->pmax];
        while (t >= 0) {
            if (h->key2[t] == 0) {
                hremove(h, t);
                --h->cnt;
                return t;
            }
            hremove(h, t);
            hinsert(h, t);
            t = h->htab[h->pmax];
        }
        if (h->pmax == 0 && h->htab[0] < 0) break;  /* reference would spin here */
    }
    return -1;
}

int64_t oracle_heap_trace(int32_t n_nodes, int64_t buckets, const int32_t* ops,
                          int64_t n_ops, int32_t* out)
{
    oheap h;
    if (oheap_init(&h, n_nodes, buckets) != 0) return -1;
    int64_t k = 0;
    for (int64_t i = 0; i < n_ops; ++i) {
        const int32_t code = ops[3 * i], node = ops[3 * i + 1], val = ops[3 * i + 2];
        if (code == 0) oheap_push(&h, node);
        else if (code == 1) oheap_modify(&h, node, val);
        else if (code == 2) out[k++] = oheap_pop(&h);
        else if (code == 3) oheap_refresh(&h, node);
        else h.pmax = val;                /* code 4: the caller sets PMAX */
    }
    oheap_free(&h);
    return k;
}
