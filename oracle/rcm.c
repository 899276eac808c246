/*
 * rcm.c -- oracle restatement of the reference's Reverse Cuthill-McKee CPU
 * baseline (rcm/src/rcm.cpp:1-154, rcm/src/queue.cpp:20-118).
 * TEST INFRASTRUCTURE ONLY.
 *
 * Semantics kept: degree = number of off-diagonal COO entries touching the
 * vertex (both directions, duplicates counted, rcm.cpp:2-12); neighbour lists in
 * COO order with both directions appended (rcm.cpp:46-57); the start of every
 * component is the lowest-degree vertex not yet in the permutation, first index
 * on ties (rcm.cpp:23-44); neighbours not yet queued are appended and then the
 * pending part of the queue is ordered by queueSort's exchange sort, which swaps
 * on `>=` (queue.cpp:39-55) -- its tie order is part of the result; finally the
 * order is reversed.  The reference reverses with a shared temporary inside an
 * OpenMP loop (rcm.cpp:145-153, a data race); this restates the race-free
 * reversal, which is what the reference computes with one thread.
 *
 * Pinned by: rcm/src/{rcm,queue}.cpp compiled into oracle/_ref and run with one
 * OpenMP thread on random graphs (tests/golden/rcm_ref.json).
 */
#include "oracle.h"
#include <stdlib.h>
#include <string.h>

typedef struct { int32_t num, degree; } qnode;

int oracle_rcm(int32_t n, int64_t nz, const int32_t* rows, const int32_t* cols,
               uint32_t* permutation)
{
    int32_t* degree = (int32_t*)calloc((size_t)n, sizeof(int32_t));
    for (int64_t i = 0; i < nz; ++i)
        if (rows[i] != cols[i]) { degree[rows[i]]++; degree[cols[i]]++; }
    int64_t* start = (int64_t*)malloc(sizeof(int64_t) * ((size_t)n + 1));
    start[0] = 0;
    for (int32_t v = 0; v < n; ++v) start[v + 1] = start[v] + degree[v];
    int32_t* nb = (int32_t*)malloc(sizeof(int32_t) * (size_t)(start[n] + 1));
    int64_t* fill = (int64_t*)malloc(sizeof(int64_t) * ((size_t)n + 1));
    memcpy(fill, start, sizeof(int64_t) * ((size_t)n + 1));
    for (int64_t i = 0; i < nz; ++i)
        if (rows[i] != cols[i]) {
            nb[fill[rows[i]]++] = cols[i];
            nb[fill[cols[i]]++] = rows[i];
        }
    uint8_t* in_perm = (uint8_t*)calloc((size_t)n, 1);
    uint8_t* in_q = (uint8_t*)calloc((size_t)n, 1);
    qnode* Q = (qnode*)malloc(sizeof(qnode) * ((size_t)n + 1));
    qnode* T = (qnode*)malloc(sizeof(qnode) * ((size_t)n + 1));
    int64_t qh = 0, qt = 0;
    int32_t added = 0, index = 0;

#define MIN_NODE() do {                                                         \
        int32_t mn = 0, first = 1;                                              \
        for (int32_t i = 0; i < n; ++i) {                                       \
            if (first && !in_perm[i]) { first = 0; mn = i; }                    \
            if (degree[i] < degree[mn] && !in_perm[i] && !in_perm[mn]) mn = i;  \
        }                                                                       \
        in_q[mn] = 1;                                                           \
        Q[qt++] = (qnode){mn, degree[mn]};                                      \
    } while (0)

    MIN_NODE();
    while (qh < qt) {
        const qnode e = Q[qh++];
        in_perm[e.num] = 1;
        permutation[index++] = (uint32_t)e.num;
        ++added;
        int64_t tt = 0;
        for (int64_t k = start[e.num]; k < start[e.num + 1]; ++k) {
            const int32_t w = nb[k];
            if (!in_perm[w] && !in_q[w]) { in_q[w] = 1; T[tt++] = (qnode){w, degree[w]}; }
        }
        for (int64_t i = 0; i < tt; ++i)
            for (int64_t j = i + 1; j < tt; ++j)
                if (T[i].degree >= T[j].degree) { qnode t = T[i]; T[i] = T[j]; T[j] = t; }
        for (int64_t i = 0; i < tt; ++i) Q[qt++] = T[i];
        if (qh == qt && added != n) MIN_NODE();
    }
#undef MIN_NODE
    for (int32_t i = 0; i < n / 2; ++i) {
        const uint32_t t = permutation[n - i - 1];
        permutation[n - i - 1] = permutation[i];
        permutation[i] = t;
    }
    free(degree); free(start); free(nb); free(fill); free(in_perm); free(in_q); free(Q); free(T);
    return 0;
}
