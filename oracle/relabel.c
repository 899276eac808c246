/*
 * relabel.c -- oracle restatement of the relabel step and the permutation check.
 * TEST INFRASTRUCTURE ONLY.
 *
 *   oracle_check_perm  include/utils.cuh:5-55 (sort, adjacent difference, range)
 *   oracle_permute     src/class.cu:3099-3116: perm[placement[k]] = k is applied
 *                      to both COO index arrays and the COO is re-sorted by
 *                      (col, row) (GRAPH::sort(false), class.cu:443-455).  In CSR
 *                      terms: entry (i, j) of M moves to (perm[i], perm[j]) and
 *                      every row is sorted by column.  The reference sort is a
 *                      Thrust merge sort (stable), so equal (row, col) pairs keep
 *                      their input order; the qsort below breaks ties by input
 *                      position to give the same order.
 *
 * Pinned by: comparison with scipy.sparse A[inv][:, inv] with sorted indices
 * (tests/test_oracle_basic.py) and invertibility (perm then inverse perm).
 */
#include "oracle.h"
#include <stdlib.h>
#include <string.h>

static int cmp_i32(const void* a, const void* b)
{
    const int32_t x = *(const int32_t*)a, y = *(const int32_t*)b;
    return (x > y) - (x < y);
}

int oracle_check_perm(const int32_t* perm, int64_t n)
{
    if (!perm || n <= 0) return -1;
    int32_t* s = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
    memcpy(s, perm, sizeof(int32_t) * (size_t)n);
    qsort(s, (size_t)n, sizeof(int32_t), cmp_i32);
    int ok = (s[0] == 0) && (s[n - 1] == n - 1);
    for (int64_t i = 1; i < n && ok; ++i) ok = (s[i] != s[i - 1]);
    free(s);
    return ok ? 0 : 1;
}

typedef struct { int32_t row, col; int64_t src; } coo_t;

static int cmp_coo(const void* a, const void* b)
{
    const coo_t* x = (const coo_t*)a;
    const coo_t* y = (const coo_t*)b;
    if (x->row != y->row) return (x->row > y->row) - (x->row < y->row);
    if (x->col != y->col) return (x->col > y->col) - (x->col < y->col);
    return (x->src > y->src) - (x->src < y->src);
}

void oracle_permute(int32_t n, const int32_t* rp, const int32_t* ci, const double* val,
                    const int32_t* perm, int32_t* rp2, int32_t* ci2, double* val2)
{
    const int64_t nnz = rp[n];
    coo_t* e = (coo_t*)malloc(sizeof(coo_t) * (size_t)(nnz + 1));
    for (int32_t i = 0; i < n; ++i)
        for (int32_t k = rp[i]; k < rp[i + 1]; ++k) {
            e[k].row = perm[i];
            e[k].col = perm[ci[k]];
            e[k].src = k;
        }
    qsort(e, (size_t)nnz, sizeof(coo_t), cmp_coo);
    memset(rp2, 0, sizeof(int32_t) * ((size_t)n + 1));
    for (int64_t k = 0; k < nnz; ++k) {
        rp2[e[k].row + 1]++;
        ci2[k] = e[k].col;
        if (val && val2) val2[k] = val[e[k].src];
    }
    for (int32_t i = 0; i < n; ++i) rp2[i + 1] += rp2[i];
    free(e);
}
