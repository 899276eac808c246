/*
 * oracle.h -- CPU restatement of the reference SpMV locality pipeline.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP product
 * library (libmr.so): only tests/, __graft_entry__.smoke() and the cpu_baseline
 * leg of bench.py may load it.  Nothing in the product links, loads or calls it.
 *
 * Every function restates one reference function (cited file:line, paths
 * relative to the reference repository root) in plain C.  Index convention:
 * the SpMV matrix M is a standard CSR (rp, ci).  The reference multiplies the
 * CSR whose row pointer is built over its COO col_ind and whose column indices
 * are its row_ind (src/class.cu:180,192,1342,1372), so
 *     reference offset2/indice2  == (rp, ci)    rows of M ("in-lists")
 *     reference offset1/indice1  == (tp, ti)    transpose of M ("out-lists")
 *
 * Pinning status of each function (what checks it besides restatement) is in
 * the header comment of the .c file that defines it.
 */
#ifndef MR_ORACLE_H
#define MR_ORACLE_H
#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- integer bucket heap: include/integer_heap.cuh ---------------------- */
typedef struct oheap {
    int32_t* htab;      /* bucket heads, -1 = empty                          */
    int64_t  htab_size;
    int32_t  pmax;
    int64_t  cnt;
    int32_t* key1;      /* per node                                          */
    int32_t* key2;
    int32_t* left;
    int32_t* right;
    int32_t  n_nodes;
} oheap;

int  oheap_init(oheap* h, int32_t n_nodes, int64_t buckets);
void oheap_free(oheap* h);
void oheap_push(oheap* h, int32_t v);
void oheap_modify(oheap* h, int32_t v, int32_t val);
int32_t oheap_pop(oheap* h);
void oheap_refresh(oheap* h, int32_t v);
/* scripted trace driver used by the golden-trace tests:
 * ops[i] = {code, node, val}: code 0 push(node), 1 modify(node,val), 2 pop.
 * Popped ids are written to out (returns count). */
int64_t oracle_heap_trace(int32_t n_nodes, int64_t buckets, const int32_t* ops,
                          int64_t n_ops, int32_t* out);

/* ---- sizes, transpose, permutation -------------------------------------- */
int32_t oracle_get_width(int64_t n, int64_t nnz);                 /* class.cu:733-750 */
void oracle_transpose(int32_t n_rows, int32_t n_cols, const int32_t* rp,
                      const int32_t* ci, int32_t* tp, int32_t* ti);
int  oracle_check_perm(const int32_t* perm, int64_t n);            /* utils.cuh:5-55 */
/* GRAPH::gorder5 (class.cu:2727-2873); returns the batch count */
int oracle_gorder5(int32_t n, const int32_t* rp, const int32_t* ci, int32_t vsize, int32_t width1,
                   int32_t width2, uint32_t* placement_gpu, uint32_t* placement_final, int32_t* perm);
/* GRAPH::order (class.cu:3158-3365): order[pos] = vertex, perm[vertex] = pos */
int oracle_tile_order(int32_t n, const int32_t* rp, const int32_t* ci, int32_t width1, int32_t width2,
                      uint32_t* order, int32_t* perm);
/* degree order and DBG (class.cu:1496-1549, 1683-1756): perm[old] = new */
void oracle_order_degree(int32_t n, const int32_t* rp, const int32_t* ci, int32_t* perm);
void oracle_order_dbg(int32_t n, const int32_t* rp, const int32_t* ci, int32_t* perm);
/* relabel: perm[old]=new; A' = P A P^T, columns sorted (class.cu:3099-3116 + 423-457) */
void oracle_permute(int32_t n, const int32_t* rp, const int32_t* ci, const double* val,
                    const int32_t* perm, int32_t* rp2, int32_t* ci2, double* val2);

/* ---- Gorder family ------------------------------------------------------- */
/* pGOrder batch schedule (class.cu:2995-3016). Returns number of batches and
 * fills up to cap entries of Pstart/Pleng/Ostart/Oleng. */
int32_t oracle_pgorder_schedule(int32_t n, int32_t width, int32_t cap, int32_t* ps,
                                int32_t* pl, int32_t* os, int32_t* ol);
/* GPU greedy-placement loop of gorder6 (class.cu:3041-3059), serial.  Writes the
 * placement order (position -> vertex) before the CPU refinement. */
int  oracle_pgorder_gpu_loop(int32_t n, const int32_t* rp, const int32_t* ci,
                             const int32_t* tp, const int32_t* ti, int32_t width,
                             uint32_t* placement);
/* the same loop with the hub cap of kernel3 (10,000 in the reference) as a
 * parameter, stopped after max_batches batches (-1 = all); returns the number
 * of batches run */
int  oracle_pgorder_gpu_loop_ex(int32_t n, const int32_t* rp, const int32_t* ci,
                                const int32_t* tp, const int32_t* ti, int32_t width, int32_t hub_cap,
                                int32_t max_batches, uint32_t* placement);
/* order_on_cpu (class.cu:2220-2314) on one slice; batch_of[v] plays Node::batch_id. */
void oracle_order_on_cpu(int32_t n, const int32_t* rp, const int32_t* ci,
                         const int32_t* tp, const int32_t* ti, int32_t* batch_of,
                         uint32_t* slice, int32_t batch_id, int32_t len,
                         int32_t window, int32_t degree);
/* full gorder6: GPU loop + CPU phase + perm[old]=new (class.cu:2955-3140) */
int  oracle_pgorder(int32_t n, const int32_t* rp, const int32_t* ci, int32_t width,
                    int32_t window, uint32_t* placement_gpu, uint32_t* placement_final,
                    int32_t* perm);
/* gorder1: classic sequential Gorder (class.cu:1825-2013); perm[old]=new */
int  oracle_gorder1(int32_t n, const int32_t* rp, const int32_t* ci, int32_t window,
                    uint32_t* order, int32_t* perm);
/* naive Gorder score of vertex z given placed order[0..id) (definition used by
 * the max-score invariant test of gorder1) */
int64_t oracle_gorder1_score(int32_t n, const int32_t* rp, const int32_t* ci,
                             const int32_t* tp, const int32_t* ti, const uint32_t* order,
                             int32_t id, int32_t window, int32_t z);

/* ---- RCM: rcm/src/rcm.cpp:1-154, rcm/src/queue.cpp:39-55 --------------- */
/* rows/cols: COO arrays as passed by class.cu:1570 (symmetrised internally). */
int  oracle_rcm(int32_t n, int64_t nz, const int32_t* rows, const int32_t* cols,
                uint32_t* permutation);

/* ---- reuse distance: class.cu:863-1261 ----------------------------------- */
int  oracle_rd_exact(const int32_t* stream, int32_t n, int64_t nnz, int32_t sector,
                     const int32_t* edges, int32_t n_edges, int64_t* counts,
                     int64_t* reuses, double* prob);
int  oracle_rd_bruteforce(const int32_t* stream, int64_t nnz, int32_t sector,
                          int64_t* rd_out);   /* -1 = cold */
uint32_t oracle_hll_hash(uint32_t x);
/* avalanche bias of oracle_hll_hash over 2^log2_n sampled inputs (checker) */
double oracle_hll_hash_bias(int32_t log2_n, uint64_t seed);
int32_t  oracle_hll_rank(uint32_t h, int32_t p);
float    oracle_hll_estimate(const int32_t* regs, int32_t m);
/* one sample of sampled_hll_rd_block_kernel (class.cu:1014-1131) */
void oracle_hll_sample(const int32_t* stream, int64_t nnz, int32_t sid, int32_t sector,
                       uint32_t seed, int32_t rd_cap, int32_t max_scan, int32_t p,
                       float* est, int32_t* status, int32_t* pos);
uint32_t oracle_hll_seed(int32_t graph_id, int32_t stream_type);
/* sampled_hll_rd_block_gpu host part (class.cu:1133-1261) */
int  oracle_rd_sampled(const int32_t* stream, int64_t nnz, int32_t samples, int32_t sector,
                       int32_t rd_cap, int32_t hll_p, int32_t max_scan, uint32_t seed,
                       const int32_t* edges, int32_t n_edges, float* est, int32_t* status,
                       int32_t* pos, int64_t* counts, int32_t* stat4, double* prob);

/* ---- predictor: include/model.cuh, ridge/out/{merbit,csr}/sampled ------- */
double oracle_predict(const double* feat, int32_t model);  /* 0 merbit, 1 csr */
double oracle_predict_preprocessed(const double* feat, const uint8_t* missing, int32_t model);

/* ---- SpMV ---------------------------------------------------------------- */
/* y = A x in FP64 with compensated (Neumaier) row sums; absrow[i] = sum|a_ij x_j| */
void oracle_spmv_f64(int32_t n_rows, const int32_t* rp, const int32_t* ci,
                     const double* val, const double* x, double* y, double* absrow);
/* k right-hand sides, row-major X[n_cols][k], Y[n_rows][k] */
void oracle_spmm_f64(int32_t n_rows, int32_t k, const int32_t* rp, const int32_t* ci,
                     const double* val, const double* x, double* y);
/* plain threaded CSR SpMV (CPU baseline): nnz-balanced static row partition */
double oracle_spmv_f32_threads(int32_t n_rows, const int32_t* rp, const int32_t* ci,
                               const float* val, const float* x, float* y,
                               int32_t threads, int32_t iters);
double oracle_spmv_f64_threads(int32_t n_rows, const int32_t* rp, const int32_t* ci,
                               const double* val, const double* x, double* y,
                               int32_t threads, int32_t iters);

/* ---- MERBIT (include/merbit.cuh) ----------------------------------------- */
/* reference 32-lane tile metadata, u32 descriptors (SIGMA < 15) */
int  oracle_merbit_ref_tiles(int32_t n, int64_t nnz, const int32_t* rp, int32_t sigma,
                             uint32_t* tile_x, uint32_t* tile_y, uint32_t* lane_desc);
void oracle_merge_search(const int32_t* rp, int32_t n, int64_t nnz, int64_t diag,
                         int64_t* cx, int64_t* cy);
/* sequential emulation of merbit_kernel's merge-path SpMV (fp64) */
void oracle_merbit_spmv(int32_t n, int64_t nnz, const int32_t* rp, const int32_t* ci,
                        const double* val, const double* x, int32_t sigma, double* y);

/* ---- Rabbit Order, one deterministic order (rabbitorder/src/rabbit_order.cpp) -- */
/* new_id[v] = position of v; returns the number of top-level communities */
int oracle_rabbit(int32_t n, const int32_t* rp, const int32_t* ci, uint32_t* new_id);

#ifdef __cplusplus
}
#endif
#endif
