/*
 * mr.h -- C-ABI of the MI355X (gfx950) SpMV locality pipeline: load a CSR,
 * reorder it (pGOrder and the CPU baselines), predict the locality headroom
 * from reuse distances, and multiply (MERBIT merge-path SpMV, CSR scalar and
 * vector SpMV, SpMM).
 *
 * The reference has no library ABI; its only operator interface is the C++
 * member set of GRAPH<V> (include/class.cuh:59-99) plus the free functions of
 * src/class.cu and include/merbit.cuh.  Every entry point below names the
 * reference interface it replaces (paths relative to the reference root).
 *
 * Conventions
 *  - Every function returns an int status (MR_OK = 0); mr_last_error() gives a
 *    thread-local message for the last failure on the calling thread.
 *  - Index arrays are int32.  Host pointers are plain C arrays owned by the
 *    caller; "device" pointers come from mr_malloc or any allocator of the same
 *    HIP runtime.
 *  - A matrix handle owns its device copy of the CSR (row_ptr, col_idx, val)
 *    and the MERBIT metadata.  The reorder entry points relabel the handle in
 *    place, as the reference's GRAPH::gorder* / relabel_* do, and accumulate the
 *    permutation (perm[old] = new) that mr_mat_perm returns.
 *  - SpMV multiplies the standard CSR A: y[i] = sum_k val[k] * x[col_idx[k]] for
 *    k in [row_ptr[i], row_ptr[i+1]).  The reference's multiplied matrix is the
 *    transpose of its MatrixMarket COO (class.cu:180,192); pass that transpose as
 *    the CSR to reproduce it (mr_mtx_load does so).
 *  - Value storage dtypes: F64, F32, F16, BF16.  F64 matrices take F64 x/y; all
 *    other matrices take F32 x/y and accumulate in F32.
 */
#ifndef MR_H
#define MR_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes --------------------------------------------------------- */
#define MR_OK                  0
#define MR_ERR_INVALID_ARG     1   /* bad argument / size                        */
#define MR_ERR_HIP             2   /* a HIP runtime call failed                  */
#define MR_ERR_OUT_OF_MEMORY   3   /* device or host allocation failed           */
#define MR_ERR_NOT_PERMUTATION 4   /* array is not a permutation of [0, n)       */
#define MR_ERR_UNSUPPORTED     5   /* combination not supported                  */
#define MR_ERR_MALFORMED_CSR   6   /* row_ptr not monotone / columns out of range */
#define MR_ERR_IO              7   /* file could not be read / parsed            */
#define MR_ERR_ADMISSION       8   /* serving: request refused (memory budget)   */
#define MR_ERR_NO_DEVICE       9   /* no HIP device visible                      */
#define MR_ERR_NUMERIC        10   /* non-finite input or result                 */
#define MR_ERR_TIMEOUT        11   /* serving: request timed out                 */
#define MR_ERR_CANCELLED      12   /* serving: request cancelled                 */

typedef enum { MR_F64 = 0, MR_F32 = 1, MR_F16 = 2, MR_BF16 = 3 } mr_dtype;
typedef enum { MR_KERNEL_MERBIT = 0, MR_KERNEL_CSR_SCALAR = 1, MR_KERNEL_CSR_VECTOR = 2 } mr_kernel;
typedef enum { MR_H2D = 0, MR_D2H = 1, MR_D2D = 2 } mr_copy_kind;

typedef struct mr_ctx_s* mr_ctx;
typedef struct mr_mat_s* mr_mat;
typedef void* mr_stream;   /* hipStream_t; NULL = the context's stream         */
typedef void* mr_event;    /* hipEvent_t                                        */

/* ---- errors, version, context -------------------------------------------- */
/* replaces the print-and-continue CHECK_CUDA/CHECK_CUSPARSE macros
 * (include/inc.cuh:101-117) with a status + thread-local message */
const char* mr_last_error(void);
const char* mr_version(void);
int mr_device_count(int* count);
/* replaces ENV (include/class.cuh:13-36); the device id is honoured */
int mr_ctx_create(int device, mr_ctx* out);
/* destroy every server, batch and matrix of the context (and free its
 * buffers) first: they refer to it until they are destroyed */
int mr_ctx_destroy(mr_ctx ctx);
int mr_ctx_stream(mr_ctx ctx, mr_stream* out);
int mr_ctx_mem_info(mr_ctx ctx, size_t* free_bytes, size_t* total_bytes);
int mr_ctx_arch(mr_ctx ctx, char* name, int cap, int* cu_count);

/* ---- device memory, streams, events (plumbing for callers without a HIP
 * runtime of their own; no reference counterpart) -------------------------- */
int mr_malloc(mr_ctx ctx, size_t bytes, void** dptr);
int mr_free(mr_ctx ctx, void* dptr);
int mr_memcpy(mr_ctx ctx, void* dst, const void* src, size_t bytes, int kind);
int mr_memset(mr_ctx ctx, void* dptr, int value, size_t bytes);
/* device-to-device streaming copy kernel (16-byte words; bytes % 16 == 0):
 * the measured-bandwidth reference of the roofline figures */
int mr_bench_copy(mr_ctx ctx, void* dst, const void* src, size_t bytes, mr_stream s);
int mr_stream_create(mr_ctx ctx, mr_stream* out);
int mr_stream_destroy(mr_ctx ctx, mr_stream s);
int mr_stream_sync(mr_ctx ctx, mr_stream s);
int mr_event_create(mr_ctx ctx, mr_event* out);
int mr_event_destroy(mr_ctx ctx, mr_event e);
int mr_event_record(mr_ctx ctx, mr_event e, mr_stream s);
int mr_event_elapsed_ms(mr_ctx ctx, mr_event start, mr_event stop, float* ms);

/* ---- host-side input generators (configurations of the benchmark; the
 * reference reads MatrixMarket files only) ---------------------------------- */
/* R-MAT edge list, 2^scale vertices, edge_factor * 2^scale edges, counter-based
 * SplitMix64 stream per edge; scramble != 0 applies a seeded bijection to the
 * vertex ids. src/dst must hold edge_factor << scale entries. */
int mr_gen_rmat_edges(int scale, int edge_factor, double a, double b, double c,
                      uint64_t seed, int scramble, int32_t* src, int32_t* dst);
/* COO -> CSR: rows sorted, columns sorted within rows, duplicates merged when
 * dedup != 0 (self loops kept). col_idx must hold m entries; *nnz_out = nnz. */
int mr_coo_to_csr(int64_t n_rows, int64_t m, const int32_t* src, const int32_t* dst,
                  int dedup, int32_t* row_ptr, int32_t* col_idx, int64_t* nnz_out);
/* values: kind 0 = uniform[lo,hi]; kind 1 = log-uniform magnitude in [lo,hi]
 * with a random sign. Counter-based: out[i] depends only on (seed, i). */
int mr_gen_values(uint64_t seed, int kind, double lo, double hi, int64_t count, double* out);
/* 27-point stencil on an nx*ny*nz grid: row_ptr[n+1], col_idx[nnz] (query the
 * sizes with NULL arrays: *n_out, *nnz_out) */
int mr_gen_stencil27(int nx, int ny, int nz, int32_t* row_ptr, int32_t* col_idx,
                     int64_t* n_out, int64_t* nnz_out);
/* random banded matrix: each row draws nnz_per_row distinct columns in
 * [i-bw, i+bw] (diagonal always present). Query with NULL arrays. */
int mr_gen_banded(int64_t n, int bandwidth, int nnz_per_row, uint64_t seed,
                  int32_t* row_ptr, int32_t* col_idx, int64_t* nnz_out);

/* ---- matrices ------------------------------------------------------------- */
/* replaces GRAPH ctors (class.cu:84-216, 330-421): copies a host CSR in.
 * val is in val_dtype (NULL = all ones, as the reference's loader forces),
 * stored as store_dtype. Validates monotone row_ptr and column range. */
int mr_csr_create(mr_ctx ctx, int64_t n_rows, int64_t n_cols, int64_t nnz,
                  const int32_t* row_ptr, const int32_t* col_idx, const void* val,
                  int val_dtype, int store_dtype, mr_mat* out);
/* MatrixMarket loader (BASEGRAPH ctor, class.cu:5-73): header "symmetric" doubles
 * every entry (diagonal included), "basezero" keeps 0-based ids, values forced
 * to 1 when force_ones != 0. The stored CSR is the reference's multiplied
 * matrix (rows over the file's column index). */
int mr_mtx_load(mr_ctx ctx, const char* path, int store_dtype, int force_ones, mr_mat* out);
/* Graph files of the reference's GAP-benchmark reader (parsb/include/
 * reader.h:50-162 and 334-400, built as builder.h:216-415 builds them), chosen
 * by suffix: .el / .net ("u v" pairs), .wel ("u v w"), .gr (DIMACS "a u v w"
 * lines, 1-based), .graph (METIS adjacency lines, 1-based; header format 1 =
 * weighted), .sg / .wsg (GAPBS serialized CSR, read as stored).  Text
 * formats: n = largest id + 1, one edge per (u, v) (the smallest weight),
 * self-loops dropped, and with symmetrize != 0 every missing reverse edge
 * added with the same weight.  Row u of the CSR holds the out-neighbours of
 * u; values are the weights (1 for unweighted files), stored as store_dtype.
 * A malformed file is MR_ERR_IO; another suffix is MR_ERR_UNSUPPORTED. */
int mr_gap_load(mr_ctx ctx, const char* path, int symmetrize, int store_dtype, mr_mat* out);
int mr_mat_destroy(mr_mat m);
int mr_mat_info(mr_mat m, int64_t* n_rows, int64_t* n_cols, int64_t* nnz, int* dtype);
/* copy the (current, possibly relabelled) CSR back to the host; val in the
 * storage dtype; any pointer may be NULL */
int mr_mat_export(mr_mat m, int32_t* row_ptr, int32_t* col_idx, void* val);
/* new handle holding the same matrix with values re-stored in dtype */
int mr_mat_convert(mr_mat m, int dtype, mr_mat* out);
/* accumulated relabelling perm[old] = new (identity before any reorder),
 * GRAPH::perm (class.cuh:78) */
int mr_mat_perm(mr_mat m, int32_t* perm);

/* ---- permutation ---------------------------------------------------------- */
/* check_perm (include/utils.cuh:5-55), host array; MR_OK or MR_ERR_NOT_PERMUTATION */
int mr_check_perm(const int32_t* perm, int64_t n);
/* symmetric relabel A' = P A P^T, columns sorted within rows; perm[old] = new,
 * host array.  Replaces the scatter + gather + GRAPH::sort(false) sequence of
 * every relabel (class.cu:3099-3116, 423-457). Requires a square matrix. */
int mr_permute(mr_mat m, const int32_t* perm);
/* vector relabel on the device: dir 0: out[perm[j]] = in[j] (x into the new
 * order); dir 1: out[i] = in[perm[i]] (y back to the old order, as the
 * reference's gather at class.cu:1440). Element size from the matrix's x/y type. */
int mr_permute_vec(mr_mat m, int dir, const void* in, void* out, mr_stream s);

/* ---- reordering ------------------------------------------------------------ */
typedef struct {
    int32_t width, window, batches;
    double gpu_ms;       /* GPU placement loop (incl. transpose)                 */
    double cpu_ms;       /* host phase (per-batch Gorder / whole CPU order); for
                            pGOrder / gorder5 it overlaps the GPU loop and is the
                            mean busy time of the host threads               */
    double relabel_ms;   /* device relabel P A P^T + MERBIT rebuild               */
    double total_ms;
} mr_reorder_stats;

/* GRAPH::get_width (class.cu:733-750): 1024, 16384 or 262144 */
int mr_get_width(int64_t n, int64_t nnz, int* width);
/* pGOrder = GRAPH::gorder6 (class.cu:2955-3140): GPU greedy placement in
 * batches (32 doubling up to width, then width), then Gorder with `window`
 * (5 in the reference) on host threads for every batch from log2(width).
 * width 0 = get_width; any power of two from 32 to 2^22 (1024, 16384 and
 * 262144 are the ones get_width gives).  Relabels the matrix in place;
 * perm_out (host, n, nullable) receives perm[old] = new. */
int mr_reorder_pgorder(mr_mat m, int width, int window, int32_t* perm_out, mr_reorder_stats* stats);
/* the two placement stages of gorder6 without relabelling (position -> vertex):
 * gpu_order after the GPU loop (class.cu:3062), final_order after the CPU
 * phase (class.cu:3093); either may be NULL */
int mr_pgorder_order(mr_mat m, int width, int window, uint32_t* gpu_order, uint32_t* final_order,
                     mr_reorder_stats* stats);
/* gorder6's CPU phase alone (order_on_cpu, class.cu:2220-2314, on every slice
 * from batch log2(width), class.cu:3074-3093, on a pool of `threads` host
 * threads as the reference's ThreadPool(128); 0 = all the process may use):
 * final_order = gpu_order refined, as mr_pgorder_order's final_order; *ms =
 * the phase's wall time without the host copies of the lists.  For timing the
 * phase apart from the GPU loop it overlaps in mr_reorder_pgorder.
 * MR_ERR_NOT_PERMUTATION unless gpu_order is a permutation of the n vertices. */
int mr_pgorder_cpu_phase(mr_mat m, int width, int window, const uint32_t* gpu_order, uint32_t* final_order,
                         int threads, double* ms);
/* GRAPH::gorder5 (class.cu:2727-2873), the earlier pGOrder schedule: batches
 * of vsize (0 = get_width; at most 2^22), start scores = column counts, the
 * GPU window = the last width1 batches (width1 * vsize <= 2^24) with no hub
 * caps, then order_on_cpu with window width2 on every batch.  The reference
 * calls it with (get_width, 1, 5) (main.cu:112-118).  Relabels. */
int mr_reorder_gorder5(mr_mat m, int vsize, int width1, int width2, int32_t* perm_out, mr_reorder_stats* stats);
/* gorder5's placements without relabelling, as mr_pgorder_order */
int mr_gorder5_order(mr_mat m, int vsize, int width1, int width2, uint32_t* gpu_order, uint32_t* final_order,
                     mr_reorder_stats* stats);
/* tile-aware greedy order = GRAPH::order (class.cu:3158-3365), sequential on
 * the host: rows sharing x entries are kept inside tiles of width1 x entries
 * (the reference names block x sigma), columns with width2 or more entries are
 * ignored; main.cu:129-133 takes both from the command line.  Relabels. */
int mr_reorder_tile_order(mr_mat m, int width1, int width2, int32_t* perm_out, mr_reorder_stats* stats);
/* sequential CPU Gorder = GRAPH::gorder1 (class.cu:1825-2013); relabels */
int mr_reorder_gorder_cpu(mr_mat m, int window, int32_t* perm_out, mr_reorder_stats* stats);
/* Reverse Cuthill-McKee = GRAPH::relabel_rcm (class.cu:1554-1602) over
 * R_Cuthill_Mckee (rcm/src/rcm.cpp:138-154); relabels */
int mr_reorder_rcm(mr_mat m, int32_t* perm_out, mr_reorder_stats* stats);
/* Rabbit Order = GRAPH::rabbitorder (class.cu:1607-1655) over rabbit_order
 * (rabbitorder/src/rabbit_order.cpp:1-451), run in one deterministic order
 * (the reference's OpenMP / CAS aggregation is timing dependent); relabels.
 * reference_scatter != 0 applies the ids the way the reference's call site
 * does (scattered as a position -> vertex sequence, i.e. the inverse map). */
int mr_reorder_rabbit(mr_mat m, int reference_scatter, int32_t* perm_out, mr_reorder_stats* stats);
/* degree order = GRAPH::relabel_degree (class.cu:1496-1549): vertices by
 * column count of M descending, ties by id; relabels */
int mr_reorder_degree(mr_mat m, int32_t* perm_out, mr_reorder_stats* stats);
/* degree-based grouping = GRAPH::dbg (class.cu:1683-1756): the key is
 * ceil(log2(column count + 1)), descending, ties by id; relabels */
int mr_reorder_dbg(mr_mat m, int32_t* perm_out, mr_reorder_stats* stats);

/* ---- plan cache ------------------------------------------------------------
 * Persisted pGOrder permutations and tuned MERBIT options, keyed by a 64-bit
 * fingerprint of the matrix pattern (dimensions, row_ptr, col_idx), the width
 * and the window.  No reference counterpart: the reference reorders from
 * scratch every run (GRAPH::gorder6, class.cu:2955-3140). */
/* fingerprint of a host CSR pattern; equals mr_mat_hash of the same matrix */
int mr_csr_hash(int64_t n_rows, int64_t n_cols, int64_t nnz, const int32_t* row_ptr, const int32_t* col_idx,
                uint64_t* hash);
/* fingerprint of the matrix's current pattern, computed on the device */
int mr_mat_hash(mr_mat m, uint64_t* hash);
/* mr_reorder_pgorder through the cache in cache_dir (an existing directory):
 * an entry for this pattern, width (0 = get_width) and window is applied with
 * its saved MERBIT options; otherwise pGOrder runs and the entry is written.
 * hit: 1 entry applied; 0 computed and stored; 2 computed, a stale or corrupt
 * entry replaced; -1 computed, the entry could not be written */
int mr_reorder_pgorder_cached(mr_mat m, int width, int window, const char* cache_dir, int32_t* perm_out,
                              mr_reorder_stats* stats, int* hit);
/* store the matrix's current MERBIT options (after mr_merbit_autotune) with
 * the entry of its last cached reorder; MR_ERR_INVALID_ARG if it was not
 * reordered through the cache or was relabelled since */
int mr_plan_save(mr_mat m);

/* ---- reuse distance and the locality predictor ---------------------------- */
#define MR_RD_EXACT   0
#define MR_RD_SAMPLED 1
#define MR_MODEL_MERBIT 0
#define MR_MODEL_CSR    1
/* Models fitted on MI355X timings: the same method as the reference's
 * (ridge/train.py:596-665), with bin edges at MI355X cache sizes.  They have
 * no counterpart in the reference. */
#define MR_MODEL_MERBIT_MI355X 2
#define MR_MODEL_CSR_MI355X    3
typedef struct {
    int32_t sector_size;     /* 32                                               */
    int32_t samples;         /* 30960                                            */
    int32_t rd_cap;          /* 589824 (= last edge)                             */
    int32_t hll_p;           /* 12                                               */
    int32_t max_scan;        /* 0 = scan to the stream start                     */
    int32_t stream_type;     /* seed component (class.cu:1203)                   */
    int32_t graph_id;        /* seed component (class.cu:1203)                   */
    int32_t n_edges;         /* 9                                                */
    int32_t bin_edges[32];   /* {1,112,512,18432,36864,73728,147456,294912,589824} */
} mr_rd_params;
typedef struct {
    int64_t reuses, cold;                      /* exact mode                     */
    int32_t found, cold_samples, large, unresolved;   /* sampled mode status counts */
    double ms;
} mr_rd_stats;
/* the fixed parameters of GRAPH::locality_predict (class.cu:1265-1277) */
int mr_rd_params_default(mr_rd_params* p);
/* the feature parameters a model was fitted with: the defaults above for the
 * reference's models, MI355X bin edges (and rd_cap = last edge) for the
 * MI355X models; no reference counterpart */
int mr_rd_params_model(int model, mr_rd_params* p);
/* exact sector RD histogram of a host access stream, compute_sector_rd_histogram
 * (class.cu:895-936); sectors sized by n as the reference does (:899).
 * counts / feat hold n_edges + 1 entries (either may be NULL). */
int mr_rd_exact_stream(const int32_t* stream, int64_t n, int64_t nnz, const mr_rd_params* p, int64_t* counts,
                       double* feat, mr_rd_stats* stats);
/* RD features of the matrix's x-access stream (col_idx in CSR order):
 * MR_RD_EXACT (host Fenwick pass) or MR_RD_SAMPLED (device HyperLogLog,
 * sampled_hll_rd_block_gpu, class.cu:1133-1261); feat: n_edges + 1 doubles */
int mr_rd_features(mr_mat m, int mode, const mr_rd_params* p, double* feat, mr_rd_stats* stats);
/* per-sample outputs of the sampled kernel (estimate, status, position) */
int mr_rd_sampled_raw(mr_mat m, const mr_rd_params* p, float* est, int32_t* status, int32_t* pos);
/* rd_merbit_model::PredictSpeedup (include/model.cuh:91-94) and the CSR model
 * of ridge/out/csr/sampled, or an MI355X model; feat holds the 10 features
 * computed with mr_rd_params_model(model); MR_ERR_NUMERIC for non-finite
 * input or result */
int mr_predict(const double* feat, int model, double* speedup);
/* GRAPH::locality_predict (class.cu:1263-1326): sampled RD features with the
 * model's parameters (mr_rd_params_model), then the model; feat_out (10
 * doubles) may be NULL */
int mr_locality_predict(mr_mat m, int model, double* speedup, double* feat_out, mr_rd_stats* stats);

/* ---- SpMV / SpMM ---------------------------------------------------------- */
/* y = A x (overwrite).  MERBIT replaces merbit_kernel (include/merbit.cuh:408-493);
 * CSR_SCALAR / CSR_VECTOR replace cusparseSpMV CSR_ALG1 (class.cu:1359-1406).
 * x, y: device pointers (F64 for F64 matrices, F32 otherwise).  Results are
 * deterministic (no floating-point atomics).  Calls on one handle that overlap
 * in time on different streams share the handle's MERBIT carry buffer and race;
 * use one stream per handle or mr_spmv_ws. */
int mr_spmv(mr_mat m, int kernel, const void* x, void* y, mr_stream s);
/* MERBIT with a caller workspace of mr_merbit_ws_bytes() bytes */
int mr_merbit_ws_bytes(mr_mat m, size_t* bytes);
int mr_spmv_ws(mr_mat m, const void* x, void* y, void* ws, mr_stream s);
/* MERBIT SpMV that allocates nothing.  No counterpart in the reference: it is
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  It runs
 * mr_spmv_ws's kernels with the carries in the library's static per-device
 * scratch array of 64 MiB instead of a workspace argument.  The plan must
 * already be built (a prior mr_spmv / mr_spmv_ws / mr_merbit_autotune on the
 * handle; else MR_ERR_INVALID_ARG), and plans whose mr_merbit_ws_bytes exceeds
 * 64 MiB are refused with MR_ERR_UNSUPPORTED.  Every call on one GPU shares
 * that GPU's scratch array: calls that overlap in time on one device (other
 * streams, other threads, other handles) race. */
int mr_spmv_noalloc(mr_mat m, const void* x, void* y, mr_stream s);
/* Y = A X with k right-hand sides, row-major X[n_cols][ldx], Y[n_rows][ldy] */
int mr_spmm(mr_mat m, int kernel, int k, const void* X, int64_t ldx, void* Y, int64_t ldy,
            mr_stream s);

/* ---- MERBIT metadata (TILE<BLOCK,SIGMA>, include/merbit.cuh:7-155) ---------- */
/* tile geometry: sigma merge-path items per lane (8, 16 or 32), 256 lanes
 * (4 wave64s) per block; rebuilds the metadata on the device */
int mr_merbit_configure(mr_mat m, int sigma);
/* load policy of the MERBIT kernel: bit 0 = non-temporal col/val loads and
 * y stores, bit 1 = XCD-contiguous block order, bit 2 = the wave-tile variant
 * with an LDS cache of the hottest columns of x (F32/F16/BF16 only; its tiles
 * are set by mr_merbit_xc_configure), bit 3 = the column-panel variant (F32/
 * F16/BF16 only; mr_merbit_panel_configure); bits 2 and 3 exclude each other.
 * Default 9 (column panels, non-temporal streams) for F32/F16/BF16 matrices
 * with at least 4,194,304 nonzeros, 1 otherwise.  Bits 0 and 1 do not change results; bits 2 and 3 change the
 * partition and hence the rounding order (results stay deterministic) */
int mr_merbit_options(mr_mat m, int policy);
int mr_merbit_get_options(mr_mat m, int* policy);
/* autotuner: times every (sigma, policy) on this matrix (reps launches each,
 * hipEvents) and keeps the fastest; no reference counterpart (the reference
 * fixes BLOCK=256, SIGMA=14 at class.cu:1409-1410) */
int mr_merbit_autotune(mr_mat m, int reps, int* best_sigma, int* best_policy, float* best_us);
/* lane-padded wave-tile / x-cache variant (no reference counterpart): the
 * merge path stored as padded (column, value) streams, tiles of 64 lanes x
 * items_per_lane items, `waves` waves per persistent workgroup (one per CU),
 * cached_cols hottest columns of x held in LDS (-1 = as many as fit); pairs
 * (8,16), (16,16), (8,8), (16,8) and, for F32, (4,16).  Rebuilds the
 * variant's metadata (about (n_rows + nnz) * (4 + |val|) bytes). */
int mr_merbit_xc_configure(mr_mat m, int items_per_lane, int waves, int cached_cols);
int mr_merbit_xc_info(mr_mat m, int* items_per_lane, int* waves, int* cached_cols, int64_t* tiles);
/* column-panel variant (no reference counterpart; replaces merbit_kernel's
 * x gathers, include/merbit.cuh:408-493, by LDS reads): the columns are cut
 * into panels of 32768; the nonzeros are stored panel-major as (16-bit panel
 * column, value) items on tiles of 64 lanes x items_per_lane, and a persistent
 * workgroup of `waves` waves per CU loads one panel of x into LDS at a time;
 * (items_per_lane, waves) is (8, 16), (8, 8) or (16, 8).  Every row's run of nonzeros in a panel gives one partial
 * sum; a second kernel adds each row's partials in panel order.  Rebuilds the
 * variant's plan (nnz * (2 + |val|) bytes plus three int32 per partial). */
int mr_merbit_panel_configure(mr_mat m, int items_per_lane, int waves);
/* tiles, partials, segments and workgroups are -1 before the plan is built */
int mr_merbit_panel_info(mr_mat m, int* items_per_lane, int* waves, int64_t* tiles, int64_t* partials,
                         int64_t* segments, int* workgroups, double* build_ms);
/* Bytes one SpMV moves in the stored format of the MERBIT variant the current
 * policy selects, each array counted once: the streams and metadata the
 * kernels read, the partial sums or carries they write and read back, x and
 * y.  No reference counterpart: the reference times its kernel only.  For
 * comparing with the canonical int32-CSR bytes (nnz * (4 + |val|) + (n + 1) *
 * 4 + n_cols * |x| + n_rows * |y|) where a variant compresses its indices.
 * -1 before that variant's plan is built. */
int mr_merbit_stored_bytes(mr_mat m, int64_t* bytes);
int mr_merbit_info(mr_mat m, int* sigma, int64_t* items, int64_t* n_blocks, double* build_ms);
/* row-end bitmap over the merge path (bit p of word p/32 set when item p ends a
 * row) and the merge coordinates at every block start (+ sentinel) */
int mr_merbit_export(mr_mat m, uint32_t* flags, int32_t* blk_x, int32_t* blk_y);

/* ---- batched SpMV / SpMM: many matrices, one persistent fused launch -------
 * The reference multiplies one matrix per GRAPH::spmv call (class.cu:1331-1454,
 * one merbit_kernel launch per iteration at :1421-1432); a batch replaces a
 * loop of those calls over many handles.  All matrices share one context and
 * one storage dtype; each is multiplied by k right-hand sides (k = 1: SpMV on
 * the MERBIT sigma grid; 2 <= k <= 64: SpMM on the 4096-item grid).  The fused
 * launch is a persistent kernel over the concatenated MERBIT tiles of all jobs
 * plus one carry fix-up; results are bit-identical to mr_spmv / mr_spmm with
 * the same sigma.  The batch owns its carry buffers, so it may run while the
 * member handles are used elsewhere; x / y buffers must not overlap between
 * jobs.  At most 1024 jobs. */
typedef struct mr_batch_s* mr_batch;
int mr_batch_create(mr_ctx ctx, int n_jobs, const mr_mat* mats, int k, mr_batch* out);
int mr_batch_destroy(mr_batch b);
/* x[j], y[j]: device pointers of job j (X[n_cols][ldx], Y[n_rows][ldy]; ldx /
 * ldy may be NULL = k).  The job table is copied on stream s; launches on the
 * same stream see it in order, and a captured graph reads the current table. */
int mr_batch_set_io(mr_batch b, const void* const* x, const int64_t* ldx, void* const* y, const int64_t* ldy,
                    mr_stream s);
/* sigma in {8,16,32} (k = 1 only), nt = non-temporal streams (k = 1 only),
 * grid = persistent workgroups (0 = one workgroup per tile) */
int mr_batch_configure(mr_batch b, int sigma, int nt, int grid);
/* times every (sigma, nt, grid) candidate with the current io (overwrites y)
 * and keeps the fastest */
int mr_batch_autotune(mr_batch b, int reps, int* sigma, int* nt, int* grid, float* best_us);
int mr_batch_run(mr_batch b, mr_stream s);
/* capture the current configuration as a hipGraph on stream s (not NULL) */
int mr_batch_graph_capture(mr_batch b, mr_stream s);
int mr_batch_graph_launch(mr_batch b, mr_stream s);
/* tiles of the fused launch, total nnz / rows / cols, and the algorithmic
 * bytes and flops of one run (canonical int32 CSR count) */
int mr_batch_info(mr_batch b, int64_t* tiles, int64_t* nnz, double* bytes, double* flops);

/* ---- serving loop ----------------------------------------------------------
 * No reference counterpart (main.cu:122-140 runs one matrix per process).
 * Reordered matrices stay resident; requests (matrix, k right-hand sides,
 * seed) are queued and a dispatcher thread batches whatever is queued into
 * fused MERBIT launches as soon as one of `slots` streams is free (continuous
 * batching: batches overlap on the device).  Requests on one matrix within a
 * batch are merged into one SpMM when `coalesce` is set.  A request is admitted
 * into a batch when its X / Y / carry bytes fit the slot's arena; one that can
 * never fit fails with MR_ERR_ADMISSION.  Right-hand sides are synthesised on
 * the device from the request seed (mr_serve_rhs gives the same values on the
 * host); results stay in the slot's arena until the slot is reused, so the
 * server reports a 64-bit hash (verify >= 1, mr_serve_hash on the host) or a
 * host copy (verify = 2).  Latency = request arrival -> the dispatcher sees its
 * batch complete.  Matrices must use one storage dtype (F32, F16 or BF16) and
 * must not be reconfigured or relabelled while served. */
typedef struct mr_server_s* mr_server;
typedef struct {
    int slots;              /* batches in flight, one HIP stream each (4)          */
    int max_jobs;           /* requests per batch (256, at most 1024)               */
    int coalesce;           /* merge requests on one matrix into one SpMM (1)       */
    int verify;             /* 0, 1 = keep result hashes, 2 = also host copies (0) */
    int sigma;              /* MERBIT tile width of single-vector jobs (8)          */
    int grid;               /* persistent workgroups per launch (0 = one per tile)  */
    int upload_threads;     /* host threads of a background reorder (0 = half)      */
    int heavy_slots;        /* slots heavy batches may hold at once (0 = 1)         */
    int64_t slot_bytes;     /* device arena per slot (1 GiB)                        */
    int64_t hbm_budget;     /* bytes uploaded matrices may hold (0 = free HBM only) */
    int64_t hbm_reserve;    /* device bytes left free when admitting uploads (4 GiB) */
    double timeout_ms;      /* deadline after arrival (0 = none): a request queued
                               past it, or a heavy one predicted to finish past it
                               (learned service rates), ends MR_ERR_TIMEOUT       */
} mr_serve_config;
typedef struct {
    double rate;            /* Poisson arrivals per second                          */
    int64_t n_requests;     /* trace length, uploads included                       */
    uint64_t seed;
    int k_choices[4];       /* k drawn uniformly from the first n_k entries         */
    int n_k;
    int upload_every;       /* every n-th arrival uploads a new R-MAT (0 = none)    */
    int upload_scale;       /* R-MAT scale of an upload (18)                        */
    int upload_edge_factor; /* (16)                                                 */
} mr_serve_load;
typedef struct {
    int64_t completed, rejected, timed_out, cancelled;
    int64_t uploads_admitted, uploads_rejected, evictions;
    double p50_us, p90_us, p99_us, mean_us, max_us;
    double wall_s;          /* first arrival -> last completion                     */
    double gflops;          /* 2 nnz k summed over completed requests / wall_s      */
    double gbytes_s;        /* per-request algorithmic bytes / wall_s               */
    double offered_rate, achieved_rate;
    double batches, mean_jobs_per_batch, mean_requests_per_batch;
    double upload_ms_mean;  /* generate + admit + pGOrder per admitted upload       */
    /* requests streaming >= 32 MiB (matrix, X and Y) are heavy: light ones are
     * dispatched first, heavy batches hold at most heavy_slots of the slots
     * (and never the last free one), and with a deadline a heavy request
     * predicted to finish late is shed */
    int64_t completed_heavy, timed_out_heavy;
    double p99_light_us, p99_heavy_us;
} mr_serve_report;
int mr_serve_config_default(mr_serve_config* c);
int mr_serve_create(mr_ctx ctx, const mr_serve_config* c, mr_server* out);
/* drains queued and running requests, then frees the server's own matrices */
int mr_serve_destroy(mr_server s);
/* a resident matrix owned by the caller (kept alive until the server ends) */
int mr_serve_add(mr_server s, mr_mat m, int* id);
/* asynchronous: generate an R-MAT (a,b,c = 0.57, 0.19, 0.19, scrambled ids,
 * uniform[-1,1] values), admit it by free HBM and the budget (evicting the
 * least recently used idle upload if needed), reorder it with pGOrder on the
 * context's stream and make it resident; mr_serve_upload_wait gives its id */
int mr_serve_upload_rmat(mr_server s, int scale, int edge_factor, uint64_t seed, int* ticket);
int mr_serve_upload_wait(mr_server s, int ticket, int* status, int* mat_id, double* ms);
int mr_serve_submit(mr_server s, int mat_id, int k, uint64_t seed, int64_t* req);
int mr_serve_cancel(mr_server s, int64_t req);   /* only while queued */
/* reorder later uploads through the plan cache in dir (NULL or "" = off), and
 * count the uploads that found their entry; no reference counterpart */
int mr_serve_set_plan_cache(mr_server s, const char* dir);
int mr_serve_plan_cache_stats(mr_server s, int64_t* hits, int64_t* misses);
int mr_serve_wait(mr_server s, int64_t req, int* status, double* latency_us, uint64_t* hash);
/* y[n_rows][k] of a completed request (verify = 2) */
int mr_serve_result(mr_server s, int64_t req, float* y);
/* open-loop Poisson load from the calling thread; blocks until every request
 * and upload of the trace has finished */
int mr_serve_run_load(mr_server s, const mr_serve_load* load, mr_serve_report* rep);
/* host copies of the device right-hand side generator and result hash */
int mr_serve_rhs(uint64_t seed, int64_t n_cols, int k, float* out);
int mr_serve_hash(const float* y, int64_t n_rows, int k, uint64_t* hash);

#ifdef __cplusplus
}
#endif
#endif
