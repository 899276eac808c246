// launch_pg.hpp -- host launchers of the pGOrder GPU loop (pgorder.hip).
#pragma once
#include <hip/hip_runtime_api.h>
#include <cstdint>

namespace mr {
namespace k {

// device-resident bookkeeping of the GPU loop
struct PgState {
    int32_t tcount[2];   // sizes of the two touched-vertex lists
    int32_t nsure, ntie; // candidate list sizes
    int32_t tb, above, need;
    int32_t ycount;      // work items listed by the window's column pass
    int32_t lo_chunk;    // first 1024-id chunk that may hold unplaced vertices
    // wide batches (|P| or the window above 1024)
    int32_t msel;        // touched vertices placed by this batch
    int32_t fill_lo;     // first chunk of the fill with untouched vertices
    int32_t fill_n;      // chunks the fill covers
    int32_t kneed;       // tie selection: rank still to find among the keys under the prefix
    uint32_t khi, klo;   // tie selection: (count, id) key prefix found so far
    int32_t overflow;    // set when a wide selection listed more than |P| vertices
};

// buffers of the wide batches; cap = largest |P| of the schedule
struct PgWide {
    int32_t cap;
    uint32_t *ka, *va, *kb, *vb;   // cap each: the two stable sorts of the selection
    void* rws;                     // radix sort workspace for cap pairs
    size_t rws_bytes;
    int32_t* coff;                 // per-chunk first fill slot ((n + 1023) / 1024)
    int32_t* lens;                 // window row lengths (largest window + 1)
    void* sws;                     // scan workspace for them
    int32_t* hist2;                // 2048 digit bins, zero on entry (each pass clears its bins)
};

int pg_hist_bins();
// capacity of the window's work-item list (items of at most 64 row-list
// entries, at most one partial item per column)
int64_t pg_item_capacity(int64_t n, int64_t nnz);
// start counts, the first touched list and its count histogram (hist must be
// zero on entry; each batch clears it and builds the next batch's)
// (always_hist: build the histogram however short the list is, as a first
// batch taken by the wide path needs)
hipError_t pg_init(int64_t n, const int32_t* rp, uint32_t* score, int32_t* t0, PgState* st, int32_t* chunk_left,
                   int32_t* hist, bool always_hist, hipStream_t s);
// batches with |P| and the window both at most this size take the one-workgroup
// selection; larger ones take the wide path (needs PgWide)
constexpr int32_t kPgNarrow = 1024;
// one batch: select |P| (pl) at ps, then score the window [os, os+ol)
hipError_t pg_batch(int64_t n, int cur, const int32_t* rp, const int32_t* ci, const int32_t* tp, const int32_t* ti,
                    uint32_t* score, int32_t* const Tl[2], PgState* st, int32_t* sure, int32_t* tie, int32_t* hist,
                    int32_t* chunk_left, uint32_t* placement, int32_t* wprefix, int32_t* h, uint64_t* items,
                    int32_t* live, int32_t ps, int32_t pl, int32_t os, int32_t ol, int32_t hub_cap,
                    const PgWide* wide, hipStream_t s);
// live[y] = full length of rows(y); then, from time to time, rows(y) keeps
// only unplaced vertices (in place, first live[y] entries of the segment)
hipError_t pg_live_init(int64_t n, const int32_t* tp, int32_t* live, hipStream_t s);
hipError_t pg_compact(int64_t n, const int32_t* tp, int32_t* ti, int32_t* live, const uint32_t* score,
                      hipStream_t s);

}  // namespace k
}  // namespace mr
