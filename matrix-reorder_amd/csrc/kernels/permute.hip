// permute.hip -- symmetric relabel A' = P A P^T of a CSR on gfx950, columns
// sorted within every row.
//
// Reference: every reorder ends with scatter(iota, placement) -> perm, a gather
// of both COO index arrays through perm and a full COO re-sort by (col, row)
// with a Thrust comparison sort (src/class.cu:3099-3116, 423-457).  Here the
// row structure is kept: new row p is old row inv[p]; its entries become
// (perm[col], position-in-old-row) keys that are sorted per row.  Because the
// position is part of the key, equal columns keep their input order, which is
// what the reference's stable merge sort gives.  Rows are binned by length:
//   class 0      len <= 1: copy
//   class 1..6   len <= 2, 4, ..., 64: bitonic sort in registers across a
//                group of 2..64 lanes of a wave64 (xor shuffles), 64/w rows
//                per wave
//   class 7      len <= 4096: one workgroup, bitonic sort in LDS
//   class 8      longer: all such rows at once, two device-wide stable radix
//                sorts (radix.hip) of their entries, by new column and then by
//                row slot (equal columns stay in row order)
// No floating point is touched: the output is bit-identical to a sequential
// relabel + stable sort.

#include "dtypes.hpp"
#include "launch_perm.hpp"
#include "launch_radix.hpp"

namespace mr {
namespace k {

namespace {

__device__ __forceinline__ int class_of(int32_t len)
{
    if (len <= 1) return 0;
    if (len <= 64) return 32 - __clz(len - 1);   // 2->1, 3..4->2, ..., 33..64->6
    if (len <= 4096) return 7;
    return 8;
}

__global__ void row_len_kernel(int64_t n, const int32_t* __restrict__ rp, const int32_t* __restrict__ inv,
                               int32_t* __restrict__ new_len)
{
    for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < n; p += (int64_t)gridDim.x * blockDim.x) {
        const int32_t o = inv[p];
        new_len[p] = rp[o + 1] - rp[o];
    }
}

__global__ __launch_bounds__(256) void class_count_kernel(int64_t n, const int32_t* __restrict__ new_len,
                                                          int32_t* __restrict__ counts)
{
    __shared__ int32_t h[kPermClasses];
    if (threadIdx.x < kPermClasses) h[threadIdx.x] = 0;
    __syncthreads();
    for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < n; p += (int64_t)gridDim.x * blockDim.x)
        atomicAdd(&h[class_of(new_len[p])], 1);
    __syncthreads();
    if (threadIdx.x < kPermClasses && h[threadIdx.x]) atomicAdd(&counts[threadIdx.x], h[threadIdx.x]);
}

__global__ __launch_bounds__(256) void class_fill_kernel(int64_t n, const int32_t* __restrict__ new_len,
                                                         const int32_t* __restrict__ offs,
                                                         int32_t* __restrict__ cursor, int32_t* __restrict__ lists)
{
    // one pass over a tile of 256 rows: local ranks in LDS, one reservation per class
    __shared__ int32_t h[kPermClasses];
    __shared__ int32_t base[kPermClasses];
    for (int64_t t0 = blockIdx.x * (int64_t)blockDim.x; t0 < n; t0 += (int64_t)gridDim.x * blockDim.x) {
        if (threadIdx.x < kPermClasses) h[threadIdx.x] = 0;
        __syncthreads();
        const int64_t p = t0 + threadIdx.x;
        int c = -1, r = 0;
        if (p < n) {
            c = class_of(new_len[p]);
            r = atomicAdd(&h[c], 1);
        }
        __syncthreads();
        if (threadIdx.x < kPermClasses)
            base[threadIdx.x] = h[threadIdx.x] ? atomicAdd(&cursor[threadIdx.x], h[threadIdx.x]) : 0;
        __syncthreads();
        if (c >= 0) lists[offs[c] + base[c] + r] = (int32_t)p;
        __syncthreads();
    }
}

template <typename E>
__global__ void copy_class0_kernel(int32_t count, const int32_t* __restrict__ rows, const int32_t* __restrict__ rp,
                                   const int32_t* __restrict__ ci, const E* __restrict__ val,
                                   const int32_t* __restrict__ perm, const int32_t* __restrict__ inv,
                                   const int32_t* __restrict__ new_rp, int32_t* __restrict__ new_ci,
                                   E* __restrict__ new_val)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= count) return;
    const int32_t p = rows[i], o = inv[p];
    const int32_t ob = rp[o], len = rp[o + 1] - ob;
    if (len == 1) {
        new_ci[new_rp[p]] = perm[ci[ob]];
        new_val[new_rp[p]] = val[ob];
    }
}

// bitonic sort of one key per lane across groups of W lanes
template <int W>
__device__ __forceinline__ uint64_t group_bitonic(uint64_t key, int li)
{
#pragma unroll
    for (int k2 = 2; k2 <= W; k2 <<= 1) {
#pragma unroll
        for (int j = k2 >> 1; j > 0; j >>= 1) {
            const uint64_t other = __shfl_xor(key, j, 64);
            const bool up = (li & k2) == 0;
            const bool lower = (li & j) == 0;
            const uint64_t mn = key < other ? key : other;
            const uint64_t mx = key < other ? other : key;
            key = (lower == up) ? mn : mx;
        }
    }
    return key;
}

template <int W, typename E>
__global__ __launch_bounds__(256) void sort_small_kernel(int32_t count, const int32_t* __restrict__ rows,
                                                         const int32_t* __restrict__ rp, const int32_t* __restrict__ ci,
                                                         const E* __restrict__ val, const int32_t* __restrict__ perm,
                                                         const int32_t* __restrict__ inv,
                                                         const int32_t* __restrict__ new_rp,
                                                         int32_t* __restrict__ new_ci, E* __restrict__ new_val)
{
    const int64_t g = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) / W;
    const int li = threadIdx.x & (W - 1);
    if (g >= count) return;   // whole groups leave together
    const int32_t p = rows[g], o = inv[p];
    const int32_t ob = rp[o], len = rp[o + 1] - ob;
    uint64_t key = ~0ull;
    if (li < len) key = ((uint64_t)(uint32_t)perm[ci[ob + li]] << 32) | (uint32_t)li;
    key = group_bitonic<W>(key, li);
    if (li < len) {
        const int32_t dst = new_rp[p] + li;
        new_ci[dst] = (int32_t)(key >> 32);
        new_val[dst] = val[ob + (int32_t)(key & 0xffffffffu)];
    }
}

// workgroup bitonic over `P` keys held at `a` (LDS or global), 256 threads
__device__ __forceinline__ void block_bitonic(uint64_t* a, int P)
{
    for (int k2 = 2; k2 <= P; k2 <<= 1) {
        for (int j = k2 >> 1; j > 0; j >>= 1) {
            for (int i = threadIdx.x; i < P / 2; i += blockDim.x) {
                // i-th compare pair of this stage: lo has bit j clear
                const int lo = ((i & ~(j - 1)) << 1) | (i & (j - 1));
                const int hi = lo | j;
                const uint64_t x = a[lo], y = a[hi];
                const bool up = (lo & k2) == 0;
                if ((x > y) == up) { a[lo] = y; a[hi] = x; }
            }
            __syncthreads();
        }
    }
}

template <typename E>
__global__ __launch_bounds__(256) void sort_block_kernel(int32_t count, const int32_t* __restrict__ rows,
                                                         const int32_t* __restrict__ rp, const int32_t* __restrict__ ci,
                                                         const E* __restrict__ val, const int32_t* __restrict__ perm,
                                                         const int32_t* __restrict__ inv,
                                                         const int32_t* __restrict__ new_rp,
                                                         int32_t* __restrict__ new_ci, E* __restrict__ new_val)
{
    __shared__ uint64_t a[4096];
    const int32_t p = rows[blockIdx.x], o = inv[p];
    const int32_t ob = rp[o], len = rp[o + 1] - ob;
    int P = 128;
    while (P < len) P <<= 1;
    for (int i = threadIdx.x; i < P; i += blockDim.x)
        a[i] = (i < len) ? (((uint64_t)(uint32_t)perm[ci[ob + i]] << 32) | (uint32_t)i) : ~0ull;
    __syncthreads();
    block_bitonic(a, P);
    const int32_t nb = new_rp[p];
    for (int i = threadIdx.x; i < len; i += blockDim.x) {
        const uint64_t key = a[i];
        new_ci[nb + i] = (int32_t)(key >> 32);
        new_val[nb + i] = val[ob + (int32_t)(key & 0xffffffffu)];
    }
}

// long rows: key = (slot << cbits) | new column, value = position in the old
// row; one block per row
// entry g = boff[i] + t of slot i: key = its new column, value = g, slot[g] = i
__global__ __launch_bounds__(256) void long_keys_kernel(int32_t count, const int32_t* __restrict__ rows,
                                                        const int32_t* __restrict__ rp, const int32_t* __restrict__ ci,
                                                        const int32_t* __restrict__ perm,
                                                        const int32_t* __restrict__ inv,
                                                        const int64_t* __restrict__ boff, uint32_t* __restrict__ keys,
                                                        uint32_t* __restrict__ ent, int32_t* __restrict__ slot)
{
    const int32_t i = blockIdx.x;
    const int32_t p = rows[i], o = inv[p];
    const int32_t ob = rp[o], len = rp[o + 1] - ob;
    const int64_t base = boff[i];
    for (int32_t t = threadIdx.x; t < len; t += blockDim.x) {
        keys[base + t] = (uint32_t)perm[ci[ob + t]];
        ent[base + t] = (uint32_t)(base + t);
        slot[base + t] = i;
    }
}

// keys of the second sort: the slot of each column-sorted entry
__global__ void long_slot_keys_kernel(int64_t total, const uint32_t* __restrict__ ent, const int32_t* __restrict__ slot,
                                      uint32_t* __restrict__ keys)
{
    for (int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; j < total; j += (int64_t)gridDim.x * blockDim.x)
        keys[j] = (uint32_t)slot[ent[j]];
}

// slot i's sorted entries sit at [boff[i], boff[i] + len): ent holds the
// entry each came from, so its old position in the row is ent - boff[i]
template <typename E>
__global__ __launch_bounds__(256) void long_write_kernel(int32_t count, const int32_t* __restrict__ rows,
                                                         const int32_t* __restrict__ rp, const int32_t* __restrict__ ci,
                                                         const int32_t* __restrict__ perm,
                                                         const int32_t* __restrict__ inv,
                                                         const int32_t* __restrict__ new_rp,
                                                         const int64_t* __restrict__ boff,
                                                         const uint32_t* __restrict__ ent, const E* __restrict__ val,
                                                         int32_t* __restrict__ new_ci, E* __restrict__ new_val)
{
    const int32_t i = blockIdx.x;
    const int32_t p = rows[i], o = inv[p];
    const int32_t ob = rp[o], len = rp[o + 1] - ob;
    const int32_t nb = new_rp[p];
    const int64_t base = boff[i];
    for (int32_t t = threadIdx.x; t < len; t += blockDim.x) {
        const int32_t tt = (int32_t)((int64_t)ent[base + t] - base);
        new_ci[nb + t] = perm[ci[ob + tt]];
        new_val[nb + t] = val[ob + tt];
    }
}

__global__ void gather_len_kernel(int32_t count, const int32_t* __restrict__ rows, const int32_t* __restrict__ new_len,
                                  int32_t* __restrict__ out)
{
    for (int32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x)
        out[i] = new_len[rows[i]];
}

unsigned bits_for(int64_t v)   // bits to hold 0..v
{
    unsigned b = 1;
    while (b < 63 && ((int64_t)1 << b) <= v) ++b;
    return b;
}

template <typename E>
hipError_t sort_class_t(int cls, int32_t count, const int32_t* rows, const int32_t* rp, const int32_t* ci,
                        const void* val, const int32_t* perm, const int32_t* inv, const int32_t* new_rp,
                        int32_t* new_ci, void* new_val, hipStream_t s)
{
    const E* v = static_cast<const E*>(val);
    E* nv = static_cast<E*>(new_val);
    auto small = [&](auto wtag) {
        constexpr int W = decltype(wtag)::value;
        const int64_t threads = (int64_t)count * W;
        sort_small_kernel<W, E><<<(unsigned)((threads + 255) / 256), 256, 0, s>>>(count, rows, rp, ci, v, perm,
                                                                                   inv, new_rp, new_ci, nv);
    };
    switch (cls) {
    case 0:
        copy_class0_kernel<E><<<(unsigned)((count + 255) / 256), 256, 0, s>>>(count, rows, rp, ci, v, perm, inv,
                                                                                new_rp, new_ci, nv);
        break;
    case 1: small(std::integral_constant<int, 2>{}); break;
    case 2: small(std::integral_constant<int, 4>{}); break;
    case 3: small(std::integral_constant<int, 8>{}); break;
    case 4: small(std::integral_constant<int, 16>{}); break;
    case 5: small(std::integral_constant<int, 32>{}); break;
    case 6: small(std::integral_constant<int, 64>{}); break;
    case 7:
        sort_block_kernel<E><<<(unsigned)count, 256, 0, s>>>(count, rows, rp, ci, v, perm, inv, new_rp, new_ci, nv);
        break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// inv[perm[i]] = i; an out-of-range entry sets *bad instead of writing
__global__ void invert_kernel(int64_t n, const int32_t* __restrict__ perm, int32_t* __restrict__ inv,
                              int32_t* __restrict__ bad)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int32_t p = perm[i];
        if ((uint32_t)p < (uint32_t)n) inv[p] = (int32_t)i;
        else *bad = 1;
    }
}

// an entry of inv still at its -1 fill means perm repeats a value
__global__ void inverse_holes_kernel(int64_t n, const int32_t* __restrict__ inv, int32_t* __restrict__ bad)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        if (inv[i] < 0) *bad = 1;
}

__global__ void col_count_kernel(int64_t nnz, const int32_t* __restrict__ ci, int32_t* __restrict__ cnt)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nnz; i += (int64_t)gridDim.x * blockDim.x)
        atomicAdd(&cnt[ci[i]], 1);
}

}  // namespace

hipError_t perm_row_lengths(int64_t n, const int32_t* rp, const int32_t* inv, int32_t* new_len, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    int64_t g = (n + 255) / 256;
    if (g > 65536) g = 65536;
    row_len_kernel<<<(unsigned)g, 256, 0, s>>>(n, rp, inv, new_len);
    return hipGetLastError();
}

hipError_t perm_class_count(int64_t n, const int32_t* new_len, int32_t* counts, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    int64_t g = (n + 255) / 256;
    if (g > 4096) g = 4096;
    class_count_kernel<<<(unsigned)g, 256, 0, s>>>(n, new_len, counts);
    return hipGetLastError();
}

hipError_t perm_class_fill(int64_t n, const int32_t* new_len, const int32_t* offs, int32_t* cursor,
                           int32_t* lists, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    int64_t g = (n + 255) / 256;
    if (g > 4096) g = 4096;
    class_fill_kernel<<<(unsigned)g, 256, 0, s>>>(n, new_len, offs, cursor, lists);
    return hipGetLastError();
}

hipError_t perm_sort_class(int cls, int32_t count, const int32_t* rows, const int32_t* rp, const int32_t* ci,
                           const void* val, int val_bytes, const int32_t* perm, const int32_t* inv,
                           const int32_t* new_rp, int32_t* new_ci, void* new_val, hipStream_t s)
{
    if (count <= 0) return hipSuccess;
    if (cls < 0 || cls >= kPermClasses - 1) return hipErrorInvalidValue;   // the long class: perm_sort_long
    switch (val_bytes) {
    case 2: return sort_class_t<uint16_t>(cls, count, rows, rp, ci, val, perm, inv, new_rp, new_ci, new_val, s);
    case 4: return sort_class_t<uint32_t>(cls, count, rows, rp, ci, val, perm, inv, new_rp, new_ci, new_val, s);
    case 8: return sort_class_t<uint64_t>(cls, count, rows, rp, ci, val, perm, inv, new_rp, new_ci, new_val, s);
    default: return hipErrorInvalidValue;
    }
}

hipError_t perm_long_lengths(int32_t count, const int32_t* rows, const int32_t* new_len, int32_t* out, hipStream_t s)
{
    if (count <= 0) return hipSuccess;
    int64_t g = ((int64_t)count + 255) / 256;
    if (g > 4096) g = 4096;
    gather_len_kernel<<<(unsigned)g, 256, 0, s>>>(count, rows, new_len, out);
    return hipGetLastError();
}

// five arrays of total entries (keys, entries, their sorted copies, slots)
static size_t long_arrays_bytes(int64_t total)
{
    return (5 * sizeof(uint32_t) * (size_t)(total > 0 ? total : 1) + 255) / 256 * 256;
}

size_t perm_long_ws_bytes(int64_t total, int32_t count, int64_t n)
{
    (void)count;
    (void)n;
    return long_arrays_bytes(total) + radix_sort_ws_bytes(total) + 256;
}

hipError_t perm_sort_long(int32_t count, const int32_t* rows, const int32_t* rp, const int32_t* ci, const void* val,
                          int val_bytes, const int32_t* perm, const int32_t* inv, const int32_t* new_rp,
                          int32_t* new_ci, void* new_val, const int64_t* boff, int64_t total, int64_t n, void* ws,
                          size_t ws_bytes, hipStream_t s)
{
    if (count <= 0 || total <= 0) return hipSuccess;
    // (slot, new column) order with equal columns in row order: stable by
    // column, then stable by slot
    if (total > ((int64_t)1 << 31) - 1) return hipErrorInvalidValue;
    const unsigned cbits = bits_for(n > 0 ? n - 1 : 0);
    const unsigned sbits = bits_for(count - 1);
    const size_t arrays = long_arrays_bytes(total);
    if (ws_bytes < arrays + radix_sort_ws_bytes(total)) return hipErrorInvalidValue;
    uint32_t* ka = static_cast<uint32_t*>(ws);
    uint32_t* ea = ka + total;
    uint32_t* kb = ea + total;
    uint32_t* eb = kb + total;
    int32_t* slot = reinterpret_cast<int32_t*>(eb + total);
    void* tmp = static_cast<char*>(ws) + arrays;
    const size_t tmp_bytes = ws_bytes - arrays;
    long_keys_kernel<<<(unsigned)count, 256, 0, s>>>(count, rows, rp, ci, perm, inv, boff, ka, ea, slot);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    e = radix_sort_pairs(ka, ea, kb, eb, total, (int)cbits, tmp, tmp_bytes, s);
    if (e != hipSuccess) return e;
    {
        int64_t g = (total + 255) / 256;
        if (g > 65536) g = 65536;
        long_slot_keys_kernel<<<(unsigned)g, 256, 0, s>>>(total, eb, slot, ka);
    }
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    e = radix_sort_pairs(ka, eb, kb, ea, total, (int)sbits, tmp, tmp_bytes, s);
    if (e != hipSuccess) return e;
    switch (val_bytes) {
    case 2:
        long_write_kernel<uint16_t><<<(unsigned)count, 256, 0, s>>>(count, rows, rp, ci, perm, inv, new_rp, boff, ea,
                                                                    static_cast<const uint16_t*>(val), new_ci,
                                                                    static_cast<uint16_t*>(new_val));
        break;
    case 4:
        long_write_kernel<uint32_t><<<(unsigned)count, 256, 0, s>>>(count, rows, rp, ci, perm, inv, new_rp, boff, ea,
                                                                    static_cast<const uint32_t*>(val), new_ci,
                                                                    static_cast<uint32_t*>(new_val));
        break;
    case 8:
        long_write_kernel<uint64_t><<<(unsigned)count, 256, 0, s>>>(count, rows, rp, ci, perm, inv, new_rp, boff, ea,
                                                                    static_cast<const uint64_t*>(val), new_ci,
                                                                    static_cast<uint64_t*>(new_val));
        break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t invert_perm(int64_t n, const int32_t* perm, int32_t* inv, int32_t* bad, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    int64_t g = (n + 255) / 256;
    if (g > 65536) g = 65536;
    hipError_t e = hipMemsetAsync(inv, 0xff, sizeof(int32_t) * (size_t)n, s);
    if (e == hipSuccess) e = hipMemsetAsync(bad, 0, sizeof(int32_t), s);
    if (e != hipSuccess) return e;
    invert_kernel<<<(unsigned)g, 256, 0, s>>>(n, perm, inv, bad);
    inverse_holes_kernel<<<(unsigned)g, 256, 0, s>>>(n, inv, bad);
    return hipGetLastError();
}

hipError_t transpose_count(int64_t nnz, const int32_t* ci, int32_t* cnt, hipStream_t s)
{
    if (nnz <= 0) return hipSuccess;
    int64_t g = (nnz + 255) / 256;
    if (g > 65536) g = 65536;
    col_count_kernel<<<(unsigned)g, 256, 0, s>>>(nnz, ci, cnt);
    return hipGetLastError();
}

}  // namespace k
}  // namespace mr
