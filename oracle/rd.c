/*
 * rd.c -- oracle restatement of the reuse-distance (RD) feature extractors.
 * TEST INFRASTRUCTURE ONLY.
 *
 *   oracle_rd_exact      Fenwick, bin_id, compute_sector_rd_histogram
 *                        (src/class.cu:864-936).  The sector count is sized by n
 *                        (:899), bins by upper_bound over the edges.
 *   oracle_rd_bruteforce O(nnz^2) distinct-sector count, the definition the
 *                        Fenwick pass computes (checker for the checker).
 *   oracle_hll_*         hllrd_hash_u32 / hllrd_rank_from_hash / hllrd_estimate
 *                        (:954-1012).  exp2f(-r) is exact for integer r and is
 *                        written ldexpf(1, -r); the sum runs in register order in
 *                        float like the reference's single thread.
 *   oracle_hll_sample    sampled_hll_rd_block_kernel (:1014-1131) for one block
 *   oracle_rd_sampled    sampled_hll_rd_block_gpu host part (:1133-1261)
 *
 * Pinned by: the src/test.cu:197-216 example stream (15 reuses, rd 0 x14 and
 * rd 1 x1); Fenwick == brute force on random streams; HLL at rd 0 gives exactly
 * 0 and at rd 1 lands in bin 1; HLL estimates within the HyperLogLog error band
 * of the brute-force rd at the same sampled positions; order independence of the
 * register file; status/bin accounting against hand-built streams
 * (tests/test_oracle_rd.py).  Known residual: the reference's small-range
 * correction uses the CUDA device logf; this uses the C library logf.  They can
 * differ by one ulp, which moves a bin only for an estimate within an ulp of an
 * edge (1, 112 or 512).
 */
#include "oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

static int32_t bin_of(double rd, const int32_t* edges, int32_t n_edges)
{
    /* std::upper_bound: first edge strictly greater than rd */
    int32_t lo = 0, hi = n_edges;
    while (lo < hi) {
        int32_t mid = (lo + hi) >> 1;
        if ((double)edges[mid] > rd) hi = mid; else lo = mid + 1;
    }
    return lo;
}

int oracle_rd_exact(const int32_t* stream, int32_t n, int64_t nnz, int32_t sector,
                    const int32_t* edges, int32_t n_edges, int64_t* counts,
                    int64_t* reuses, double* prob)
{
    if (sector <= 0 || n_edges <= 0) return -1;
    const int32_t nsec = (n + sector - 1) / sector;
    int64_t* last = (int64_t*)malloc(sizeof(int64_t) * ((size_t)nsec + 1));
    int32_t* bit = (int32_t*)calloc((size_t)nnz + 2, sizeof(int32_t));
    for (int32_t s = 0; s < nsec; ++s) last[s] = -1;
    const int32_t nb = n_edges + 1;
    for (int32_t b = 0; b < nb; ++b) counts[b] = 0;
    int64_t valid = 0;
#define BIT_ADD(idx, d) do { for (int64_t i_ = (idx) + 1; i_ <= nnz; i_ += i_ & -i_) bit[i_] += (d); } while (0)
    for (int64_t k = 0; k < nnz; ++k) {
        const int32_t s = stream[k] / sector;
        const int64_t old = last[s];
        if (old < 0) { BIT_ADD(k, 1); last[s] = k; continue; }
        BIT_ADD(old, -1);
        /* range_sum(old+1, k-1) = sum(k-1) - sum(old) */
        int64_t rd = 0;
        if (k - 1 >= old + 1) {
            int64_t sr = 0, sl = 0;
            for (int64_t i = k; i > 0; i -= i & -i) sr += bit[i];          /* sum(k-1) */
            for (int64_t i = old + 1; i > 0; i -= i & -i) sl += bit[i];    /* sum(old) */
            rd = sr - sl;
        }
        counts[bin_of((double)rd, edges, n_edges)]++;
        ++valid;
        BIT_ADD(k, 1);
        last[s] = k;
    }
#undef BIT_ADD
    for (int32_t b = 0; b < nb; ++b) prob[b] = valid > 0 ? (double)counts[b] / (double)valid : 0.0;
    if (reuses) *reuses = valid;
    free(last); free(bit);
    return 0;
}

int oracle_rd_bruteforce(const int32_t* stream, int64_t nnz, int32_t sector, int64_t* rd_out)
{
    for (int64_t k = 0; k < nnz; ++k) {
        const int32_t s = stream[k] / sector;
        int64_t p = k - 1;
        while (p >= 0 && stream[p] / sector != s) --p;
        if (p < 0) { rd_out[k] = -1; continue; }
        /* distinct sectors strictly between p and k */
        int64_t d = 0;
        for (int64_t j = p + 1; j < k; ++j) {
            const int32_t sj = stream[j] / sector;
            int seen = 0;
            for (int64_t q = p + 1; q < j; ++q) if (stream[q] / sector == sj) { seen = 1; break; }
            d += !seen;
        }
        rd_out[k] = d;
    }
    return 0;
}

uint32_t oracle_hll_hash(uint32_t x)
{
    x ^= x >> 16;
    x *= 0x7feb352dU;
    x ^= x >> 15;
    x *= 0x846ca68bU;
    x ^= x >> 16;
    return x;
}

/* Avalanche bias of oracle_hll_hash, as the hash's published source
 * (C. Wellons, hash-prospector, where lowbias32 is listed with bias
 * 0.17353355999581582 over all 2^32 inputs) measures it: for 2^log2_n inputs x
 * and every input bit j, count how often output bit k flips when bit j of x
 * flips; bias = 1000 * sqrt(mean over (j, k) of ((count - n/2) / (n/2))^2).
 * The sampling noise of that mean, 1/n for independent fair bits, is
 * subtracted.  A checker of the restated constants, not part of any path. */
double oracle_hll_hash_bias(int32_t log2_n, uint64_t seed)
{
    const int64_t n = (int64_t)1 << log2_n;
    static int64_t bins[32][32];
    memset(bins, 0, sizeof(bins));
    uint64_t s = seed;
    for (int64_t i = 0; i < n; ++i) {
        s += 0x9E3779B97F4A7C15ull;            /* SplitMix64 inputs */
        uint64_t z = s;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        const uint32_t x = (uint32_t)(z ^ (z >> 31));
        const uint32_t h0 = oracle_hll_hash(x);
        for (int j = 0; j < 32; ++j) {
            const uint32_t flip = h0 ^ oracle_hll_hash(x ^ (1u << j));
            for (int k = 0; k < 32; ++k) bins[j][k] += (flip >> k) & 1u;
        }
    }
    double mean = 0.0;
    for (int j = 0; j < 32; ++j)
        for (int k = 0; k < 32; ++k) {
            const double d = ((double)bins[j][k] - (double)n / 2.0) / ((double)n / 2.0);
            mean += d * d / 1024.0;
        }
    mean -= 1.0 / (double)n;
    return 1000.0 * sqrt(mean > 0.0 ? mean : 0.0);
}

int32_t oracle_hll_rank(uint32_t h, int32_t p)
{
    const uint32_t w = h >> p;
    const int32_t max_rank = 32 - p + 1;
    if (w == 0u) return max_rank;
    int32_t r = __builtin_clz(w) - p + 1;
    if (r < 1) r = 1;
    if (r > max_rank) r = max_rank;
    return r;
}

float oracle_hll_estimate(const int32_t* regs, int32_t hll_m)
{
    float sum = 0.0f;
    int32_t zeros = 0;
    for (int32_t i = 0; i < hll_m; ++i) {
        const int32_t r = regs[i];
        if (r == 0) ++zeros;
        sum += ldexpf(1.0f, -r);
    }
    const float m = (float)hll_m;
    float alpha;
    if (hll_m == 16) alpha = 0.673f;
    else if (hll_m == 32) alpha = 0.697f;
    else if (hll_m == 64) alpha = 0.709f;
    else alpha = 0.7213f / (1.0f + 1.079f / m);
    float estimate = alpha * m * m / sum;
    if (estimate <= 2.5f * m && zeros > 0) estimate = m * logf(m / (float)zeros);
    return estimate;
}

uint32_t oracle_hll_seed(int32_t graph_id, int32_t stream_type)
{
    return 246813579U + (uint32_t)graph_id * 1009U + (uint32_t)stream_type * 9176U;
}

void oracle_hll_sample(const int32_t* stream, int64_t nnz, int32_t sid, int32_t sector,
                       uint32_t seed, int32_t rd_cap, int32_t max_scan, int32_t p,
                       float* est, int32_t* status, int32_t* pos)
{
    const int32_t m = 1 << p;
    const int32_t block = 256;
    int32_t* regs = (int32_t*)calloc((size_t)m, sizeof(int32_t));
    const uint32_t r = oracle_hll_hash(seed ^ (uint32_t)((uint32_t)sid * 0x9e3779b9U + 17U));
    const int64_t sp = 1 + (int64_t)(r % (uint32_t)(nnz - 1));
    const int64_t lo = (max_scan > 0) ? ((sp - max_scan > 0) ? sp - max_scan : 0) : 0;
    int32_t st = -99;
    float e = 0.0f;
    int32_t chunks = 0;
    const int32_t target = stream[sp] / sector;
    for (int64_t hi = sp - 1; hi >= lo && st == -99; hi -= block) {
        /* nearest previous occurrence inside this chunk */
        int64_t found = -1;
        for (int32_t t = 0; t < block; ++t) {
            const int64_t idx = hi - t;
            if (idx < lo) break;
            if (stream[idx] / sector == target) { found = idx; break; }
        }
        for (int32_t t = 0; t < block; ++t) {
            const int64_t idx = hi - t;
            if (idx < lo) break;
            const int32_t sec = stream[idx] / sector;
            if (sec != target && (found < 0 || idx > found)) {
                const uint32_t h = oracle_hll_hash((uint32_t)sec);
                const int32_t rid = (int32_t)(h & (uint32_t)(m - 1));
                const int32_t rk = oracle_hll_rank(h, p);
                if (rk > regs[rid]) regs[rid] = rk;
            }
        }
        ++chunks;
        if (found >= 0) {
            e = oracle_hll_estimate(regs, m);
            st = 0;
        } else if (rd_cap > 0 && (chunks % 16 == 0)) {
            const float t = oracle_hll_estimate(regs, m);
            if (t > (float)rd_cap * 1.03f) { e = t; st = 2; }
        }
    }
    if (st == -99) {
        e = oracle_hll_estimate(regs, m);
        st = (lo == 0) ? 1 : 3;
    }
    *est = e;
    *status = st;
    *pos = (int32_t)sp;
    free(regs);
}

int oracle_rd_sampled(const int32_t* stream, int64_t nnz, int32_t samples, int32_t sector,
                      int32_t rd_cap, int32_t hll_p, int32_t max_scan, uint32_t seed,
                      const int32_t* edges, int32_t n_edges, float* est, int32_t* status,
                      int32_t* pos, int64_t* counts, int32_t* stat4, double* prob)
{
    const int32_t nb = n_edges + 1;
    for (int32_t b = 0; b < nb; ++b) { counts[b] = 0; prob[b] = 0.0; }
    for (int i = 0; i < 4; ++i) stat4[i] = 0;
    if (samples <= 0 || nnz <= 1) return 0;
    if (sector <= 0 || n_edges <= 0) return -1;
    if (rd_cap > 0 && rd_cap < edges[n_edges - 1]) return -1;
    int32_t p = hll_p <= 0 ? 12 : hll_p;
    if (p < 4) p = 4;
    if (p > 14) p = 14;
    for (int32_t s = 0; s < samples; ++s) {
        oracle_hll_sample(stream, nnz, s, sector, seed, rd_cap, max_scan, p,
                          &est[s], &status[s], &pos[s]);
        switch (status[s]) {
        case 0: stat4[0]++; counts[bin_of((double)est[s], edges, n_edges)]++; break;
        case 1: stat4[1]++; break;
        case 2: stat4[2]++; counts[nb - 1]++; break;
        default: stat4[3]++; break;
        }
    }
    const int32_t valid = stat4[0] + stat4[2];
    if (valid > 0)
        for (int32_t b = 0; b < nb; ++b) prob[b] = (double)counts[b] / (double)valid;
    return 0;
}
