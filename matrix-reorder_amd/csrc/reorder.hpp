// reorder.hpp -- internal reordering helpers shared by reorder.cpp,
// cpu_order.cpp and the serving loop.
#pragma once
#include <cstdint>
#include <vector>

#include "common.hpp"

namespace mr {

// CSC of the matrix on the device (row ids ascending inside each column)
int gpu_transpose(mr_mat m, DevBuf& tp, DevBuf& ti, hipStream_t s);
// apply perm[old] = new on the device (P A P^T), accumulate m->perm, rebuild MERBIT
int relabel(mr_mat m, const std::vector<int32_t>& perm, double* ms_out);
int32_t pg_schedule(int32_t n, int32_t W, std::vector<int32_t>& Ps, std::vector<int32_t>& Pl,
                    std::vector<int32_t>& Os, std::vector<int32_t>& Ol);
// sequential CPU Gorder (gorder1) on host CSR + CSC; order[pos] = vertex
void gorder_cpu(int32_t n, const int32_t* rp, const int32_t* ci, const int32_t* tp, const int32_t* ti,
                int32_t window, uint32_t* order);
// tile-aware greedy order (GRAPH::order) on host CSR + CSC; order[pos] = vertex
void tile_order(int32_t n, const int32_t* rp, const int32_t* ci, const int32_t* tp, const int32_t* ti,
                int32_t width1, int32_t width2, uint32_t* order);
// Reverse Cuthill-McKee over a COO edge list (both directions are neighbours)
void rcm_order(int32_t n, int64_t nz, const int32_t* rows, const int32_t* cols, uint32_t* order);
// Rabbit Order on the symmetrised pattern of M (rp, ci): new_id[v] = position of v
void rabbit_order(int32_t n, const int32_t* rp, const int32_t* ci, uint32_t* new_id);

}  // namespace mr
