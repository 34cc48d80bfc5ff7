// Sparse-grid index construction (1D hierarchy connectivity, cell 1D indices
// and the GPU cell-pair pattern scan).
#pragma once

#include <cstdint>
#include <vector>

#include "akm_internal.hpp"

namespace akm
{
//! connect_1d: CSR over the 2^max_level 1D cells, columns ascending.
int build_connect_1d(int max_level, bool full, std::vector<int> &pntr, std::vector<int> &indx,
                     std::vector<int> &diag);

//! Dense M x M table (M = 2^max_level) of connect_1d positions, -1 when the
//! two 1D cells are not connected.
int build_offset_table(int max_level, std::vector<int> &table);

//! elem[c*d + k] = 1D index of cell c in dimension k (0 for level 0, else
//! 2^(level-1) + position).
void cell_1d_indices(int d, int64_t num_cells, const int *table, std::vector<int> &elem);

//! Cell-pair CSR on the GPU: pntr (host) and the column ids (device), each
//! row's columns ascending, or in the order of order_host (a permutation of
//! the cells) when it is not null.
int device_sparse_pattern(int d, int64_t num_cells, const int *table_host,
                          std::vector<int64_t> &pntr, dev_buffer<int> &indx_dev,
                          const int *order_host = nullptr);

//! The cells sorted lexicographically by their 1D indices (dimension 0
//! first), ties by cell id.
void lexicographic_order(int d, int64_t num_cells, const int *elem, std::vector<int> &order);
} // namespace akm
