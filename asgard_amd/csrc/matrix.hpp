// The kronmult matrix object behind the akm_matrix handle.
#pragma once

#include <memory>
#include <vector>

#include "akm_internal.hpp"
#include "kron_kernels.hpp"

struct akm_gmres_workspace;

struct akm_matrix_s
{
  akm_prec prec = AKM_F64;
  int d = 0, n = 0, rows = 0, cols = 0, T = 0;
  int mode = akm::mode_ia;
  int64_t N = 1;   // tensor size n^d
  int64_t nnz = 0; // connected (row, col) pairs
  size_t esize = 8;

  // sparse: CSR (host row pointer kept for work decomposition)
  std::vector<int64_t> pntr;
  akm::dev_buffer<int> indx, iA;
  akm::dev_buffer<unsigned char> vA;
  int64_t num_vA = 0;
  // grid mode
  akm::dev_buffer<int> elem, off1d;
  int M = 0;
  // dense mode
  std::vector<akm::dev_buffer<unsigned char>> terms;
  akm::dev_buffer<void *> term_ptrs;
  int nb = 0, row_offset = 0, col_offset = 0;

  // work decomposition
  int64_t chunk_columns = 0; // 0 = automatic
  int force_generic = 0;
  bool chunks_ready = false;
  int num_chunks = 0;
  akm::dev_buffer<int> chunk_row, row_chunk_begin, counters;
  akm::dev_buffer<int64_t> chunk_beg, chunk_end;
  akm::dev_buffer<unsigned char> partial;

  // workspaces for host-resident vectors
  akm::dev_buffer<unsigned char> xws, yws;

  hipStream_t stream = nullptr;

  std::unique_ptr<akm_gmres_workspace, void (*)(akm_gmres_workspace *)> gmres{nullptr,
                                                                             nullptr};

  int64_t input_size() const { return N * cols; }
  int64_t output_size() const { return N * rows; }
  int64_t flops() const { return N * n * d * int64_t(T) * nnz * 2; }
};

namespace akm
{
//! y = alpha*A*x + beta*y + gamma*xg on device pointers (xg row-indexed).
int apply_device(akm_matrix_s &m, double alpha, const void *x, double beta, void *y,
                 double gamma, const void *xg);

//! Columns per chunk and lanes per column of the kernel the matrix uses.
int lanes_per_column(akm_prec prec, int d, int n, int force_generic);
} // namespace akm
