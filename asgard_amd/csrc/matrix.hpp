// The kronmult matrix object behind the akm_matrix handle.
#pragma once

#include <memory>
#include <vector>

#include "akm_internal.hpp"
#include "kron_kernels.hpp"

struct akm_gmres_workspace;
struct akm_global_s;

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
  // advanced by every call that rewrites vA, so tables derived from it know
  // when to rebuild
  uint64_t vA_stamp = 0;
  // the first implicit solve's timed operator apply (ms; -1 before it),
  // which decides whether GMRES steps are enqueued ahead
  double gmres_apply_ms = -1;
  // grid mode
  akm::dev_buffer<int> elem, off1d;
  int M = 0;
  int max_level = 0;
  // grid mode, LDS factor tables: factors as [dim][connection][term][nnp],
  // connect_1d row pointer, packed cell 1D indices
  akm::dev_buffer<unsigned char> vAk;
  akm::dev_buffer<int> conn_pntr;
  akm::dev_buffer<unsigned long long> epack;
  std::vector<int> conn_pntr_host;
  int nconn = 0, nnp = 0;
  int64_t max_table_elems = 0;
  int64_t max_table_last4 = 0; // largest per-row table of the last four dimensions
  int lds_cap = -1; // -1 automatic, 0 disables the LDS tables
  // dense mode
  std::vector<akm::dev_buffer<unsigned char>> terms;
  akm::dev_buffer<void *> term_ptrs;
  int nb = 0, row_offset = 0, col_offset = 0;

  // work decomposition: one workgroup per row, rows longest first
  int waves = 0;        // wavefronts per row workgroup (0 = automatic)
  int force_generic = 0;
  int family = 0;       // 0 register / LDS-table kernels, 1 matrix-core kernel
  akm::dev_buffer<int64_t> pntr_dev;
  akm::dev_buffer<int> row_order;

  // workspaces for host-resident vectors
  akm::dev_buffer<unsigned char> xws, yws;

  // operator update on the device: connect_1d (rows, then columns) in vA
  // order, and the device copy of the caller's operator pointers
  akm::dev_buffer<int> op_conn;
  akm::dev_buffer<const void *> op_ptrs;
  // IMEX step vectors (four of output_size), on the implicit handle
  akm::dev_buffer<double> imex_ws;

  hipStream_t stream = nullptr;

  std::unique_ptr<akm_gmres_workspace, void (*)(akm_gmres_workspace *)> gmres{nullptr,
                                                                             nullptr};
  // global Kronecker operator (mode_global)
  std::unique_ptr<akm_global_s, void (*)(akm_global_s *)> gk{nullptr, nullptr};

  int64_t input_size() const { return N * cols; }
  int64_t output_size() const { return N * rows; }
  int64_t flops() const;
};

namespace akm
{
//! y = alpha*A*x + beta*y + gamma*xg on device pointers (xg row-indexed).
int apply_device(akm_matrix_s &m, double alpha, const void *x, double beta, void *y,
                 double gamma, const void *xg);

//! Wavefronts per row workgroup used for the matrix (explicit or automatic).
int waves_per_row(akm_matrix_s const &m);

//! Global operator: y = alpha*A*x + beta*y + gamma*xg on device pointers,
//! divided by (1 - jdt * pc) with the reference's Jacobi diagonal when
//! `jacobi` is set.
int global_apply(akm_matrix_s &m, double alpha, const void *x, double beta, void *y,
                 double gamma, const void *xg, bool jacobi, double jdt);
//! Global operator: new factor values (same vA layout), Jacobi diagonal rebuilt.
int global_update(akm_matrix_s &m, const void *vA);
//! Global operator: Jacobi diagonal rebuilt on the device from the stored
//! factor blocks (after they were written on the device).
int global_device_jacobi(akm_matrix_s &m);
//! Global operator: flops per apply as the reference counts them.
int64_t global_flops(akm_matrix_s const &m);
//! Global operator: row layout in use, as akm_set_tuning takes it (1, 8, 2,
//! 3, 4 or 5).
int global_layout(akm_matrix_s const &m);
//! Global operator: times the lane layouts on the operator, keeps the faster.
int global_autotune(akm_matrix_s &m, const void *x, void *y);
//! Global operator: row layout 1 (lane per DOF), 8 (8 lanes per DOF), 2
//! (wavefront per cell), 3 (wavefront per cell, chains sharing their first
//! stage's input read it once), 4 / 5 (2 / 3 with each XCD taking contiguous
//! cells), 0 (unchanged).
int global_set_layout(akm_matrix_s &m, int layout);

//! Kernel arguments of an FP64 matrix (vectors and scalars left empty).
kron_args<double> kron_args_f64(akm_matrix_s const &m);
} // namespace akm
