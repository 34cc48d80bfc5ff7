// Kronmult kernels for gfx950 (CDNA4, wave64).
//
// Operator: y[row] = alpha * sum_{col in row} sum_t (A_{t,0} x ... x A_{t,d-1}) x[col]
//                    + beta * y[row] + gamma * xg[row]
// Tensors are n^d entries per cell, row-major (dimension 0 slowest); every
// n x n factor is column-major (A[j*n + s] = A(s, j)), as in the reference
// (src/device/asgard_spkronmult_cpu.cpp:84-433).  The gamma term fuses the
// identity of the implicit operator (I - dt A) into the apply.
//
// Work decomposition: one workgroup of W wavefronts owns one output row (a
// cell); rows are dispatched longest first.  The W wavefronts take turns over
// the row's column list, CPW columns per wavefront step.  Inside a
// wavefront, G = n^g lanes share one column: lane `sub` of the group holds
// the R = n^(d-g) tensor entries whose leading g digits equal `sub`.
// Contraction stages over the trailing d-g dimensions run in place in
// registers; stages over the leading g dimensions exchange values inside the
// lane group (DPP for pairs, ds_bpermute otherwise).  At the end the
// per-lane accumulators are summed through LDS in a fixed order (wavefront by
// wavefront, slot by slot), and the row is stored once.  No atomics, no
// cross-workgroup traffic: results are deterministic run to run.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace akm
{
constexpr int wave_size = 64;

template<int n, int p>
struct ipow
{
  static constexpr int value = n * ipow<n, p - 1>::value;
};
template<int n>
struct ipow<n, 0>
{
  static constexpr int value = 1;
};

//! How the factor blocks of a (pair, term, dim) are addressed.
enum addr_mode : int
{
  //! explicit per-pair index list iA (the reference's sparse constructor)
  mode_ia = 0,
  //! sparse-grid: 1D offsets from the cell 1D indices and a connect_1d table
  mode_grid = 1,
  //! dense mode: all column cells, per-term block arrays (make_kronmult_dense)
  mode_dense = 2,
  //! global Kronecker operator (global_kron_matrix), one dimension at a time
  mode_global = 3,
};

template<typename P>
struct kron_args
{
  // factor data
  const P *vA;           // mode_ia / mode_grid: factor blocks
  const int *iA;         // mode_ia: (pair*T + t)*d + k -> offset into vA
  const int64_t *pntr;   // mode_ia / mode_grid: CSR row pointer
  const int *indx;       // mode_ia / mode_grid: column cell ids (CSR)
  const int *elem;       // mode_grid / mode_dense: 1D index per (cell, dim)
  const int *off1d;      // mode_grid: connect_1d position, [M*M], -1 if none
  int M;                 // mode_grid: 2^max_level
  const P *const *terms; // mode_dense: per-term block arrays (device)
  int nb;                // mode_dense: num_1d_blocks
  int row_offset;        // mode_dense
  int col_offset;        // mode_dense
  int cols;              // mode_dense: number of column cells
  int T;                 // number of terms
  // mode_grid with LDS factor tables
  const P *vAk;                     // factors as [dim][connection][term][nnp]
  const int *conn_pntr;             // connect_1d row pointer (M + 1)
  const unsigned long long *epack;  // 1D indices of a cell, 10 bits per dim
  int nconn;                        // connect_1d connections
  int nnp;                          // padded n*n block size in vAk
  int64_t max_table_elems;          // largest per-row factor table (elements)
  int64_t max_table_last4;          // the same for the last four dimensions only
  int lds_cap;                      // LDS bytes a workgroup may use (0: no tables)
  // rows in dispatch order (longest first)
  const int *row_order;
  // kernel family: 0 register / LDS-table kernels, 1 the sum-factorised
  // matrix-core kernel (FP64, sparse-grid, d = 3, n = 3)
  int family;
  // vectors
  const P *x;
  P *y;
  const P *xg; // gamma term input (row-indexed), may be null when gamma == 0
  P alpha, beta, gamma;
};

//! Vectors of one system of a batched solve: the Krylov basis V (restart + 1
//! columns of n), the solution X, the size and the time step.
struct batch_system
{
  double *V;
  double *X;
  int64_t n;
  double dt;
};

//! Padded n*n block size of the re-laid-out factor array (16-byte rows).
template<typename P>
__host__ __device__ constexpr int padded_block(int n)
{
  return (sizeof(P) == 8) ? (n * n + 1) / 2 * 2 : (n * n + 3) / 4 * 4;
}

__device__ __forceinline__ double swap_pair(double v)
{
  // quad_perm [1,0,3,2]: exchange with lane ^ 1
  int lo = __double2loint(v), hi = __double2hiint(v);
  lo = __builtin_amdgcn_update_dpp(0, lo, 0xB1, 0xF, 0xF, false);
  hi = __builtin_amdgcn_update_dpp(0, hi, 0xB1, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}
__device__ __forceinline__ float swap_pair(float v)
{
  int b = __float_as_int(v);
  b = __builtin_amdgcn_update_dpp(0, b, 0xB1, 0xF, 0xF, false);
  return __int_as_float(b);
}

template<typename P>
__device__ __forceinline__ P lane_read(P v, int src_lane)
{
  return __shfl(v, src_lane, wave_size);
}

//! Final combine of a row: y = alpha*s + beta*y + gamma*xg.
template<typename P>
__device__ __forceinline__ void store_row(kron_args<P> const &a, int64_t pos, P s)
{
  P out = a.alpha * s;
  if (a.beta != P{0})
    out += a.beta * a.y[pos];
  if (a.gamma != P{0})
    out += a.gamma * a.xg[pos];
  a.y[pos] = out;
}

} // namespace akm
