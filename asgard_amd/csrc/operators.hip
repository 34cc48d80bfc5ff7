// Operator update on the GPU: the stored 1D factor blocks of a kronmult
// handle rebuilt from dense 1D operator matrices that are already on the
// device, with no host copy of the blocks (update_kronmult_coefficients,
// src/asgard_kronmult_matrix.cpp:710-811).
//
// Sparse-grid and global handles store one n x n block per (connect_1d
// connection, term, dimension) in the vA layout [connection][term][dim][n*n]
// (make_kronmult_sparse, src/asgard_kronmult_matrix.cpp:369-393); dense
// handles store per term [dim][column block][row block][n*n].  Every block is
// column-major: entry (s, c) of the block of 1D cells (row, col) is
// A(n*row + s, n*col + c).  One thread per stored value; the reads are
// gathers from the operator matrices, the writes coalesced.
#include "akm_internal.hpp"
#include "grid.hpp"
#include "matrix.hpp"

#include <vector>

namespace akm
{
constexpr int op_threads = 256;

static int op_blocks(int64_t work)
{
  const int64_t b = (work + op_threads - 1) / op_threads;
  return static_cast<int>(b < 8192 ? (b > 0 ? b : 1) : 8192);
}

//! vA[((j*T + t)*d + k)*n*n + c*n + s] = A_{t,k}(n*row_j + s, n*col_j + c).
template<typename P>
__global__ void __launch_bounds__(op_threads)
    gather_grid_blocks(const P *const *ops, int64_t lda, const int *conn_row, const int *conn_col,
                       int64_t nconn, int T, int d, int n, P *vA)
{
  const int nn = n * n;
  const int64_t total = nconn * T * d * nn;
  for (int64_t idx = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; idx < total;
       idx += int64_t(gridDim.x) * blockDim.x)
  {
    const int q = static_cast<int>(idx % nn);
    int64_t rest = idx / nn;
    const int k = static_cast<int>(rest % d);
    rest /= d;
    const int t = static_cast<int>(rest % T);
    const int64_t j = rest / T;
    const int c = q / n, s = q - c * n;
    vA[idx] = ops[t * d + k][int64_t(n) * conn_row[j] + s + lda * (int64_t(n) * conn_col[j] + c)];
  }
}

//! The LDS-table copy of the factors: [dim][connection][term][nnp], the
//! padding of every block zero.
template<typename P>
__global__ void __launch_bounds__(op_threads)
    relayout_blocks(const P *vA, int64_t nconn, int T, int d, int n, int nnp, P *vAk)
{
  const int nn = n * n;
  const int64_t total = int64_t(d) * nconn * T * nnp;
  for (int64_t idx = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; idx < total;
       idx += int64_t(gridDim.x) * blockDim.x)
  {
    const int q = static_cast<int>(idx % nnp);
    int64_t rest = idx / nnp;
    const int t = static_cast<int>(rest % T);
    rest /= T;
    const int64_t j = rest % nconn;
    const int k = static_cast<int>(rest / nconn);
    vAk[idx] = (q < nn) ? vA[((j * T + t) * d + k) * nn + q] : P(0);
  }
}

//! Dense handles, term t: term[((k*nb + ocol)*nb + orow)*n*n + c*n + s] =
//! A_{t,k}(n*orow + s, n*ocol + c) (src/asgard_kronmult_matrix.cpp:741-754).
template<typename P>
__global__ void __launch_bounds__(op_threads)
    gather_dense_blocks(const P *const *ops, int64_t lda, int nb, int d, int n, int t, P *term)
{
  const int nn = n * n;
  const int64_t total = int64_t(d) * nb * nb * nn;
  for (int64_t idx = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; idx < total;
       idx += int64_t(gridDim.x) * blockDim.x)
  {
    const int q = static_cast<int>(idx % nn);
    int64_t rest = idx / nn;
    const int orow = static_cast<int>(rest % nb);
    rest /= nb;
    const int ocol = static_cast<int>(rest % nb);
    const int k = static_cast<int>(rest / nb);
    const int c = q / n, s = q - c * n;
    term[idx] = ops[t * d + k][int64_t(n) * orow + s + lda * (int64_t(n) * ocol + c)];
  }
}

template<typename P>
static int update_operators_typed(akm_matrix_s &m, const void *const *ops_dev, int64_t lda)
{
  const P *const *ops = reinterpret_cast<const P *const *>(ops_dev);
  const int d = m.d, n = m.n, T = m.T;
  if (m.mode == mode_dense)
  {
    const int64_t work = int64_t(d) * m.nb * m.nb * n * n;
    for (int t = 0; t < T; t++)
    {
      gather_dense_blocks<P><<<op_blocks(work), op_threads, 0, m.stream>>>(
          ops, lda, m.nb, d, n, t, reinterpret_cast<P *>(m.terms[t].data()));
      AKM_TRY(check_launch("operator update (dense blocks)"));
    }
    return AKM_SUCCESS;
  }
  const int64_t nconn = m.op_conn.size() / 2;
  P *vA = reinterpret_cast<P *>(m.vA.data());
  gather_grid_blocks<P><<<op_blocks(m.num_vA), op_threads, 0, m.stream>>>(
      ops, lda, m.op_conn.data(), m.op_conn.data() + nconn, nconn, T, d, n, vA);
  AKM_TRY(check_launch("operator update (factor blocks)"));
  if (m.mode == mode_grid && m.vAk.data() != nullptr)
  {
    relayout_blocks<P><<<op_blocks(int64_t(d) * nconn * T * m.nnp), op_threads, 0, m.stream>>>(
        vA, nconn, T, d, n, m.nnp, reinterpret_cast<P *>(m.vAk.data()));
    AKM_TRY(check_launch("operator update (LDS tables)"));
  }
  if (m.mode == mode_global)
    AKM_TRY(global_device_jacobi(m));
  return AKM_SUCCESS;
}

} // namespace akm

using namespace akm;

akm_status akm_update_operators(akm_matrix mat, const void *const *ops, int64_t lda)
{
  if (mat == nullptr)
    return set_error(AKM_ERR_INVALID, "operator update: null handle");
  if (mat->T == 0)
    return AKM_SUCCESS;
  if (ops == nullptr)
    return set_error(AKM_ERR_INVALID, "operator update: null operator list");
  mat->vA_stamp++;
  const int nops = mat->T * mat->d;
  for (int i = 0; i < nops; i++)
    if (ops[i] == nullptr)
      return set_error(AKM_ERR_INVALID, "operator update: operator %d is null", i);
  int64_t need = 0;
  if (mat->mode == mode_dense)
    need = int64_t(mat->n) * mat->nb;
  else if (mat->mode == mode_grid || mat->mode == mode_global)
    need = int64_t(mat->n) << mat->max_level;
  else
    return set_error(AKM_ERR_UNSUPPORTED,
                     "operator update: needs a sparse-grid, global or dense handle (this one "
                     "was made from an explicit index list)");
  if (lda < need)
    return set_error(AKM_ERR_INVALID, "operator update: lda %lld below the operator size %lld",
                     static_cast<long long>(lda), static_cast<long long>(need));
  if (mat->mode != mode_dense && mat->op_conn.size() == 0)
  {
    // connect_1d as (row, column) per connection, in vA order
    std::vector<int> p1, i1, dg;
    AKM_TRY(build_connect_1d(mat->max_level, true, p1, i1, dg));
    const int64_t nconn = static_cast<int64_t>(i1.size());
    if (nconn * mat->T * mat->d * mat->n * mat->n != mat->num_vA)
      return set_error(AKM_ERR_INVALID, "operator update: stored blocks do not match connect_1d");
    std::vector<int> rc(2 * nconn);
    for (int r = 0; r + 1 < static_cast<int>(p1.size()); r++)
      for (int j = p1[r]; j < p1[r + 1]; j++)
      {
        rc[j] = r;
        rc[nconn + j] = i1[j];
      }
    AKM_TRY(mat->op_conn.upload(rc.data(), 2 * nconn, mat->stream));
  }
  std::vector<const void *> ptrs(ops, ops + nops);
  AKM_TRY(mat->op_ptrs.upload(ptrs.data(), nops, mat->stream));
  if (mat->prec == AKM_F64)
    AKM_TRY(update_operators_typed<double>(*mat, mat->op_ptrs.data(), lda));
  else
    AKM_TRY(update_operators_typed<float>(*mat, mat->op_ptrs.data(), lda));
  AKM_HIP(hipStreamSynchronize(mat->stream));
  return AKM_SUCCESS;
}
