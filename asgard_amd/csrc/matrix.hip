// C ABI of the kronmult matrix: construction, apply, timing, destruction.
#include "matrix.hpp"

#include "grid.hpp"

#include <algorithm>
#include <cstring>
#include <map>
#include <mutex>
#include <string>

namespace akm
{
namespace
{
thread_local std::string last_error;
} // namespace

int set_error(int code, const char *fmt, ...)
{
  char buf[1024];
  va_list args;
  va_start(args, fmt);
  vsnprintf(buf, sizeof(buf), fmt, args);
  va_end(args);
  last_error = buf;
  return code;
}

int check_launch(const char *what)
{
  hipError_t e = hipGetLastError();
  if (e != hipSuccess)
    return set_error(AKM_ERR_HIP, "%s: kernel launch failed: %s", what, hipGetErrorString(e));
  return AKM_SUCCESS;
}

static int ipow_int(int n, int d)
{
  int r = 1;
  for (int k = 0; k < d; k++)
    r *= n;
  return r;
}

int waves_per_row(akm_matrix_s const &m)
{
  // 16 wavefronts per row exist only for the matrix-core kernels
  if (m.waves > 0)
    return (m.waves > 8 && m.family != 1) ? 8 : m.waves;
  // enough wavefronts to cover the 1024 SIMDs several times over
  const int64_t target = 8 * 1024;
  int w = 1;
  while (w < 8 && int64_t(m.rows) * w < target)
    w *= 2;
  return w;
}

//! Row dispatch order: longest column list first (stable), so the heaviest
//! workgroups start first and the tail of the launch is short.
static int build_schedule(akm_matrix_s &m)
{
  std::vector<int> order(m.rows);
  for (int r = 0; r < m.rows; r++)
    order[r] = r;
  if (m.mode != mode_dense)
  {
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
      return (m.pntr[a + 1] - m.pntr[a]) > (m.pntr[b + 1] - m.pntr[b]);
    });
    AKM_TRY(m.pntr_dev.upload(m.pntr.data(), int64_t(m.pntr.size()), m.stream));
  }
  AKM_TRY(m.row_order.upload(order.data(), int64_t(order.size()), m.stream));
  return AKM_SUCCESS;
}

//! Re-lays the reference's vA ([connection][term][dim][n*n]) out as
//! [dim][connection][term][nnp] for the LDS-table kernel: the blocks one
//! row's 1D cell can touch in one dimension become one contiguous range.
template<typename P>
static void relayout_vA(akm_matrix_s const &m, const P *vA, std::vector<P> &out)
{
  const int d = m.d, T = m.T, nn = m.n * m.n, nnp = m.nnp;
  out.assign(size_t(d) * m.nconn * T * nnp, P{0});
  for (int64_t j = 0; j < m.nconn; j++)
    for (int t = 0; t < T; t++)
      for (int k = 0; k < d; k++)
      {
        const P *src = vA + ((j * T + t) * d + k) * nn;
        P *dst = out.data() + ((int64_t(k) * m.nconn + j) * T + t) * nnp;
        for (int q = 0; q < nn; q++)
          dst[q] = src[q];
      }
}

static int upload_vAk(akm_matrix_s &m, const void *vA)
{
  if (m.prec == AKM_F64)
  {
    std::vector<double> v;
    relayout_vA(m, static_cast<const double *>(vA), v);
    return m.vAk.upload(reinterpret_cast<const unsigned char *>(v.data()),
                        int64_t(v.size() * sizeof(double)), m.stream);
  }
  std::vector<float> v;
  relayout_vA(m, static_cast<const float *>(vA), v);
  return m.vAk.upload(reinterpret_cast<const unsigned char *>(v.data()),
                      int64_t(v.size() * sizeof(float)), m.stream);
}

//! LDS bytes the sparse-grid kernel may use: the whole 160 KB of a CU minus
//! a margin unless the caller set a cap.
static int lds_cap_bytes(akm_matrix_s const &m)
{
  if (m.lds_cap >= 0)
    return m.lds_cap;
  return 160 * 1024 - 4096;
}

template<typename P>
static kron_args<P> matrix_args(akm_matrix_s const &m)
{
  kron_args<P> a{};
  a.vA = reinterpret_cast<const P *>(m.vA.data());
  a.iA = m.iA.data();
  a.pntr = m.pntr_dev.data();
  a.indx = m.indx.data();
  a.elem = m.elem.data();
  a.off1d = m.off1d.data();
  a.M = m.M;
  a.terms = reinterpret_cast<const P *const *>(m.term_ptrs.data());
  a.nb = m.nb;
  a.row_offset = m.row_offset;
  a.col_offset = m.col_offset;
  a.cols = m.cols;
  a.T = m.T;
  a.vAk = reinterpret_cast<const P *>(m.vAk.data());
  a.conn_pntr = m.conn_pntr.data();
  a.epack = m.epack.data();
  a.nconn = m.nconn;
  a.nnp = m.nnp;
  a.max_table_elems = m.max_table_elems;
  a.max_table_last4 = m.max_table_last4;
  a.lds_cap = (m.vAk.data() != nullptr) ? lds_cap_bytes(m) : 0;
  a.row_order = m.row_order.data();
  a.family = m.family;
  return a;
}

kron_args<double> kron_args_f64(akm_matrix_s const &m) { return matrix_args<double>(m); }

} // namespace akm

int64_t akm_matrix_s::flops() const
{
  if (mode == akm::mode_global)
    return akm::global_flops(*this);
  return N * n * d * int64_t(T) * nnz * 2;
}

namespace akm
{

template<typename P>
static int apply_typed(akm_matrix_s &m, P alpha, const P *x, P beta, P *y, P gamma,
                       const P *xg)
{
  if (m.T == 0 || alpha == P{0} || m.rows == 0)
    return launch_scale<P>(m.output_size(), beta, y, gamma, xg, m.stream);
  kron_args<P> a = matrix_args<P>(m);
  a.x = x;
  a.y = y;
  a.xg = xg;
  a.alpha = alpha;
  a.beta = beta;
  a.gamma = gamma;
  return launch_kronmult<P>(m.d, m.n, m.mode, a, m.rows, waves_per_row(m), m.force_generic,
                            m.stream);
}

int apply_device(akm_matrix_s &m, double alpha, const void *x, double beta, void *y,
                 double gamma, const void *xg)
{
  if (gamma != 0 && m.rows != m.cols)
    return set_error(AKM_ERR_INVALID, "identity term needs a square operator");
  if (m.mode == mode_global)
    return global_apply(m, alpha, x, beta, y, gamma, xg, false, 0.0);
  if (m.prec == AKM_F64)
    return apply_typed<double>(m, alpha, static_cast<const double *>(x), beta,
                               static_cast<double *>(y), gamma,
                               static_cast<const double *>(xg));
  return apply_typed<float>(m, static_cast<float>(alpha), static_cast<const float *>(x),
                            static_cast<float>(beta), static_cast<float *>(y),
                            static_cast<float>(gamma), static_cast<const float *>(xg));
}

static int check_common(int d, int n, int rows, int cols, int T, akm_prec prec,
                        akm_matrix *mat)
{
  if (mat == nullptr)
    return set_error(AKM_ERR_INVALID, "null matrix handle pointer");
  *mat = nullptr;
  if (d < 1 || d > 6)
    return set_error(AKM_ERR_UNSUPPORTED, "kronmult: d = %d not in [1, 6]", d);
  if (n < 1)
    return set_error(AKM_ERR_INVALID, "kronmult: n = %d must be positive", n);
  if (ipow_int(n, d) > 4096)
    return set_error(AKM_ERR_UNSUPPORTED, "kronmult: n^d = %d exceeds 4096", ipow_int(n, d));
  if (rows < 0 || cols < 0 || T < 0)
    return set_error(AKM_ERR_INVALID, "kronmult: negative size");
  if (prec != AKM_F64 && prec != AKM_F32)
    return set_error(AKM_ERR_INVALID, "kronmult: unknown precision %d", int(prec));
  return AKM_SUCCESS;
}

static akm_matrix_s *new_matrix(int d, int n, int rows, int cols, int T, akm_prec prec)
{
  auto *m = new akm_matrix_s();
  m->prec = prec;
  m->esize = (prec == AKM_F64) ? 8 : 4;
  m->d = d;
  m->n = n;
  m->rows = rows;
  m->cols = cols;
  m->T = T;
  m->N = ipow_int(n, d);
  return m;
}

} // namespace akm

using namespace akm;

extern "C" {

int akm_last_error(char *buffer, size_t size)
{
  const std::string &e = akm::last_error;
  if (buffer != nullptr && size > 0)
  {
    size_t k = std::min(size - 1, e.size());
    std::memcpy(buffer, e.data(), k);
    buffer[k] = '\0';
  }
  return static_cast<int>(e.size());
}

const char *akm_version(void) { return "asgard_amd 0.1.0 (gfx950)"; }

akm_status akm_device_count(int *count)
{
  if (count == nullptr)
    return set_error(AKM_ERR_INVALID, "null count");
  int c = 0;
  hipError_t e = hipGetDeviceCount(&c);
  *count = (e == hipSuccess) ? c : 0;
  return AKM_SUCCESS;
}

akm_status akm_set_stream(akm_matrix mat, void *stream)
{
  if (mat == nullptr)
    return set_error(AKM_ERR_INVALID, "null matrix");
  mat->stream = static_cast<hipStream_t>(stream);
  // tables derived from vA are rebuilt on the new stream
  mat->vA_stamp++;
  return AKM_SUCCESS;
}

akm_status akm_synchronize(akm_matrix mat)
{
  if (mat == nullptr)
    return set_error(AKM_ERR_INVALID, "null matrix");
  AKM_HIP(hipStreamSynchronize(mat->stream));
  return AKM_SUCCESS;
}

akm_status akm_connect_1d(int max_level, int full, int *pntr, int *indx, int *diag,
                          int *num_connections)
{
  std::vector<int> p, i, dg;
  AKM_TRY(build_connect_1d(max_level, full != 0, p, i, dg));
  if (num_connections)
    *num_connections = static_cast<int>(i.size());
  if (pntr)
    std::copy(p.begin(), p.end(), pntr);
  if (indx)
    std::copy(i.begin(), i.end(), indx);
  if (diag)
    std::copy(dg.begin(), dg.end(), diag);
  return AKM_SUCCESS;
}

akm_status akm_sparse_pattern(int d, int64_t num_cells, const int *table, int64_t *pntr,
                              int *indx)
{
  if (table == nullptr || pntr == nullptr)
    return set_error(AKM_ERR_INVALID, "sparse pattern: null table or pntr");
  if (d < 1 || d > 6)
    return set_error(AKM_ERR_UNSUPPORTED, "sparse pattern: d = %d not in [1, 6]", d);
  std::vector<int64_t> p;
  dev_buffer<int> idx;
  AKM_TRY(device_sparse_pattern(d, num_cells, table, p, idx));
  std::copy(p.begin(), p.end(), pntr);
  if (indx != nullptr && p.back() > 0)
    AKM_HIP(hipMemcpy(indx, idx.data(), sizeof(int) * p.back(), hipMemcpyDeviceToHost));
  return AKM_SUCCESS;
}

akm_status akm_create_sparse(int d, int n, int rows, int cols, int T, const int64_t *pntr,
                             const int *indx, const int *iA, const void *vA, int64_t num_vA,
                             akm_prec prec, akm_matrix *mat)
{
  AKM_TRY(check_common(d, n, rows, cols, T, prec, mat));
  if (pntr == nullptr || (rows > 0 && pntr[rows] > 0 && (indx == nullptr || iA == nullptr)))
    return set_error(AKM_ERR_INVALID, "sparse kronmult: null index arrays");
  if (T > 0 && (vA == nullptr || num_vA <= 0))
    return set_error(AKM_ERR_INVALID, "sparse kronmult: empty vA");
  const int64_t nnz = pntr[rows];
  for (int r = 0; r < rows; r++)
    if (pntr[r + 1] < pntr[r])
      return set_error(AKM_ERR_INVALID, "sparse kronmult: row pointer not monotone");
  // every index must stay inside x and vA
  for (int64_t j = 0; j < nnz; j++)
    if (indx[j] < 0 || indx[j] >= cols)
      return set_error(AKM_ERR_INVALID, "sparse kronmult: column %d out of range", indx[j]);
  const int64_t nn = int64_t(n) * n;
  for (int64_t q = 0; q < nnz * T * d; q++)
    if (iA[q] < 0 || iA[q] + nn > num_vA)
      return set_error(AKM_ERR_INVALID, "sparse kronmult: iA entry %d outside vA", iA[q]);

  std::unique_ptr<akm_matrix_s> m(new_matrix(d, n, rows, cols, T, prec));
  m->mode = mode_ia;
  m->nnz = nnz;
  m->pntr.assign(pntr, pntr + rows + 1);
  AKM_TRY(m->indx.upload(indx, nnz));
  AKM_TRY(m->iA.upload(iA, nnz * T * d));
  m->num_vA = num_vA;
  AKM_TRY(m->vA.upload(static_cast<const unsigned char *>(vA), num_vA * m->esize));
  AKM_TRY(build_schedule(*m));
  *mat = m.release();
  return AKM_SUCCESS;
}

akm_status akm_create_sparse_grid(int d, int n, int64_t num_cells, const int *table,
                                  int max_level, int T, const void *vA, int64_t num_vA,
                                  akm_prec prec, akm_matrix *mat)
{
  if (num_cells > (int64_t(1) << 31) - 2)
    return set_error(AKM_ERR_INVALID, "sparse-grid kronmult: too many cells");
  AKM_TRY(check_common(d, n, int(num_cells), int(num_cells), T, prec, mat));
  if (table == nullptr)
    return set_error(AKM_ERR_INVALID, "sparse-grid kronmult: null table");
  int lmax = 0;
  for (int64_t c = 0; c < num_cells; c++)
    for (int k = 0; k < d; k++)
    {
      const int L = table[c * 2 * d + k], p = table[c * 2 * d + d + k];
      if (L < 0 || p < 0 || (L <= 1 && p != 0) || (L > 1 && p >= (1 << (L - 1))))
        return set_error(AKM_ERR_INVALID, "sparse-grid kronmult: invalid cell (%d, %d)", L, p);
      lmax = std::max(lmax, L);
    }
  if (lmax > max_level)
    return set_error(AKM_ERR_INVALID, "sparse-grid kronmult: level %d above max_level %d",
                     lmax, max_level);
  std::vector<int> offtab, p1, i1, dg;
  AKM_TRY(build_offset_table(max_level, offtab));
  AKM_TRY(build_connect_1d(max_level, true, p1, i1, dg));
  const int64_t nconn = static_cast<int64_t>(i1.size());
  const int64_t expect_vA = nconn * d * T * int64_t(n) * n;
  if (T > 0 && (vA == nullptr || num_vA != expect_vA))
    return set_error(AKM_ERR_INVALID,
                     "sparse-grid kronmult: vA has %lld values, connect_1d(%d) needs %lld",
                     static_cast<long long>(num_vA), max_level,
                     static_cast<long long>(expect_vA));

  std::unique_ptr<akm_matrix_s> m(new_matrix(d, n, int(num_cells), int(num_cells), T, prec));
  m->mode = mode_grid;
  std::vector<int> elem, lex;
  cell_1d_indices(d, num_cells, table, elem);
  // each row's columns in lexicographic order of their 1D indices: the
  // sum-factorised kernels close a group of columns when its prefix changes
  lexicographic_order(d, num_cells, elem.data(), lex);
  AKM_TRY(device_sparse_pattern(d, num_cells, table, m->pntr, m->indx, lex.data()));
  m->nnz = m->pntr[num_cells];
  AKM_TRY(m->elem.upload(elem.data(), int64_t(elem.size())));
  AKM_TRY(m->off1d.upload(offtab.data(), int64_t(offtab.size())));
  m->M = 1 << max_level;
  m->max_level = max_level;
  m->num_vA = num_vA;
  m->nconn = static_cast<int>(nconn);
  m->conn_pntr_host = p1;
  if (T > 0)
  {
    AKM_TRY(m->vA.upload(static_cast<const unsigned char *>(vA), num_vA * m->esize));
    if (max_level <= 10)
    {
      // LDS-table kernel inputs: packed cell words, connect_1d row pointer,
      // factors re-laid out per dimension, largest per-row table
      std::vector<unsigned long long> pack(num_cells);
      for (int64_t c = 0; c < num_cells; c++)
      {
        unsigned long long w = 0;
        for (int k = 0; k < d; k++)
          w |= static_cast<unsigned long long>(elem[c * d + k]) << (10 * k);
        pack[c] = w;
      }
      AKM_TRY(m->epack.upload(pack.data(), num_cells));
      AKM_TRY(m->conn_pntr.upload(p1.data(), int64_t(p1.size())));
      m->nnp = (m->prec == AKM_F64) ? padded_block<double>(n) : padded_block<float>(n);
      AKM_TRY(upload_vAk(*m, vA));
      int64_t mx = 0, mx2 = 0;
      for (int64_t c = 0; c < num_cells; c++)
      {
        int64_t s = 0, s2 = 0;
        for (int k = 0; k < d; k++)
        {
          const int r = elem[c * d + k];
          const int64_t len = int64_t(p1[r + 1] - p1[r]) * T * m->nnp;
          s += len;
          if (k >= d - 4)
            s2 += len;
        }
        mx = std::max(mx, s);
        mx2 = std::max(mx2, s2);
      }
      m->max_table_elems = mx;
      m->max_table_last4 = mx2;
    }
  }
  AKM_TRY(build_schedule(*m));
  *mat = m.release();
  return AKM_SUCCESS;
}

akm_status akm_create_dense(int d, int n, int rows, int cols, int T, const void *const *terms,
                            const int *elem, int num_indexes, int row_offset, int col_offset,
                            int num_1d_blocks, akm_prec prec, akm_matrix *mat)
{
  AKM_TRY(check_common(d, n, rows, cols, T, prec, mat));
  if (elem == nullptr || (T > 0 && terms == nullptr))
    return set_error(AKM_ERR_INVALID, "dense kronmult: null inputs");
  if (row_offset < 0 || col_offset < 0 || row_offset + rows > num_indexes ||
      col_offset + cols > num_indexes)
    return set_error(AKM_ERR_INVALID, "dense kronmult: offsets outside elem");
  for (int64_t q = 0; q < int64_t(num_indexes) * d; q++)
    if (elem[q] < 0 || elem[q] >= num_1d_blocks)
      return set_error(AKM_ERR_INVALID, "dense kronmult: elem entry %d out of range", elem[q]);
  std::unique_ptr<akm_matrix_s> m(new_matrix(d, n, rows, cols, T, prec));
  m->mode = mode_dense;
  m->nnz = int64_t(rows) * cols;
  m->nb = num_1d_blocks;
  m->row_offset = row_offset;
  m->col_offset = col_offset;
  AKM_TRY(m->elem.upload(elem, int64_t(num_indexes) * d));
  const int64_t tsize = int64_t(d) * num_1d_blocks * num_1d_blocks * n * n;
  m->terms.resize(T);
  std::vector<void *> ptrs(T);
  for (int t = 0; t < T; t++)
  {
    AKM_TRY(m->terms[t].upload(static_cast<const unsigned char *>(terms[t]), tsize * m->esize));
    ptrs[t] = m->terms[t].data();
  }
  if (T > 0)
    AKM_TRY(m->term_ptrs.upload(ptrs.data(), T));
  AKM_TRY(build_schedule(*m));
  *mat = m.release();
  return AKM_SUCCESS;
}

akm_status akm_update_coefficients(akm_matrix mat, const void *values)
{
  if (mat == nullptr || values == nullptr)
    return set_error(AKM_ERR_INVALID, "update coefficients: null argument");
  mat->vA_stamp++;
  if (mat->mode == mode_dense)
  {
    const void *const *terms = static_cast<const void *const *>(values);
    const int64_t tsize = int64_t(mat->d) * mat->nb * mat->nb * mat->n * mat->n;
    for (int t = 0; t < mat->T; t++)
    {
      AKM_HIP(hipMemcpyAsync(mat->terms[t].data(), terms[t], tsize * mat->esize,
                             hipMemcpyHostToDevice, mat->stream));
    }
  }
  else if (mat->mode == mode_global)
  {
    if (mat->num_vA > 0)
      AKM_TRY(global_update(*mat, values));
  }
  else if (mat->num_vA > 0)
  {
    AKM_HIP(hipMemcpyAsync(mat->vA.data(), values, mat->num_vA * mat->esize,
                           hipMemcpyHostToDevice, mat->stream));
    if (mat->vAk.data() != nullptr)
      AKM_TRY(upload_vAk(*mat, values));
  }
  AKM_HIP(hipStreamSynchronize(mat->stream));
  return AKM_SUCCESS;
}

akm_status akm_apply(akm_matrix mat, double alpha, const void *x, double beta, void *y,
                     int on_device)
{
  if (mat == nullptr || x == nullptr || y == nullptr)
    return set_error(AKM_ERR_INVALID, "apply: null argument");
  if (on_device)
    return apply_device(*mat, alpha, x, beta, y, 0.0, nullptr);
  // host vectors: stage through the matrix's device workspaces
  const int64_t nx = mat->input_size() * mat->esize, ny = mat->output_size() * mat->esize;
  AKM_TRY(mat->xws.allocate(nx));
  AKM_TRY(mat->yws.allocate(ny));
  AKM_HIP(hipMemcpyAsync(mat->xws.data(), x, nx, hipMemcpyHostToDevice, mat->stream));
  if (beta != 0)
    AKM_HIP(hipMemcpyAsync(mat->yws.data(), y, ny, hipMemcpyHostToDevice, mat->stream));
  AKM_TRY(apply_device(*mat, alpha, mat->xws.data(), beta, mat->yws.data(), 0.0, nullptr));
  AKM_HIP(hipMemcpyAsync(y, mat->yws.data(), ny, hipMemcpyDeviceToHost, mat->stream));
  AKM_HIP(hipStreamSynchronize(mat->stream));
  return AKM_SUCCESS;
}

akm_status akm_apply_euler(akm_matrix mat, double dt, double alpha, const void *x, double beta,
                           void *y)
{
  if (mat == nullptr || x == nullptr || y == nullptr)
    return set_error(AKM_ERR_INVALID, "apply_euler: null argument");
  return apply_device(*mat, -dt * alpha, x, beta, y, alpha, x);
}

akm_status akm_matrix_info(akm_matrix mat, int64_t *tensor_size, int64_t *input_size,
                           int64_t *output_size, int64_t *num_pairs, int64_t *flops)
{
  if (mat == nullptr)
    return set_error(AKM_ERR_INVALID, "null matrix");
  if (tensor_size)
    *tensor_size = mat->N;
  if (input_size)
    *input_size = mat->input_size();
  if (output_size)
    *output_size = mat->output_size();
  if (num_pairs)
    *num_pairs = mat->nnz;
  if (flops)
    *flops = mat->flops();
  return AKM_SUCCESS;
}

akm_status akm_set_tuning(akm_matrix mat, int waves_per_row, int lds_cap_bytes,
                          int force_generic)
{
  if (mat == nullptr)
    return set_error(AKM_ERR_INVALID, "null matrix");
  if (mat->mode == mode_global)
    return global_set_layout(*mat, waves_per_row);
  if (waves_per_row != 0 && waves_per_row != 1 && waves_per_row != 2 && waves_per_row != 4 &&
      waves_per_row != 8 && waves_per_row != 16)
    return set_error(AKM_ERR_INVALID, "waves per row must be 0, 1, 2, 4, 8 or 16 (got %d)",
                     waves_per_row);
  if (lds_cap_bytes > 160 * 1024)
    return set_error(AKM_ERR_INVALID, "LDS cap %d above the 160 KB of a CU", lds_cap_bytes);
  mat->waves = waves_per_row;
  mat->lds_cap = (lds_cap_bytes < 0) ? -1 : lds_cap_bytes;
  mat->force_generic = force_generic;
  return AKM_SUCCESS;
}

akm_status akm_time_apply(akm_matrix mat, double alpha, const void *x, double beta, void *y,
                          int iterations, double *ms_per_apply)
{
  if (mat == nullptr || ms_per_apply == nullptr || iterations <= 0)
    return set_error(AKM_ERR_INVALID, "time_apply: invalid argument");
  AKM_TRY(apply_device(*mat, alpha, x, beta, y, 0.0, nullptr)); // warm-up
  hipEvent_t start, stop;
  AKM_HIP(hipEventCreate(&start));
  AKM_HIP(hipEventCreate(&stop));
  AKM_HIP(hipEventRecord(start, mat->stream));
  for (int i = 0; i < iterations; i++)
    AKM_TRY(apply_device(*mat, alpha, x, beta, y, 0.0, nullptr));
  AKM_HIP(hipEventRecord(stop, mat->stream));
  AKM_HIP(hipEventSynchronize(stop));
  float ms = 0;
  AKM_HIP(hipEventElapsedTime(&ms, start, stop));
  AKM_HIP(hipEventDestroy(start));
  AKM_HIP(hipEventDestroy(stop));
  *ms_per_apply = double(ms) / iterations;
  return AKM_SUCCESS;
}

akm_status akm_autotune(akm_matrix mat, const void *x, void *y, int *waves_per_row,
                        int *lds_tables)
{
  if (mat == nullptr || x == nullptr || y == nullptr)
    return set_error(AKM_ERR_INVALID, "autotune: null argument");
  if (mat->mode == mode_global)
  {
    // row layout (reported in waves_per_row as akm_set_tuning takes it)
    AKM_TRY(global_autotune(*mat, x, y));
    if (waves_per_row)
      *waves_per_row = global_layout(*mat);
    if (lds_tables)
      *lds_tables = 0;
    return AKM_SUCCESS;
  }
  struct choice
  {
    int waves, lds, family;
  };
  static std::mutex mtx;
  static std::map<std::vector<int64_t>, choice> cache;
  const std::vector<int64_t> key = {int64_t(mat->prec), mat->d,      mat->n,  mat->T,
                                    mat->max_level,     mat->rows,   mat->nnz, int64_t(mat->mode)};
  choice best{0, -1, 0};
  bool cached = false;
  {
    std::lock_guard<std::mutex> lock(mtx);
    auto it = cache.find(key);
    if (it != cache.end())
    {
      best = it->second;
      cached = true;
    }
  }
  if (!cached)
  {
    // enough applies per candidate for ~2 ms of timing, at most 50
    double probe = 0;
    mat->waves = 4;
    mat->lds_cap = -1;
    mat->family = 0;
    AKM_TRY(akm_time_apply(mat, 1.0, x, 0.0, y, 1, &probe));
    const int iters = std::max(1, std::min(50, int(2.0 / std::max(probe, 1e-3))));
    double best_ms = 1e300;
    const int lds_opts[2] = {-1, 0};
    for (int li = 0; li < (mat->vAk.data() != nullptr ? 2 : 1); li++)
      for (int w : {1, 2, 4, 8})
      {
        mat->waves = w;
        mat->lds_cap = (mat->vAk.data() != nullptr) ? lds_opts[li] : 0;
        double ms = 0;
        AKM_TRY(akm_time_apply(mat, 1.0, x, 0.0, y, iters, &ms));
        if (ms < best_ms)
        {
          best_ms = ms;
          best = {w, mat->lds_cap, 0};
        }
      }
    if (mat->prec == AKM_F64 &&
        mfma_family_applies(mat->d, mat->n, mat->mode, mat->T, mat->epack.data() != nullptr))
    {
      const int saved_lds = mat->lds_cap;
      mat->family = 1;
      for (int w : {1, 2, 4, 8, 16})
      {
        mat->waves = w;
        double ms = 0;
        AKM_TRY(akm_time_apply(mat, 1.0, x, 0.0, y, iters, &ms));
        if (ms < best_ms)
        {
          best_ms = ms;
          best = {w, saved_lds, 1};
        }
      }
      mat->family = 0;
    }
    std::lock_guard<std::mutex> lock(mtx);
    cache[key] = best;
  }
  mat->waves = best.waves;
  mat->lds_cap = best.lds;
  mat->family = best.family;
  if (waves_per_row)
    *waves_per_row = best.waves;
  if (lds_tables)
    *lds_tables = (best.family == 1) ? 2 : (best.lds != 0 && mat->vAk.data() != nullptr) ? 1 : 0;
  return AKM_SUCCESS;
}

akm_status akm_set_kernel_family(akm_matrix mat, int family)
{
  if (mat == nullptr)
    return set_error(AKM_ERR_INVALID, "kernel family: null matrix");
  if (family != 0 && family != 1)
    return set_error(AKM_ERR_INVALID, "kernel family must be 0 or 1 (got %d)", family);
  if (family == 1 && !(mat->prec == AKM_F64 &&
                       mfma_family_applies(mat->d, mat->n, mat->mode, mat->T,
                                           mat->epack.data() != nullptr)))
    return set_error(AKM_ERR_UNSUPPORTED,
                     "kernel family 1 needs an FP64 sparse-grid matrix with d = 3, n = 3, "
                     "1 to 8 terms and max_level <= 10");
  mat->family = family;
  return AKM_SUCCESS;
}

akm_status akm_destroy(akm_matrix mat)
{
  if (mat == nullptr)
    return AKM_SUCCESS;
  (void)hipStreamSynchronize(mat->stream);
  delete mat;
  return AKM_SUCCESS;
}

} // extern "C"
