// C ABI of the kronmult matrix: construction, apply, timing, destruction.
#include "matrix.hpp"

#include "grid.hpp"

#include <algorithm>
#include <cstring>
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

int lanes_per_column(akm_prec prec, int d, int n, int force_generic)
{
  // mirrors lead_dims() + the register-kernel table of kron_kernels.hip;
  // 0 means the block-wide LDS kernel
  static const int reg_cases[][2] = {{1, 1}, {1, 2}, {1, 3}, {1, 4}, {1, 5}, {2, 1}, {2, 2},
                                     {2, 3}, {2, 4}, {2, 5}, {3, 1}, {3, 2}, {3, 3}, {3, 4},
                                     {3, 5}, {4, 1}, {4, 2}, {4, 3}, {4, 4}, {5, 1}, {5, 2},
                                     {5, 3}, {6, 1}, {6, 2}, {6, 3}};
  if (force_generic)
    return 0;
  bool listed = false;
  for (auto const &c : reg_cases)
    if (c[0] == d && c[1] == n)
      listed = true;
  if (!listed)
    return 0;
  const int budget = (prec == AKM_F64) ? 32 : 64;
  int R = ipow_int(n, d), G = 1;
  while (R > budget)
  {
    R /= n;
    G *= n;
  }
  return (G <= wave_size) ? G : 0;
}

//! Cuts every row's column range into chunks of about `chunk_columns`
//! columns (a multiple of the columns a wavefront processes per step).
static int build_chunks(akm_matrix_s &m)
{
  const int G = lanes_per_column(m.prec, m.d, m.n, m.force_generic);
  const int64_t cpw = (G > 0) ? wave_size / G : 1;
  int64_t cc = m.chunk_columns;
  if (cc <= 0)
  {
    // about 16 wavefronts per SIMD (1024 SIMDs), 1 to 64 steps per chunk
    const int64_t target = 16 * 1024;
    int64_t steps = (m.nnz + cpw * target - 1) / (cpw * target);
    steps = std::max<int64_t>(1, std::min<int64_t>(steps, 64));
    if (G == 0)
      steps = std::max<int64_t>(1, std::min<int64_t>(m.nnz / (4 * 1024), 32));
    cc = cpw * steps;
  }
  cc = std::max<int64_t>(cpw, (cc + cpw - 1) / cpw * cpw);

  std::vector<int> crow, rbeg(m.rows + 1, 0);
  std::vector<int64_t> cb, ce;
  for (int r = 0; r < m.rows; r++)
  {
    const int64_t b = (m.mode == mode_dense) ? 0 : m.pntr[r];
    const int64_t e = (m.mode == mode_dense) ? m.cols : m.pntr[r + 1];
    rbeg[r] = static_cast<int>(crow.size());
    const int64_t len = e - b;
    const int64_t nch = std::max<int64_t>(1, (len + cc - 1) / cc);
    int64_t piece = (len + nch - 1) / nch;
    piece = std::max<int64_t>(cpw, (piece + cpw - 1) / cpw * cpw);
    int64_t s = b;
    do
    {
      const int64_t f = std::min(e, s + piece);
      crow.push_back(r);
      cb.push_back(s);
      ce.push_back(f);
      s = f;
    } while (s < e);
    if (static_cast<int64_t>(crow.size()) > (int64_t(1) << 31) - 1)
      return set_error(AKM_ERR_INVALID, "kronmult: too many work chunks");
  }
  rbeg[m.rows] = static_cast<int>(crow.size());
  m.num_chunks = static_cast<int>(crow.size());
  AKM_TRY(m.chunk_row.upload(crow.data(), crow.size(), m.stream));
  AKM_TRY(m.chunk_beg.upload(cb.data(), cb.size(), m.stream));
  AKM_TRY(m.chunk_end.upload(ce.data(), ce.size(), m.stream));
  AKM_TRY(m.row_chunk_begin.upload(rbeg.data(), rbeg.size(), m.stream));
  AKM_TRY(m.counters.allocate(m.rows));
  AKM_HIP(hipMemsetAsync(m.counters.data(), 0, sizeof(int) * m.rows, m.stream));
  AKM_TRY(m.partial.allocate(int64_t(m.num_chunks) * m.N * int64_t(m.esize)));
  AKM_HIP(hipStreamSynchronize(m.stream));
  m.chunks_ready = true;
  return AKM_SUCCESS;
}

template<typename P>
static int apply_typed(akm_matrix_s &m, P alpha, const P *x, P beta, P *y, P gamma,
                       const P *xg)
{
  if (m.T == 0 || alpha == P{0} || m.rows == 0)
    return launch_scale<P>(m.output_size(), beta, y, gamma, xg, m.stream);
  if (!m.chunks_ready)
    AKM_TRY(build_chunks(m));
  kron_args<P> a{};
  a.vA = reinterpret_cast<const P *>(m.vA.data());
  a.iA = m.iA.data();
  a.indx = m.indx.data();
  a.elem = m.elem.data();
  a.off1d = m.off1d.data();
  a.M = m.M;
  a.terms = reinterpret_cast<const P *const *>(m.term_ptrs.data());
  a.nb = m.nb;
  a.row_offset = m.row_offset;
  a.col_offset = m.col_offset;
  a.T = m.T;
  a.num_chunks = m.num_chunks;
  a.chunk_row = m.chunk_row.data();
  a.chunk_beg = m.chunk_beg.data();
  a.chunk_end = m.chunk_end.data();
  a.row_chunk_begin = m.row_chunk_begin.data();
  a.counters = m.counters.data();
  a.partial = reinterpret_cast<P *>(m.partial.data());
  a.x = x;
  a.y = y;
  a.xg = xg;
  a.alpha = alpha;
  a.beta = beta;
  a.gamma = gamma;
  return launch_kronmult<P>(m.d, m.n, m.mode, a, m.force_generic, m.stream);
}

int apply_device(akm_matrix_s &m, double alpha, const void *x, double beta, void *y,
                 double gamma, const void *xg)
{
  if (gamma != 0 && m.rows != m.cols)
    return set_error(AKM_ERR_INVALID, "identity term needs a square operator");
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
  AKM_TRY(build_chunks(*m));
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
  AKM_TRY(device_sparse_pattern(d, num_cells, table, m->pntr, m->indx));
  m->nnz = m->pntr[num_cells];
  std::vector<int> elem;
  cell_1d_indices(d, num_cells, table, elem);
  AKM_TRY(m->elem.upload(elem.data(), int64_t(elem.size())));
  AKM_TRY(m->off1d.upload(offtab.data(), int64_t(offtab.size())));
  m->M = 1 << max_level;
  m->num_vA = num_vA;
  if (T > 0)
    AKM_TRY(m->vA.upload(static_cast<const unsigned char *>(vA), num_vA * m->esize));
  AKM_TRY(build_chunks(*m));
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
  AKM_TRY(build_chunks(*m));
  *mat = m.release();
  return AKM_SUCCESS;
}

akm_status akm_update_coefficients(akm_matrix mat, const void *values)
{
  if (mat == nullptr || values == nullptr)
    return set_error(AKM_ERR_INVALID, "update coefficients: null argument");
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
  else if (mat->num_vA > 0)
  {
    AKM_HIP(hipMemcpyAsync(mat->vA.data(), values, mat->num_vA * mat->esize,
                           hipMemcpyHostToDevice, mat->stream));
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

akm_status akm_set_tuning(akm_matrix mat, int64_t chunk_columns, int force_generic)
{
  if (mat == nullptr)
    return set_error(AKM_ERR_INVALID, "null matrix");
  AKM_HIP(hipStreamSynchronize(mat->stream));
  mat->chunk_columns = chunk_columns;
  mat->force_generic = force_generic;
  mat->chunks_ready = false;
  return build_chunks(*mat);
}

akm_status akm_time_apply(akm_matrix mat, double alpha, const void *x, double beta, void *y,
                          int iterations, double *ms_per_apply)
{
  if (mat == nullptr || ms_per_apply == nullptr || iterations <= 0)
    return set_error(AKM_ERR_INVALID, "time_apply: invalid argument");
  AKM_TRY(apply_device(*mat, alpha, x, beta, y, 0.0, nullptr)); // builds chunks
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

akm_status akm_destroy(akm_matrix mat)
{
  if (mat == nullptr)
    return AKM_SUCCESS;
  (void)hipStreamSynchronize(mat->stream);
  delete mat;
  return AKM_SUCCESS;
}

} // extern "C"
