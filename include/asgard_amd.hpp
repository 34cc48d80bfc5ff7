/*
 * asgard_amd: the C++ call surface of ASGarD's kronmult path over the C ABI
 * of include/asgard_amd.h (header only; link -lasgard_amd).
 *
 * It keeps the names, argument meaning and error behaviour of the reference
 * (project-asgard/asgard, KRON_MODE_GLOBAL off):
 *   asgard::kronmult_matrix<P>       src/asgard_kronmult_matrix.hpp:111-682
 *   asgard::matrix_list<P>           src/asgard_kronmult_matrix.hpp:1041-1279
 *   asgard::matrix_entry             src/asgard_kronmult_matrix.hpp:736-744
 *   asgard::gmres_info<P>            src/pde/pde_base.hpp:94-98
 *   asgard::solver::simple_gmres_euler  src/solver.hpp:34-40, src/solver.cpp:119-135
 *   asgard::solver::simple_gmres (dense) src/solver.hpp:17-22, src/solver.cpp:54-74
 *   asgard::solver::default_gmres_restarts  src/solver.cpp:142-152
 * Errors become the reference's exception types: std::invalid_argument for
 * invalid arguments (restart > n), std::runtime_error otherwise.
 *
 * Vectors are any contiguous container with data() and size() (the
 * reference's fk::vector, std::vector, ...); `resource` says whether their
 * data() is host or device memory.  Device memory is not allocated here:
 * the library owns every device buffer of a matrix.
 */
#ifndef ASGARD_AMD_HPP
#define ASGARD_AMD_HPP

#include "asgard_amd.h"

#include <cstdint>
#include <functional>
#include <limits>
#include <memory>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

namespace asgard
{
//! Where the data of a vector lives (the reference's resource).
enum class resource
{
  host,
  device
};

//! Expressive indexing for the matrices of a matrix_list.
enum matrix_entry
{
  regular = 0,
  imex_explicit = 1,
  imex_implicit = 2
};

//! Outcome of a GMRES solve: the final residual estimate and the number of
//! operator applications.
template<typename P>
struct gmres_info
{
  P error;
  int iterations;
};

//! The reference's "no user value" sentinels (src/program_options.hpp:207-208).
constexpr int no_user_value = -1;
constexpr double no_user_value_fp = std::numeric_limits<double>::min();

//! Throws the reference's exception type for a failed call.
inline void akm_check(akm_status s)
{
  if (s == AKM_SUCCESS || s == AKM_ERR_NOT_CONVERGED)
    return;
  char msg[1024];
  akm_last_error(msg, sizeof(msg));
  if (s == AKM_ERR_INVALID)
    throw std::invalid_argument(msg);
  throw std::runtime_error(msg);
}

template<typename precision>
class kronmult_matrix
{
  static_assert(std::is_same_v<precision, double> || std::is_same_v<precision, float>,
                "kronmult_matrix is FP64 or FP32");

public:
  static constexpr akm_prec akm_precision =
      std::is_same_v<precision, double> ? AKM_F64 : AKM_F32;

  //! An unset matrix (evaluates to false).
  kronmult_matrix() = default;

  //! A matrix with no terms: apply gives y = beta*y (hpp:123-130).
  kronmult_matrix(int num_dimensions, int kron_size, int num_rows, int num_cols)
      : d_(num_dimensions), n_(kron_size), rows_(num_rows), cols_(num_cols)
  {
    std::vector<int64_t> pntr(num_rows + 1, 0);
    akm_matrix h = nullptr;
    akm_check(akm_create_sparse(d_, n_, rows_, cols_, 0, pntr.data(), nullptr, nullptr, nullptr,
                                0, akm_precision, &h));
    adopt(h);
  }

  //! Sparse single-call constructor, CPU format (hpp:235-272): row_indx is
  //! the CSR row pointer (num_rows + 1), col_indx the column cell ids,
  //! index_A the factor offsets into values_A, (pair*T + t)*d + dim.
  kronmult_matrix(int num_dimensions, int kron_size, int num_rows, int num_cols, int num_terms,
                  std::vector<int> &&row_indx, std::vector<int> &&col_indx,
                  std::vector<int> &&index_A, std::vector<precision> &&values_A)
      : d_(num_dimensions), n_(kron_size), rows_(num_rows), cols_(num_cols), T_(num_terms)
  {
    if (static_cast<int>(row_indx.size()) != num_rows + 1)
      throw std::invalid_argument("kronmult_matrix: row_indx needs num_rows + 1 entries");
    std::vector<int64_t> pntr(row_indx.begin(), row_indx.end());
    akm_matrix h = nullptr;
    akm_check(akm_create_sparse(d_, n_, rows_, cols_, T_, pntr.data(), col_indx.data(),
                                index_A.data(), values_A.data(),
                                static_cast<int64_t>(values_A.size()), akm_precision, &h));
    adopt(h);
  }

  //! Dense constructor (hpp:132-176): every (row, column) pair; terms[t]
  //! holds d*nb*nb*n*n factor values of term t, elem the 1D indices.
  kronmult_matrix(int num_dimensions, int kron_size, int num_rows, int num_cols, int num_terms,
                  std::vector<std::vector<precision>> &&terms, std::vector<int> &&elem,
                  int row_offset, int col_offset, int num_1d_blocks)
      : d_(num_dimensions), n_(kron_size), rows_(num_rows), cols_(num_cols), T_(num_terms),
        dense_(true)
  {
    std::vector<const void *> tp(terms.size());
    for (size_t t = 0; t < terms.size(); t++)
      tp[t] = terms[t].data();
    akm_matrix h = nullptr;
    akm_check(akm_create_dense(d_, n_, rows_, cols_, T_, tp.data(), elem.data(),
                               static_cast<int>(elem.size()) / num_dimensions, row_offset,
                               col_offset, num_1d_blocks, akm_precision, &h));
    adopt(h);
  }

  //! make_kronmult_sparse (src/asgard_kronmult_matrix.cpp:330-690) from the
  //! element table (cells x 2d ints, get_active_table()) and the vA block
  //! array it fills; the cell-pair list is built on the GPU.
  static kronmult_matrix sparse_grid(int num_dimensions, int kron_size,
                                     std::vector<int> const &table, int max_level, int num_terms,
                                     std::vector<precision> const &values_A)
  {
    kronmult_matrix m;
    const int64_t cells = static_cast<int64_t>(table.size()) / (2 * num_dimensions);
    m.d_ = num_dimensions;
    m.n_ = kron_size;
    m.rows_ = m.cols_ = static_cast<int>(cells);
    m.T_ = num_terms;
    akm_matrix h = nullptr;
    akm_check(akm_create_sparse_grid(num_dimensions, kron_size, cells, table.data(), max_level,
                                     num_terms, values_A.data(),
                                     static_cast<int64_t>(values_A.size()), akm_precision, &h));
    m.adopt(h);
    return m;
  }

  //! y = alpha*A*x + beta*y (hpp:340-478); host vectors go through the
  //! matrix's own device workspaces, device vectors are used in place.
  template<resource rec = resource::host>
  void apply(precision alpha, precision const x[], precision beta, precision y[]) const
  {
    akm_check(akm_apply(handle(), alpha, x, beta, y, rec == resource::device ? 1 : 0));
  }

  //! Replaces the factor values keeping the index lists (hpp:520-565).
  void update_stored_coefficients(std::vector<precision> &&values_A)
  {
    akm_check(akm_update_coefficients(handle(), values_A.data()));
  }
  //! Dense matrices: new term arrays (one per term).
  void update_stored_coefficients(std::vector<std::vector<precision>> &&values_A)
  {
    std::vector<const void *> tp(values_A.size());
    for (size_t t = 0; t < values_A.size(); t++)
      tp[t] = values_A[t].data();
    akm_check(akm_update_coefficients(handle(), tp.data()));
  }

  int num_dimensions() const { return d_; }
  int kron_size() const { return n_; }
  int num_terms() const { return T_; }
  int64_t num_batch() const { return int64_t{rows_} * int64_t{cols_}; }
  int tensor_size() const { return static_cast<int>(info(0)); }
  int input_size() const { return static_cast<int>(info(1)); }
  int output_size() const { return static_cast<int>(info(2)); }
  //! Connected (row, column) pairs the apply visits.
  int64_t num_pairs() const { return info(3); }
  //! Flops per apply as the reference's flops() counts them (without the
  //! factor 2 of compute_flops).
  int64_t flops() const { return info(4) / 2; }
  //! compute_flops (hpp:507-512).
  static int64_t compute_flops(int num_dimensions, int kron_size, int num_terms,
                               int64_t num_batch)
  {
    int64_t t = kron_size;
    for (int d = 1; d < num_dimensions; d++)
      t *= kron_size;
    return t * kron_size * num_dimensions * num_terms * num_batch * 2;
  }
  bool is_dense() const { return dense_; }
  explicit operator bool() const { return d_ > 0 && handle_ != nullptr; }

  //! The library handle (for the solvers and akm_* calls).
  akm_matrix handle() const
  {
    if (!handle_)
      throw std::runtime_error("kronmult_matrix: the matrix is not set");
    return handle_.get();
  }

private:
  void adopt(akm_matrix h) { handle_.reset(h, [](akm_matrix p) { akm_destroy(p); }); }
  int64_t info(int which) const
  {
    int64_t v[5] = {0, 0, 0, 0, 0};
    akm_check(akm_matrix_info(handle(), &v[0], &v[1], &v[2], &v[3], &v[4]));
    return v[which];
  }

  int d_ = 0, n_ = 0, rows_ = 0, cols_ = 0, T_ = 0;
  bool dense_ = false;
  std::shared_ptr<akm_matrix_s> handle_;
};

/*!
 * The three matrices of a time step (regular, IMEX explicit, IMEX implicit)
 * grouped as the reference groups them.  The reference builds a matrix from
 * its PDE and grid inside make(); here the caller passes that step as a
 * factory, and the coefficient refresh of reset_coefficients as an update.
 * Each matrix owns its device workspaces (the reference's xdev/ydev).
 */
template<typename precision>
struct matrix_list
{
  using matrix_type = kronmult_matrix<precision>;
  using factory_type = std::function<matrix_type()>;
  using update_type = std::function<void(matrix_type &)>;

  matrix_list() : matrices(3) {}

  matrix_type &operator[](matrix_entry entry) { return matrices[static_cast<int>(entry)]; }

  template<resource rec = resource::host>
  void apply(matrix_entry entry, precision alpha, precision const x[], precision beta,
             precision y[])
  {
    matrices[static_cast<int>(entry)].template apply<rec>(alpha, x, beta, y);
  }

  int64_t flops(matrix_entry entry) { return matrices[static_cast<int>(entry)].flops(); }

  //! Makes the matrix of `entry` unless it is already set.
  void make(matrix_entry entry, factory_type const &factory)
  {
    if (!(*this)[entry])
      (*this)[entry] = factory();
  }

  //! Makes the matrix when unset, otherwise refreshes only its coefficients.
  void reset_coefficients(matrix_entry entry, factory_type const &factory,
                          update_type const &update)
  {
    if (!(*this)[entry])
      make(entry, factory);
    else
      update((*this)[entry]);
  }

  void clear(matrix_entry entry)
  {
    if (matrices[static_cast<int>(entry)])
      matrices[static_cast<int>(entry)] = matrix_type();
  }
  void clear_all()
  {
    for (auto &m : matrices)
      if (m)
        m = matrix_type();
  }

  std::vector<matrix_type> matrices;
};

namespace solver
{
//! clamp(512 MB / size, min(10, n), min(200, n)) (src/solver.cpp:142-152).
template<typename P>
int default_gmres_restarts(int num_cols)
{
  const int minimum = std::min(10, num_cols);
  const int maximum = std::min(200, num_cols);
  const double mb = double(num_cols) * sizeof(P) / (1024.0 * 1024.0);
  const int guess = static_cast<int>(512.0 / mb);
  return std::max(minimum, std::min(guess, maximum));
}

namespace detail
{
template<typename P>
double tolerance_arg(P tolerance)
{
  return (static_cast<double>(tolerance) == no_user_value_fp || tolerance <= P{0})
             ? -1.0
             : static_cast<double>(tolerance);
}
inline int count_arg(int v) { return (v == no_user_value || v <= 0) ? -1 : v; }
} // namespace detail

//! Solves (I - dt*mat) x = b with restarted GMRES, x holding the initial
//! guess; restart, max_iter and tolerance take the reference's sentinels
//! for their defaults.  Not converging is reported in the result, not thrown.
template<typename P, resource resrc, typename vector_type>
gmres_info<P> simple_gmres_euler(const P dt, kronmult_matrix<P> const &mat, vector_type &x,
                                 vector_type const &b, int const restart, int const max_iter,
                                 P const tolerance)
{
  if (static_cast<int64_t>(x.size()) != mat.output_size() ||
      static_cast<int64_t>(b.size()) != mat.output_size())
    throw std::invalid_argument("simple_gmres_euler: x and b must match the matrix size");
  double err = 0;
  int iters = 0;
  akm_check(akm_gmres_euler(mat.handle(), dt, x.data(), b.data(), detail::count_arg(restart),
                            detail::count_arg(max_iter), detail::tolerance_arg(tolerance),
                            resrc == resource::device ? 1 : 0, &err, &iters));
  return gmres_info<P>{static_cast<P>(err), iters};
}

//! The node-local test version of GMRES on an explicit column-major n x n
//! matrix A, with M (empty = none) as left preconditioner through its LU
//! factors.  Host containers; n = x.size().
template<typename P, typename matrix_type, typename vector_type>
gmres_info<P> simple_gmres(matrix_type const &A, vector_type &x, vector_type const &b,
                           matrix_type const &M, int const restart, int const max_iter,
                           P const tolerance)
{
  const int n = static_cast<int>(x.size());
  if (static_cast<int64_t>(A.size()) != int64_t(n) * n || static_cast<int>(b.size()) != n)
    throw std::invalid_argument("simple_gmres: A must be n x n and b of size n");
  if (M.size() != 0 && static_cast<int64_t>(M.size()) != int64_t(n) * n)
    throw std::invalid_argument("simple_gmres: M must be n x n or empty");
  constexpr akm_prec prec = std::is_same_v<P, double> ? AKM_F64 : AKM_F32;
  double err = 0;
  int iters = 0;
  akm_check(akm_gmres_dense(n, A.data(), x.data(), b.data(), M.size() ? M.data() : nullptr,
                            detail::count_arg(restart), detail::count_arg(max_iter),
                            detail::tolerance_arg(tolerance), prec, &err, &iters));
  return gmres_info<P>{static_cast<P>(err), iters};
}
} // namespace solver
} // namespace asgard

#endif
