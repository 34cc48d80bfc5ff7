/*
 * CPU oracle for the sparse-grid kronmult apply and the implicit-step GMRES.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
 * The product (asgard_amd/) never links, loads or calls anything here.
 *
 * Every function restates the reference algorithm it names (file:line in
 * project-asgard/asgard).  How each one is pinned is listed beside it:
 *   - element table order     : reference fixtures element_table/{table,ids}_*.dat
 *   - connect_1d              : known answers of src/asgard_indexset_tests.cpp:83-155
 *   - check_connected / CSR   : symmetry, diagonal, agreement with connect_1d
 *                               (every connected pair has a 1D offset), brute
 *                               force over the 1D supports, nnz totals
 *   - kronmult (sparse/dense) : explicit Kronecker product + gemv, the
 *                               known answers of asgard_kronmult_matrix_tests.cpp:178-193
 *   - GMRES                   : 3x3 gold of src/solver_tests.cpp:141-213 and a
 *                               dense LU solve of (I - dt A) x = b
 */
#ifndef ASGARD_ORACLE_H
#define ASGARD_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- grid --- */

/* Number of level tuples selected by get_lequal_multi (sparse) or
 * get_max_multi (full grid); full_grid != 0 selects the latter. */
int64_t orc_num_level_tuples(int d, const int *levels, int full_grid);

/* Writes the level tuples (num_tuples x d, row-major) in the order of
 * elements::table, src/elements.cpp:254-281. */
int orc_level_tuples(int d, const int *levels, int full_grid, int *tuples);

/* Number of cells of the element table. */
int64_t orc_num_cells(int d, const int *levels, int full_grid);

/* Element table: table[c*2d + k] = level of dim k, table[c*2d + d + k] = cell
 * position of dim k; ids[c] = elements::map_to_id with max_level (may be NULL).
 * Restates elements::table::table, src/elements.cpp:249-350 and
 * get_cell_index_set, src/elements.cpp:408-465. */
int orc_element_table(int d, const int *levels, int full_grid, int max_level,
                      int *table, int64_t *ids);

/* connect_1d(max_level, mode) of src/asgard_grid_1d.hpp:159-317.
 * mode 0 = hierarchy::volume, 1 = hierarchy::full.
 * Returns the number of connections; pntr (2^L + 1), indx, diag (2^L) are
 * written when non-NULL (query the size with NULL first). */
int orc_connect_1d(int max_level, int mode, int *pntr, int *indx, int *diag);

/* connect_1d::get_offset, src/asgard_grid_1d.hpp:92-122 (binary search). */
int orc_connect_1d_offset(const int *pntr, const int *indx, int row, int col);

/* check_connected of src/asgard_kronmult_matrix.cpp:230-300. */
int orc_check_connected_1d(int L1, int p1, int L2, int p2);
int orc_check_connected(int d, const int *row_coords, const int *col_coords);

/* Connectivity scan of compute_mem_usage (src/asgard_kronmult_matrix.cpp:918-938):
 * row pointer of the cell-pair CSR (num_cells + 1 entries), returns nnz. */
int64_t orc_sparse_pattern_count(int d, int64_t num_cells, const int *table,
                                 int64_t *pntr);

/* Column lists of make_kronmult_sparse (src/asgard_kronmult_matrix.cpp:406-466):
 * indx[nnz] = column cell ids; iA (nnz*T*d, may be NULL) =
 * offset_1d * (d*T*n*n) + (t*d + dim) * n*n. Needs pntr from the count. */
int orc_sparse_pattern_fill(int d, int n, int T, int max_level, int64_t num_cells,
                            const int *table, const int64_t *pntr, int *indx,
                            int *iA);

/* Same as the fill above for a subset of row cells only (rows[num_sel]):
 * pntr (num_sel + 1), indx and iA of those rows; indx/iA may be NULL to
 * count.  Returns the number of pairs or a negative code. */
int64_t orc_sparse_pattern_rows(int d, int n, int T, int max_level,
                                int64_t num_cells, const int *table,
                                const int *rows, int num_sel, int64_t *pntr,
                                int *indx, int *iA);

/* 1D index of every cell and dimension: elem[c*d + k] = 0 if l == 0 else
 * 2^(l-1) + p  (src/asgard_kronmult_matrix.cpp:121-132). */
void orc_cell_1d_index(int d, int64_t num_cells, const int *table, int *elem);

/* vA of make_kronmult_sparse (src/asgard_kronmult_matrix.cpp:369-393):
 * for every connect_1d(L) connection (row, col), for t, for dim, the n x n
 * block of ops[t*d + dim] (column-major, leading dimension lda) at block
 * (row, col), stored column-major.  vA has num_connections*d*T*n*n entries. */
void orc_build_vA(int d, int n, int T, int max_level, const double *const *ops,
                  int64_t lda, double *vA);

/* Dense-mode term arrays of make_kronmult_dense (src/asgard_kronmult_matrix.cpp:98-114):
 * terms[t][dim*nb*nb*n*n + n*n*(ocol*nb + orow) + i*n + s] = op(n*orow+s, n*ocol+i). */
void orc_build_dense_terms(int d, int n, int T, int num_1d_blocks,
                           const double *const *ops, int64_t lda,
                           double *const *terms);

/* ------------------------------------------------------------ kronmult --- */

/* kronmult::cpu_sparse, src/device/asgard_spkronmult_cpu.cpp:84-433 (n > 1)
 * and cpu_n0 (:26-60, n == 1), dispatch :606-768.  OpenMP over rows.
 * y[row] = beta*y[row] + alpha * sum_{j in row} sum_t (A_{t,0} x ... x A_{t,d-1}) x[indx[j]] */
int orc_kronmult_sparse(int d, int n, int num_rows, const int64_t *pntr,
                        const int *indx, int T, const int *iA,
                        const double *vA, double alpha, const double *x,
                        double beta, double *y);
int orc_kronmult_sparse_f32(int d, int n, int num_rows, const int64_t *pntr,
                            const int *indx, int T, const int *iA,
                            const float *vA, float alpha, const float *x,
                            float beta, float *y);

/* kronmult::cpu_dense, src/device/asgard_kronmult_cpu.cpp:158-511, 655-818:
 * all row/col pairs, factor for dim k = terms[t][k*nb*nb*n*n + n*n*(elem[col*d+k]*nb + elem[row*d+k])]. */
int orc_kronmult_dense(int d, int n, int num_rows, int num_cols, int T,
                       const int *elem, int row_offset, int col_offset,
                       const double *const *terms, int num_1d_blocks,
                       double alpha, const double *x, double beta, double *y);

/* Explicit Kronecker product (src/asgard_kronmult_tests.hpp:39-58):
 * result = A (m x m) kron B (n x n), column-major. */
void orc_kronecker(int m, const double *A, int n, const double *B, double *result);

/* reference_kronmult_one (src/asgard_kronmult_tests.hpp:63-76):
 * y += (A_0 x ... x A_{d-1}) x with the product formed explicitly. */
void orc_kronmult_one_bruteforce(int d, int n, const double *const *A,
                                 const double *x, double *y);

/* ---------------------------------------------------------------- GMRES --- */

typedef void (*orc_matvec_fn)(void *ctx, double alpha, const double *x,
                              double beta, double *y);
typedef void (*orc_precond_fn)(void *ctx, double *x);

/* solver::simple_gmres, src/solver.cpp:156-298: restarted GMRES(restart),
 * single-pass classical Gram-Schmidt, Givens rotations, packed triangular
 * back-solve, absolute tolerance on the (preconditioned) residual estimate.
 * precond may be NULL (no_op_preconditioner).  Returns 0 or a negative
 * code (-1: restart > n).  err/iterations mirror gmres_info. */
int orc_simple_gmres(int n, orc_matvec_fn mat, void *mat_ctx,
                     orc_precond_fn precond, void *pc_ctx, double *x,
                     const double *b, int restart, int max_outer, double tol,
                     double *err, int *iterations);

/* simple_gmres with a dense column-major matrix and an optional dense
 * left preconditioner M (LU-factored, solver.cpp:11-38); M may be NULL. */
int orc_simple_gmres_dense(int n, const double *A, const double *M, double *x,
                           const double *b, int restart, int max_outer,
                           double tol, double *err, int *iterations);

/* simple_gmres_euler (src/solver.cpp:119-135): solves (I - dt*A) x = b with
 * A the sparse kronmult operator above (rows == cols). */
int orc_gmres_euler_sparse(int d, int n, int num_cells, const int64_t *pntr,
                           const int *indx, int T, const int *iA,
                           const double *vA, double dt, double *x,
                           const double *b, int restart, int max_outer,
                           double tol, double *err, int *iterations);

/* Dense LU with partial pivoting solve (gesv), used as the gold for GMRES. */
int orc_dense_solve(int n, double *A, double *b);

/* Number of OpenMP threads the kronmult uses. */
int orc_num_threads(void);
void orc_set_num_threads(int n);

#ifdef __cplusplus
}
#endif

#endif
