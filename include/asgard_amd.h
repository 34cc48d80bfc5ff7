/*
 * asgard_amd: MI355X (gfx950) kronmult operator apply and implicit-step GMRES.
 *
 * C ABI.  Plain pointers and sizes; no exceptions cross it.  Every entry point
 * returns an akm_status (0 on success); akm_last_error() describes the most
 * recent failure of the calling thread.
 *
 * Each declaration names the reference interface (project-asgard/asgard) it
 * replaces.  Precision: a matrix is created in FP64 or FP32 (akm_prec) and
 * every x, y, b, vA pointer passed with it has that element type; scalars are
 * passed as double and converted.
 *
 * Vectors: "on_device" != 0 means x/y/b are device pointers (the reference's
 * apply<resource::device>); 0 means host pointers, copied in and out through
 * device workspaces owned by the matrix, exactly as the reference's
 * apply<resource::host> does with its xdev/ydev workspaces
 * (src/asgard_kronmult_matrix.hpp:351-358, 451-452).
 *
 * Threading: a matrix is not thread-safe (the reference's apply is not,
 * src/asgard_kronmult_matrix.hpp:336).  Work is enqueued on the matrix's
 * stream (default: the null stream) and is asynchronous for device pointers.
 */
#ifndef ASGARD_AMD_H
#define ASGARD_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int akm_status;
#define AKM_SUCCESS 0
#define AKM_ERR_INVALID -1     /* invalid argument (std::invalid_argument) */
#define AKM_ERR_UNSUPPORTED -2 /* unsupported (d, n) (std::runtime_error) */
#define AKM_ERR_HIP -3         /* HIP runtime failure */
#define AKM_ERR_MEMORY -4      /* device allocation failure */
#define AKM_ERR_NOT_CONVERGED 1 /* GMRES stopped above the tolerance (info) */

typedef enum
{
  AKM_F64 = 0,
  AKM_F32 = 1
} akm_prec;

typedef struct akm_matrix_s *akm_matrix;

/* ------------------------------------------------------------ utilities -- */

/* Copies the calling thread's last error message; returns its length. */
int akm_last_error(char *buffer, size_t size);

/* Library version string. */
const char *akm_version(void);

/* Number of visible HIP devices (0 when none). */
akm_status akm_device_count(int *count);

/* Selects the stream the matrix enqueues work on (a hipStream_t, NULL = the
 * null stream).  The reference always used the default stream. */
akm_status akm_set_stream(akm_matrix mat, void *stream);

/* Blocks until the matrix's stream is idle. */
akm_status akm_synchronize(akm_matrix mat);

/* --------------------------------------------------- index construction -- */

/* connect_1d(max_level, full ? hierarchy::full : hierarchy::volume)
 * (src/asgard_grid_1d.hpp:34-46, 159-317): CSR of the 1D cell hierarchy.
 * pntr has 2^max_level + 1 entries, indx *num_connections, diag 2^max_level;
 * call with NULL arrays to query num_connections.  Host arrays. */
akm_status akm_connect_1d(int max_level, int full, int *pntr, int *indx, int *diag,
                          int *num_connections);

/* Cell-pair sparsity of the kronmult operator on the GPU: replaces the
 * O(cells^2) check_connected scans of compute_mem_usage and
 * make_kronmult_sparse (src/asgard_kronmult_matrix.cpp:918-938, 411-466).
 * table is the flattened element table (num_cells x 2d host ints:
 * levels then cell positions, elements::table::get_active_table()).
 * Writes pntr (num_cells + 1, int64, host) and, when indx is not NULL, the
 * column cell ids (host, nnz entries, ascending in each row).  Call once with
 * indx = NULL to get nnz = pntr[num_cells], then again with indx. */
akm_status akm_sparse_pattern(int d, int64_t num_cells, const int *table, int64_t *pntr,
                              int *indx);

/* ----------------------------------------------------- matrix lifecycle -- */

/* Sparse single-call constructor of kronmult_matrix<P>
 * (src/asgard_kronmult_matrix.hpp:235-272, CPU format): pntr (rows + 1) and
 * indx (nnz) are the CSR of column cell ids, iA (nnz*T*d) the factor offsets
 * into vA (index (pair*T + t)*d + dim, dim 0 applied to the slowest tensor
 * index), vA (num_vA values) the n x n column-major blocks.  All host arrays;
 * the matrix copies them to the device (the reference moves them in). */
akm_status akm_create_sparse(int d, int n, int rows, int cols, int num_terms,
                             const int64_t *pntr, const int *indx, const int *iA,
                             const void *vA, int64_t num_vA, akm_prec prec,
                             akm_matrix *mat);

/* Sparse-grid factory: make_kronmult_sparse (src/asgard_kronmult_matrix.cpp:330-690)
 * without the per-pair iA list.  From the element table (num_cells x 2d host
 * ints) it builds the cell-pair CSR on the GPU and derives every factor offset
 * on the fly from the cells' 1D indices and connect_1d(max_level).  vA is the
 * reference's block array: for each connect_1d(max_level) connection, for
 * each term, for each dim, one n x n column-major block (num_connections*d*T*n*n
 * values, as make_kronmult_sparse fills it at :369-393).  Host arrays. */
akm_status akm_create_sparse_grid(int d, int n, int64_t num_cells, const int *table,
                                  int max_level, int num_terms, const void *vA,
                                  int64_t num_vA, akm_prec prec, akm_matrix *mat);

/* Global Kronecker factory: make_global_kron_matrix + set_specific_mode
 * (src/asgard_kronmult_matrix.cpp:1198-1426), the reference's default
 * single-node operator (global_kron_matrix<P>, src/asgard_kronmult_matrix.hpp:780-989).
 * Same table and vA as akm_create_sparse_grid.  active (num_terms x d host
 * ints, NULL = all) marks the non-identity factors: a factor marked 0 is
 * taken as the identity, as the reference does for pure mass terms
 * (check_identity_term, :1175-1184).  flux_dim (num_terms host ints, NULL =
 * none) is the reference's per-term flux direction, which it moves to the
 * front of the active list (:1313-1331).  The cell set is completed with its
 * missing ancestors (compute_ancestry_completion); the operator is applied
 * one dimension at a time (global_cpu, src/device/asgard_glkronmult_cpu.cpp:151-189).
 * The handle works with akm_apply, akm_apply_euler, akm_update_coefficients
 * (which also rebuilds the Jacobi diagonal), akm_time_apply, akm_matrix_info
 * (flops counted as the reference does, num_pairs = 0) and the solvers; its
 * akm_gmres_euler is the reference's left-Jacobi-preconditioned solve
 * (simple_gmres_euler for global_kron_matrix, src/solver.cpp:94-118). */
akm_status akm_create_global_grid(int d, int n, int64_t num_cells, const int *table,
                                  int max_level, int num_terms, const void *vA,
                                  int64_t num_vA, const int *active, const int *flux_dim,
                                  akm_prec prec, akm_matrix *mat);

/* Global operator sizes: padding cells added to the grid, chains (one per
 * upper/lower choice of each term), stages per apply.  No counterpart in the
 * reference (its constructor prints the padding count). */
akm_status akm_global_info(akm_matrix mat, int64_t *padded_cells, int *chains, int *stages);

/* The slab layout of a global operator (akm_set_tuning layout 6): term
 * groups handled by the slab kernels (terms with the same one or two
 * non-identity dimensions), chains left to the cell-layout stages, and the
 * vectors the final sum reads (group sums, remaining chains' results, x for
 * identity terms).  No counterpart in the reference. */
akm_status akm_global_slab_info(akm_matrix mat, int *groups, int *rest_chains,
                                int *sum_sources);

/* The Jacobi diagonal of a global operator (build_preconditioner,
 * src/asgard_kronmult_matrix.cpp:1016-1067), one FP64 value per active DOF,
 * copied to the host array pc. */
akm_status akm_global_jacobi(akm_matrix mat, double *pc);

/* Dense constructor of kronmult_matrix<P>
 * (src/asgard_kronmult_matrix.hpp:132-176): every row/column pair, term t's
 * factors in terms[t] (d*nb*nb*n*n values each, block (ocol, orow) of dim k at
 * k*nb*nb*n*n + n*n*(ocol*nb + orow)), elem the cells' 1D indices
 * (num_indexes x d), row/col offsets into elem.  Host arrays. */
akm_status akm_create_dense(int d, int n, int rows, int cols, int num_terms,
                            const void *const *terms, const int *elem, int num_indexes,
                            int row_offset, int col_offset, int num_1d_blocks,
                            akm_prec prec, akm_matrix *mat);

/* update_stored_coefficients (src/asgard_kronmult_matrix.hpp:520-565):
 * replaces vA (sparse) or the term arrays (dense; terms is then a
 * const void *const * of num_terms arrays) keeping the index lists. */
akm_status akm_update_coefficients(akm_matrix mat, const void *values);

/* update_kronmult_coefficients (src/asgard_kronmult_matrix.cpp:710-811) on
 * the GPU: the stored factor blocks are gathered on the device from the
 * dense 1D operator matrices, with no host copy.  ops is a host array of
 * num_terms * d device pointers; ops[t*d + k] is the lda x lda column-major
 * operator of term t in dimension k (in the handle's precision), lda >=
 * n * 2^max_level for sparse-grid and global handles, >= n * num_1d_blocks
 * for dense ones.  Global handles also get their Jacobi diagonal rebuilt on
 * the device.  Handles made from an explicit iA list (akm_create_sparse)
 * return AKM_ERR_UNSUPPORTED. */
akm_status akm_update_operators(akm_matrix mat, const void *const *ops, int64_t lda);

/* kronmult_matrix<P>::apply<resource>(alpha, x, beta, y)
 * (src/asgard_kronmult_matrix.hpp:340-478): y = alpha*A*x + beta*y.
 * With no terms: y = beta*y.  beta == 0 never reads y. */
akm_status akm_apply(akm_matrix mat, double alpha, const void *x, double beta, void *y,
                     int on_device);

/* The implicit operator of simple_gmres_euler (src/solver.cpp:126-131) in
 * one pass: y = alpha*(x - dt*A*x) + beta*y (device pointers). */
akm_status akm_apply_euler(akm_matrix mat, double dt, double alpha, const void *x,
                           double beta, void *y);

/* Sizes: tensor_size n^d, input_size, output_size (hpp:481-490), number of
 * connected pairs, and flops per apply in the ×2 convention of
 * kronmult_matrix::compute_flops (hpp:507-512). */
akm_status akm_matrix_info(akm_matrix mat, int64_t *tensor_size, int64_t *input_size,
                           int64_t *output_size, int64_t *num_pairs, int64_t *flops);

/* Tuning knobs: wavefronts per row workgroup (0 = automatic, else 1, 2, 4,
 * 8 or 16; 16 only for the matrix-core kernels, the others run 8), the LDS
 * bytes a workgroup may hold for the row's factor tables
 * (-1 = automatic, 0 = read factors from global memory) and whether to
 * force the block-wide LDS kernel (for tests).  Takes effect at the next
 * apply.  On a global operator only waves_per_row counts and picks the row
 * layout: 1 = one lane per DOF, 8 = eight lanes per DOF, 2 = one wavefront
 * per cell (n^d <= 64), 3 = as 2 with the chains whose first stage reads
 * the same input along the same dimension grouped (one read for up to four
 * chains), 4 / 5 = as 2 / 3 with the workgroups of each XCD taking
 * contiguous cells (neighbours shared in one L2), 6 = slabs: the terms
 * with one or two non-identity dimensions grouped by those dimensions, one
 * workgroup per bundle of slabs (cells equal outside the group's
 * dimensions) running all the group's chains in LDS, the other terms as 2
 * (or 1 when n^d > 64), 0 = unchanged; akm_autotune times them and reports
 * its choice the same way.  No counterpart in the reference (its launch
 * shape is fixed at build time, CMakeLists.txt:231-232). */
akm_status akm_set_tuning(akm_matrix mat, int waves_per_row, int lds_cap_bytes,
                          int force_generic);

/* Autotuner: times every workgroup shape (1, 2, 4, 8 wavefronts per row)
 * with and without the LDS factor tables on the matrix itself (device x/y,
 * y is overwritten), and the sum-factorised matrix-core kernel where it
 * applies (akm_set_kernel_family), and keeps the fastest.  Choices are
 * cached per (precision, d, n, terms, max level, rows, pairs) for the
 * process, so a second matrix of the same shape reuses them without timing.
 * Writes the chosen waves per row and kernel (0 = factors read from global
 * memory, 1 = LDS factor tables, 2 = matrix cores) when the pointers are not
 * NULL.  No counterpart in the reference (fixed launch shapes). */
akm_status akm_autotune(akm_matrix mat, const void *x, void *y, int *waves_per_row,
                        int *lds_tables);

/* Kernel family of the apply: 0 = the register / LDS-table kernels, 1 = the
 * sum-factorised matrix-core kernels for FP64 sparse-grid matrices with
 * (d, n) = (3, 3), (4, 2) or (6, 2), 1 to 8 terms and max_level <= 10: the
 * row's columns are taken in lexicographic order of their 1D indices, the
 * trailing dimensions are contracted per column on the FP64 matrix cores and
 * summed over the columns that share the leading 1D indices before the
 * leading dimensions are contracted once per such group.
 * AKM_ERR_UNSUPPORTED for any other matrix.  Takes effect at the next apply;
 * akm_set_tuning's waves_per_row sets its wavefronts per row.  No
 * counterpart in the reference. */
akm_status akm_set_kernel_family(akm_matrix mat, int family);

/* Times `iterations` back-to-back applies (device x/y) with HIP events on the
 * matrix's stream; returns the average milliseconds per apply. */
akm_status akm_time_apply(akm_matrix mat, double alpha, const void *x, double beta,
                          void *y, int iterations, double *ms_per_apply);

/* Releases the matrix and all of its device memory. */
akm_status akm_destroy(akm_matrix mat);

/* ---------------------------------------------------------------- GMRES -- */

/* solver::simple_gmres_euler<P, resource> (src/solver.cpp:119-135) on top of
 * solver::simple_gmres (:156-298): solves (I - dt*A) x = b with restarted
 * GMRES(restart), classical Gram-Schmidt, Givens rotations and an absolute
 * tolerance on the residual estimate, x holding the initial guess on entry.
 * restart/max_iter/tol <= 0 select the reference defaults
 * (default_gmres_restarts, n outer iterations, 1e-12 FP64 / 1e-6 FP32).
 * All vector work stays on the GPU.  err and iterations mirror gmres_info.
 * Returns AKM_ERR_INVALID if restart > n, as the reference throws, and
 * AKM_ERR_UNSUPPORTED if restart > 2000 (the reference's default is at most
 * 200).  On a
 * global operator (akm_create_global_grid) the residual and every new basis
 * vector are divided by (1 - dt*pc), pc its Jacobi diagonal, as the
 * reference's global-mode solve does (src/solver.cpp:76-118); tol and err
 * then refer to that preconditioned residual. */
akm_status akm_gmres_euler(akm_matrix mat, double dt, void *x, const void *b, int restart,
                           int max_iter, double tol, int on_device, double *err,
                           int *iterations);

/* Device scratch of akm_gmres_euler_noalloc on each GPU, in bytes. */
#define AKM_NOALLOC_SCRATCH_BYTES (32ll << 20)

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): akm_gmres_euler on device vectors x and b, with the
 * same algorithm, results, defaults and errors, that allocates nothing.  The
 * Krylov basis, the small work array and the step state live in a fixed
 * per-GPU device array of AKM_NOALLOC_SCRATCH_BYTES, which each GPU holds
 * until the process ends: about 8*(n*(restart+1) + restart^2/2) bytes for
 * FP64 must fit, else AKM_ERR_INVALID names the need and the limit
 * (restart <= 0 takes the reference default, lowered until it fits).  Every
 * call on one GPU uses that GPU's array, so calls that overlap on one GPU
 * (several threads or streams) race; each call writes every value it reads.
 * The host waits for each Arnoldi step (no step is enqueued ahead). */
akm_status akm_gmres_euler_noalloc(akm_matrix mat, double dt, void *x, const void *b,
                                   int restart, int max_iter, double tol, double *err,
                                   int *iterations);

/* solver::simple_gmres(fk::matrix A, fk::vector x, fk::vector b, fk::matrix M,
 * restart, max_iter, tol) (src/solver.hpp:17-22, src/solver.cpp:54-74): the
 * reference's node-local GMRES on an explicit n x n matrix, with the same
 * algorithm, defaults, absolute tolerance and error behaviour as
 * akm_gmres_euler.  A and M are column-major host arrays (fk::matrix layout),
 * x (initial guess in, solution out) and b host vectors, all of precision
 * prec.  M (NULL = none) is a left preconditioner applied as M^-1 through its
 * LU factors with partial pivoting, as dense_preconditioner does with
 * getrf/getrs (src/solver.cpp:11-38), to the residual and to every new
 * Krylov vector; n <= 4096 with M.  A singular M returns AKM_ERR_INVALID.
 * The products and solves run on the GPU (null stream); the call returns
 * with x on the host. */
akm_status akm_gmres_dense(int n, const void *A, void *x, const void *b, const void *M,
                           int restart, int max_iter, double tol, akm_prec prec, double *err,
                           int *iterations);

/* --------------------------------------------------- DG operators (1D) -- */

#define AKM_DG_MASS 0    /* int g phi_j psi_i */
#define AKM_DG_DIV 1     /* -int g phi_j psi_i' + flux {{g u}} + s |g|/2 [[u]] */
#define AKM_DG_GRAD 2    /* -(div, flux reversed, Dirichlet/Neumann swapped)^T */
#define AKM_DG_PENALTY 3 /* the jump part s |g|/2 [[u]] of the div flux only */

#define AKM_DG_PERIODIC 0
#define AKM_DG_DIRICHLET 1 /* no boundary flux (penalty: the jump at the wall) */
#define AKM_DG_NEUMANN 2   /* outflow flux at the wall */

/* One partial term of a 1D DG operator.  gq: device array of the coefficient
 * g at the nq Gauss-Legendre points (ascending) of each of the 2^level cells,
 * cell-major (not read for penalty terms); ge: device array of g at the
 * 2^level + 1 cell edges a + e*h (all but mass terms).  flux_scale s: -1
 * downwind, 0 central, 1 upwind.  left/right: the boundary conditions at a
 * and b (partial_term::set_bilinear_boundary, src/pde/pde_base.hpp:240-270). */
typedef struct
{
  int kind;
  double flux_scale;
  const double *gq;
  const double *ge;
  int left;
  int right;
} akm_dg_partial;

/* generate_coefficients (src/coefficients.cpp:18-245) on the GPU for one
 * 1D term on a domain [a, b] of 2^level cells with n basis
 * functions: the product of its partial terms (rechain_dimension), each
 * multiplied by the inverse unit mass matrix, rotated to the multiwavelet
 * basis, out = F (R_1 ... R_p) F^T.  F is the caller's N x N column-major
 * transform (operator_two_scale), out an N x N column-major device matrix,
 * N = n * 2^level <= 8192, ready for akm_update_operators.  Enqueued on
 * `stream` (hipStream_t, NULL = null stream) and synchronized. */
akm_status akm_dg_term(int n, int level, double a, double b, int nq,
                       const akm_dg_partial *partials, int num_partials, const double *F,
                       double *out, void *stream);

/* ----------------------------------------------------------------- IMEX -- */

/* Events of an IMEX step at which the caller's hook runs. */
#define AKM_IMEX_EXPLICIT_1 0 /* before E f0; f = f0 */
#define AKM_IMEX_IMPLICIT_1 1 /* before the stage-1 solve; f = its right-hand side */
#define AKM_IMEX_EXPLICIT_2 2 /* before E f1; f = the stage-1 result */
#define AKM_IMEX_IMPLICIT_2 3 /* before the stage-2 solve; f = its right-hand side */

/* Caller code run inside an IMEX step with the device vector of the event
 * (the stream is synchronized first); returns 0 to continue. */
typedef int (*akm_imex_hook)(void *user, int event, const void *f);

/* imex_advance (src/time_advance.cpp:760-1000) with imex_resrc =
 * resource::device: one SSP-RK2 IMEX step on the FP64 device vector f,
 *   f1s = f0 + dt E f0;  solve (I - dt B) f1 = f1s (initial guess f1s);
 *   f2s = (f0 + f1 + dt E f1) / 2;  solve (I - dt/2 B) f = f2s (guess f1),
 * E = explicit_op, B = implicit_op (square, same size).  Without
 * `collision` the two solves are skipped (f1 = f1s, f = f2s), as in the
 * reference's non-collision branch.  `hook` (may be NULL) runs at the four
 * events above, where the reference recomputes moments and regenerates and
 * resets coefficients (calculate_moments, generate_all_coefficients,
 * reset_coefficients); it may update either handle (akm_update_operators).
 * restart, max_iter and tol are those of akm_gmres_euler; errors[2] and
 * iterations[2] (may be NULL) receive each solve's gmres_info.  Returns
 * AKM_ERR_NOT_CONVERGED when a solve stopped above the tolerance. */
akm_status akm_imex_step(akm_matrix explicit_op, akm_matrix implicit_op, double dt, void *f,
                         int restart, int max_iter, double tol, int collision,
                         akm_imex_hook hook, void *user, double *errors, int *iterations);

/* A right preconditioner of the flexible solver (see akm_mlp_create). */
typedef struct akm_precond_s *akm_precond;

/* Outcome of akm_gmres_euler_mixed. */
typedef struct
{
  double rel_residual;     /* |b - (I - dt*A)x| / |b| with the FP64 operator, at exit */
  double abs_residual;     /* |b - (I - dt*A)x| */
  double rhs_norm;         /* |b| */
  int iterations;          /* Krylov iterations over all cycles */
  int cycles;              /* refinement (restart) cycles */
  int applies_f64;         /* FP64 operator applications (residuals included) */
  int applies_f32;         /* FP32 operator applications; with a cheap operator
                              (under 0.25 ms per apply) each cycle also runs one
                              speculative step whose result is discarded */
  int precond_applies;     /* preconditioner applications */
  int fallback;            /* 1 if plain FP64 GMRES took over */
  int fallback_iterations; /* Krylov iterations of the plain FP64 fallback */
  int converged;           /* 1 if rel_residual <= tol */
} akm_solve_stats;

/* The implicit step (I - dt*A) x = b of simple_gmres_euler
 * (src/solver.cpp:119-135) by FP64 iterative refinement around flexible
 * GMRES(restart) cycles.  mat is the FP64 operator; it computes every
 * residual, and the Krylov basis, Gram-Schmidt and Hessenberg least squares
 * are FP64.  Inside a cycle the operator is mat32 (an FP32 matrix of the same
 * operator; NULL = mat itself) and, when pre is not NULL, every basis vector
 * goes through the right preconditioner first (flexible GMRES).  A cycle with
 * mat32 stops once its residual estimate is inner_tol times its starting
 * residual.  After each cycle the FP64 residual must at least halve; if it
 * does not, the cycle is undone and plain FP64 GMRES continues to the end.  The
 * tolerance is relative: |b - (I - dt*A)x| <= tol*|b| with the FP64
 * operator, which is also how the answer is judged.  x (initial guess in,
 * solution out) and b are FP64 device vectors.  tol <= 0 selects 1e-10,
 * inner_tol <= 0 1e-6, restart <= 0 default_gmres_restarts, max_cycles <= 0
 * n.  Returns AKM_ERR_NOT_CONVERGED (1) when max_cycles ran out.
 * Extends the reference, whose GMRES runs in one precision and without a
 * preconditioner (the euler operator's preconditioner is a no-op,
 * src/solver.cpp:133). */
akm_status akm_gmres_euler_mixed(akm_matrix mat, akm_matrix mat32, akm_precond pre, double dt,
                                 double *x, const double *b, int restart, int max_cycles,
                                 double tol, double inner_tol, akm_solve_stats *stats);

/* --------------------------------------------------- batched solves -- */
/* No counterpart in the reference, which solves one implicit step at a time
 * (time_advance::implicit_advance, src/time_advance.cpp:432-452, one
 * simple_gmres_euler each).  A batch of independent implicit steps
 * (I - dt_s A_s) x_s = b_s over FP64 sparse-grid matrices of any sizes and
 * (d, n) (n in 2..4; d <= 6 with n = 2, d <= 5 with n = 3, d <= 3 with
 * n = 4): restarted GMRES(restart) in lockstep, every operator apply one
 * persistent launch per (d, n) group over a (system, row) work list, every
 * restart cycle three hipGraphs (start, Arnoldi iteration, update).  The
 * matrices must outlive the batch and keep their coefficients. */
typedef struct akm_batch_s *akm_batch;

/* Builds the batch: work lists, device workspaces, per-group workgroup shape
 * chosen by timing (1, 2, 4 or 8 wavefronts per row), captured graphs.
 * restart <= 0 selects 50 (at most 255 and the smallest system size). */
akm_status akm_batch_create(int count, const akm_matrix *mats, const double *dts, int restart,
                            akm_batch *batch);

/* restart used, number of (d, n) groups and each group's wavefronts per row
 * (group_waves has num_groups entries; groups ordered by (d, n)). */
akm_status akm_batch_info(akm_batch batch, int *restart, int *num_groups, int *group_waves);

/* Solves every system to |b_s - (I - dt_s A_s) x_s| <= tol*|b_s| (true FP64
 * residual, tol <= 0 selects 1e-10), at most max_cycles restart cycles
 * (<= 0: 1000).  x[s] (initial guess in, solution out) and rhs[s] are FP64
 * device vectors; the call synchronizes the device first and returns with the
 * solutions in place.  Per-system outputs (host arrays of count, may be NULL):
 * relative residual, operator applications, converged flag.  Returns
 * AKM_ERR_NOT_CONVERGED if any system ran out of cycles. */
akm_status akm_batch_solve(akm_batch batch, double *const *x, const double *const *rhs,
                           int max_cycles, double tol, double *rel_residual, int *iterations,
                           int *converged);

akm_status akm_batch_destroy(akm_batch batch);

/* ------------------------------------------- learned preconditioner -- */
/* No counterpart in the reference.  A per-cell MLP right preconditioner:
 * for cell c with residual block r_c (n^d values) and s = |r_c| it reads
 * [r_c/s, f_c] (f_c: per dimension (level + (position+1/2)/cells of the
 * level)/(max_level+1)), has two ReLU hidden layers of `hidden` units and a
 * linear output o, and returns z_c = r_c + s*o.  Parameters are one FP32
 * array [W1 (H x K1) | b1 (H) | W2 (H x H) | b2 (H) | W3 (N3 x H) | b3 (N3)],
 * W row-major (output x input), K1 = roundup16(n^d + d), N3 = roundup16(n^d);
 * padding rows/columns are zero.  Training runs in FP32; the preconditioner
 * runs with BF16 weights on the matrix cores, FP32 accumulation. */

/* Creates the network for the grid `table` (num_cells x 2d host ints, as
 * akm_create_sparse_grid takes it): He-normal hidden layers from `seed`,
 * zero biases, zero output layer (identity preconditioner).  hidden <= 0
 * selects 256 (multiple of 16, <= 512); n^d <= 256. */
akm_status akm_mlp_create(int d, int n, int64_t num_cells, const int *table, int max_level,
                          int hidden, uint64_t seed, akm_precond *pre);

/* Parameter count and layer widths (inputs n^d + d, padded K1, hidden,
 * outputs n^d, padded N3). */
akm_status akm_mlp_info(akm_precond pre, int64_t *num_params, int *inputs, int *padded_inputs,
                        int *hidden, int *outputs, int *padded_outputs);

/* Stream of the training and parameter calls (hipStream_t, NULL = null stream). */
akm_status akm_mlp_set_stream(akm_precond pre, void *stream);

/* Copies the FP32 parameters to / from a host array of num_params floats. */
akm_status akm_mlp_get_params(akm_precond pre, float *params);
akm_status akm_mlp_set_params(akm_precond pre, const float *params);

/* FP32 forward: O (rows x N3, device) = network outputs o for the inputs X
 * (rows x K1, device). */
akm_status akm_mlp_forward(akm_precond pre, const float *X, int64_t rows, float *O);

/* Mean squared error over rows x n^d outputs of o(X) against targets T
 * (rows x N3, device), and its gradient with respect to every parameter
 * (num_params floats, device; may be NULL). */
akm_status akm_mlp_gradients(akm_precond pre, const float *X, const float *T, int64_t rows,
                             double *loss, float *grad);

/* One Adam step (beta1 0.9, beta2 0.999, eps 1e-8) with the gradient
 * `grad` (device, num_params floats); the step counter is kept by pre. */
akm_status akm_mlp_adam_step(akm_precond pre, const float *grad, double lr);

/* Trains on num_vectors (residual, correction) pairs of whole-grid FP64
 * device vectors (num_vectors x num_cells*n^d each, pair q at offset
 * q*num_cells*n^d): every cell of every pair is one sample with input
 * [r_c/s, f_c] and target (e_c - r_c)/s.  Shuffled mini-batches of `batch`
 * samples, Adam with learning rate lr, `epochs` passes; writes the mean loss
 * of the last epoch.  Refreshes the BF16 weights. */
akm_status akm_mlp_train(akm_precond pre, const double *residuals, const double *corrections,
                         int64_t num_vectors, int epochs, int batch, double lr, uint64_t seed,
                         double *final_loss);

/* z = P(v) for FP64 device vectors of num_cells*n^d values, enqueued on
 * `stream` (hipStream_t, NULL = null stream). */
akm_status akm_precond_apply(akm_precond pre, const double *v, double *z, void *stream);

/* Releases the preconditioner. */
akm_status akm_precond_destroy(akm_precond pre);

#ifdef __cplusplus
}
#endif

#endif
