/*
 * Oracle: restarted GMRES of the reference's implicit step.
 *
 * TEST INFRASTRUCTURE ONLY (the checker; never linked by the product).
 *
 * Restates solver::simple_gmres, src/solver.cpp:156-298:
 *   - outer: r = b - A x (one matvec, counted), left preconditioner, norm;
 *   - inner: w = A v_j (counted), preconditioner, single-pass classical
 *     Gram-Schmidt through two gemv calls, norm, normalise, apply the
 *     previous Givens rotations to the new Hessenberg column, rotg,
 *     absolute residual estimate |s_j * g_j|; the right-hand side g is
 *     rotated only when the estimate is still above the tolerance
 *     (solver.cpp:261-267), exactly as the reference does;
 *   - update: packed upper-triangular back-solve (tpsv 'U','N','N') and
 *     x += V y.
 * BLAS-1/2 semantics follow the reference BLAS (rot, rotg) since the
 * reference links any vendor BLAS.
 */
#include "asgard_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

static double dnrm2(int64_t n, const double *x)
{
  double s = 0;
  for (int64_t i = 0; i < n; i++)
    s += x[i] * x[i];
  return sqrt(s);
}

/* reference BLAS drot: x' = c x + s y, y' = c y - s x */
static void drot(double *x, double *y, double c, double s)
{
  double t = c * (*x) + s * (*y);
  *y = c * (*y) - s * (*x);
  *x = t;
}

/* reference BLAS drotg */
static void drotg(double *a, double *b, double *c, double *s)
{
  double roe = (fabs(*a) > fabs(*b)) ? *a : *b;
  double scale = fabs(*a) + fabs(*b);
  double r, z;
  if (scale == 0)
  {
    *c = 1;
    *s = 0;
    r = 0;
    z = 0;
  }
  else
  {
    double as = *a / scale, bs = *b / scale;
    r = scale * sqrt(as * as + bs * bs);
    r = (roe >= 0 ? 1.0 : -1.0) * r;
    *c = *a / r;
    *s = *b / r;
    z = 1;
    if (fabs(*a) > fabs(*b))
      z = *s;
    if (fabs(*b) >= fabs(*a) && *c != 0)
      z = 1.0 / *c;
  }
  *a = r;
  *b = z;
}

static int pos_from_indices(int i, int j) { return i + j * (j + 1) / 2; }

int orc_simple_gmres(int n, orc_matvec_fn mat, void *mat_ctx,
                     orc_precond_fn precond, void *pc_ctx, double *x,
                     const double *b, int restart, int max_outer, double tol,
                     double *err, int *iterations)
{
  if (restart > n)
    return -1;
  if (restart <= 0 || max_outer <= 0)
    return -2;
  double *basis = (double *)malloc(sizeof(double) * (size_t)n * (restart + 1));
  double *proj = (double *)calloc((size_t)restart * (restart + 1) / 2, sizeof(double));
  double *sines = (double *)calloc((size_t)restart + 1, sizeof(double));
  double *cosines = (double *)calloc((size_t)restart + 1, sizeof(double));
  double *gsol = (double *)calloc((size_t)restart + 1, sizeof(double));

  int total_iterations = 0, outer_iterations = 0, inner_iterations = 0;
  double inner_res = 0, outer_res = tol + 1;
  while (outer_res > tol && outer_iterations < max_outer)
  {
    double *scaled = basis;
    memcpy(scaled, b, sizeof(double) * (size_t)n);
    mat(mat_ctx, -1.0, x, 1.0, scaled);
    if (precond)
      precond(pc_ctx, scaled);
    ++total_iterations;

    inner_res = dnrm2(n, scaled);
    {
      double inv = 1.0 / inner_res;
      for (int i = 0; i < n; i++)
        scaled[i] *= inv;
    }
    gsol[0] = inner_res;

    inner_iterations = 0;
    while (inner_res > tol && inner_iterations < restart)
    {
      const double *tmp = basis + (int64_t)inner_iterations * n;
      double *w = basis + (int64_t)(inner_iterations + 1) * n;
      mat(mat_ctx, 1.0, tmp, 0.0, w);
      if (precond)
        precond(pc_ctx, w);
      ++total_iterations;

      double *coeffs = proj + pos_from_indices(0, inner_iterations);
      /* coeffs = V^T w (gemv 't'), then w = w - V coeffs (gemv 'n') */
      for (int k = 0; k <= inner_iterations; k++)
      {
        const double *v = basis + (int64_t)k * n;
        double s = 0;
        for (int i = 0; i < n; i++)
          s += v[i] * w[i];
        coeffs[k] = s;
      }
      for (int i = 0; i < n; i++)
      {
        double s = 0;
        for (int k = 0; k <= inner_iterations; k++)
          s += basis[(int64_t)k * n + i] * coeffs[k];
        w[i] = w[i] - s;
      }
      double nrm = dnrm2(n, w);
      {
        double inv = 1.0 / nrm;
        for (int i = 0; i < n; i++)
          w[i] *= inv;
      }
      for (int k = 0; k < inner_iterations; k++)
        drot(&coeffs[k], &coeffs[k + 1], cosines[k], sines[k]);

      double beta = nrm;
      drotg(&coeffs[inner_iterations], &beta, &cosines[inner_iterations],
            &sines[inner_iterations]);

      inner_res = fabs(sines[inner_iterations] * gsol[inner_iterations]);

      if (inner_res > tol && inner_iterations < restart)
      {
        gsol[inner_iterations + 1] = 0.;
        drot(&gsol[inner_iterations], &gsol[inner_iterations + 1],
             cosines[inner_iterations], sines[inner_iterations]);
      }
      ++inner_iterations;
    }

    if (inner_iterations > 0)
    {
      /* tpsv: packed upper triangular, no transpose, non-unit diagonal */
      int m = inner_iterations;
      for (int j = m - 1; j >= 0; j--)
      {
        gsol[j] /= proj[pos_from_indices(j, j)];
        for (int i = 0; i < j; i++)
          gsol[i] -= gsol[j] * proj[pos_from_indices(i, j)];
      }
      for (int i = 0; i < n; i++)
      {
        double s = 0;
        for (int k = 0; k < m; k++)
          s += basis[(int64_t)k * n + i] * gsol[k];
        x[i] += s;
      }
    }
    ++outer_iterations;
    outer_res = inner_res;
  }
  if (err)
    *err = outer_res;
  if (iterations)
    *iterations = total_iterations;
  free(basis);
  free(proj);
  free(sines);
  free(cosines);
  free(gsol);
  return 0;
}

/* ---------------------------------------------------- dense variants --- */

struct dense_ctx
{
  int n;
  const double *A;
};

static void dense_matvec(void *vctx, double alpha, const double *x,
                         double beta, double *y)
{
  struct dense_ctx *ctx = (struct dense_ctx *)vctx;
  int n = ctx->n;
  for (int i = 0; i < n; i++)
  {
    double s = 0;
    for (int j = 0; j < n; j++)
      s += ctx->A[(int64_t)j * n + i] * x[j];
    y[i] = alpha * s + ((beta == 0) ? 0.0 : beta * y[i]);
  }
}

struct lu_ctx
{
  int n;
  double *LU;
  int *piv;
};

/* getrf with partial pivoting, column-major */
static int lu_factor(int n, double *A, int *piv)
{
  for (int k = 0; k < n; k++)
  {
    int p = k;
    double mx = fabs(A[(int64_t)k * n + k]);
    for (int i = k + 1; i < n; i++)
      if (fabs(A[(int64_t)k * n + i]) > mx)
      {
        mx = fabs(A[(int64_t)k * n + i]);
        p = i;
      }
    piv[k] = p;
    if (mx == 0)
      return k + 1;
    if (p != k)
      for (int j = 0; j < n; j++)
      {
        double t = A[(int64_t)j * n + k];
        A[(int64_t)j * n + k] = A[(int64_t)j * n + p];
        A[(int64_t)j * n + p] = t;
      }
    double inv = 1.0 / A[(int64_t)k * n + k];
    for (int i = k + 1; i < n; i++)
      A[(int64_t)k * n + i] *= inv;
    for (int j = k + 1; j < n; j++)
    {
      double akj = A[(int64_t)j * n + k];
      if (akj != 0)
        for (int i = k + 1; i < n; i++)
          A[(int64_t)j * n + i] -= A[(int64_t)k * n + i] * akj;
    }
  }
  return 0;
}

static void lu_solve(int n, const double *LU, const int *piv, double *b)
{
  for (int k = 0; k < n; k++)
    if (piv[k] != k)
    {
      double t = b[k];
      b[k] = b[piv[k]];
      b[piv[k]] = t;
    }
  for (int j = 0; j < n; j++)
    for (int i = j + 1; i < n; i++)
      b[i] -= LU[(int64_t)j * n + i] * b[j];
  for (int j = n - 1; j >= 0; j--)
  {
    b[j] /= LU[(int64_t)j * n + j];
    for (int i = 0; i < j; i++)
      b[i] -= LU[(int64_t)j * n + i] * b[j];
  }
}

static void lu_precond(void *vctx, double *x)
{
  struct lu_ctx *ctx = (struct lu_ctx *)vctx;
  lu_solve(ctx->n, ctx->LU, ctx->piv, x);
}

int orc_simple_gmres_dense(int n, const double *A, const double *M, double *x,
                           const double *b, int restart, int max_outer,
                           double tol, double *err, int *iterations)
{
  struct dense_ctx dctx = {n, A};
  if (M)
  {
    struct lu_ctx lctx;
    lctx.n = n;
    lctx.LU = (double *)malloc(sizeof(double) * (size_t)n * n);
    lctx.piv = (int *)malloc(sizeof(int) * (size_t)n);
    memcpy(lctx.LU, M, sizeof(double) * (size_t)n * n);
    int info = lu_factor(n, lctx.LU, lctx.piv);
    int rc = -3;
    if (info == 0)
      rc = orc_simple_gmres(n, dense_matvec, &dctx, lu_precond, &lctx, x, b,
                            restart, max_outer, tol, err, iterations);
    free(lctx.LU);
    free(lctx.piv);
    return rc;
  }
  return orc_simple_gmres(n, dense_matvec, &dctx, NULL, NULL, x, b, restart,
                          max_outer, tol, err, iterations);
}

int orc_dense_solve(int n, double *A, double *b)
{
  int *piv = (int *)malloc(sizeof(int) * (size_t)n);
  int info = lu_factor(n, A, piv);
  if (info == 0)
    lu_solve(n, A, piv, b);
  free(piv);
  return info;
}

/* ------------------------------------------------- implicit Euler step --- */

struct euler_ctx
{
  int d, n, num_cells, T;
  const int64_t *pntr;
  const int *indx;
  const int *iA;
  const double *vA;
  double dt;
  int64_t size;
};

/* The operator of simple_gmres_euler (src/solver.cpp:119-135):
 * y <- alpha*x - dt*alpha*A x + beta*y, as apply(-dt*alpha, x, beta, y)
 * followed by axpy(alpha, x, y). */
static void euler_matvec(void *vctx, double alpha, const double *x,
                         double beta, double *y)
{
  struct euler_ctx *c = (struct euler_ctx *)vctx;
  orc_kronmult_sparse(c->d, c->n, c->num_cells, c->pntr, c->indx, c->T, c->iA,
                      c->vA, -c->dt * alpha, x, beta, y);
  for (int64_t i = 0; i < c->size; i++)
    y[i] += alpha * x[i];
}

int orc_gmres_euler_sparse(int d, int n, int num_cells, const int64_t *pntr,
                           const int *indx, int T, const int *iA,
                           const double *vA, double dt, double *x,
                           const double *b, int restart, int max_outer,
                           double tol, double *err, int *iterations)
{
  int64_t N = 1;
  for (int k = 0; k < d; k++)
    N *= n;
  struct euler_ctx ctx = {d, n, num_cells, T, pntr, indx, iA, vA, dt,
                          N * num_cells};
  return orc_simple_gmres((int)(N * num_cells), euler_matvec, &ctx, NULL, NULL,
                          x, b, restart, max_outer, tol, err, iterations);
}
