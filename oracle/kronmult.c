/*
 * Oracle: CPU kronmult, restating kronmult::cpu_sparse and kronmult::cpu_dense
 * of the reference, plus the explicit-Kronecker brute force of its tests.
 *
 * TEST INFRASTRUCTURE ONLY (the checker and the CPU baseline; never linked
 * by the product).
 *
 * Loop structure follows src/device/asgard_spkronmult_cpu.cpp:84-433:
 *   OpenMP over output rows; beta applied to the row first; for each
 *   connected column and each term, d contraction stages on a stack tensor,
 *   stage k applying the k-th factor (iA[ma++], column-major n x n, entry
 *   A[j*n + s] = A(s, j)) to tensor index k (index 0 is the slowest);
 *   then y += alpha * Y with the alpha = 1 / -1 special cases.
 * The (d, n) in {1..6} x {2, 3, 4} cases are instantiated with constant
 * sizes so the compiler unrolls them, as the reference's templates do.
 */
#include "asgard_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_INLINE static inline __attribute__((always_inline))

enum scalar_case
{
  SC_ZERO,
  SC_ONE,
  SC_NEG_ONE,
  SC_OTHER
};

static enum scalar_case classify(double v)
{
  if (v == 0)
    return SC_ZERO;
  if (v == 1)
    return SC_ONE;
  if (v == -1)
    return SC_NEG_ONE;
  return SC_OTHER;
}

static int64_t ipow_i(int n, int d)
{
  int64_t r = 1;
  for (int k = 0; k < d; k++)
    r *= n;
  return r;
}

/* One contraction stage: tensor viewed as (outer, n, inner),
 * out[o, s, i] = sum_j A[j*n + s] * in[o, j, i]. */
#define DEFINE_STAGE(T, NAME)                                                  \
  ORC_INLINE void NAME(int n, int64_t outer, int64_t inner, const T *A,        \
                       const T *in, T *out)                                    \
  {                                                                            \
    for (int64_t o = 0; o < outer; o++)                                        \
      for (int s = 0; s < n; s++)                                              \
        for (int64_t i = 0; i < inner; i++)                                    \
        {                                                                      \
          T acc = 0;                                                           \
          for (int j = 0; j < n; j++)                                          \
            acc += A[j * n + s] * in[(o * n + j) * inner + i];                 \
          out[(o * n + s) * inner + i] = acc;                                  \
        }                                                                      \
  }

DEFINE_STAGE(double, stage_f64)
DEFINE_STAGE(float, stage_f32)

/* y-row update with the alpha special cases */
#define DEFINE_ACCUM(T, NAME)                                                  \
  ORC_INLINE void NAME(int64_t N, enum scalar_case ac, T alpha, const T *Y,    \
                       T *y)                                                   \
  {                                                                            \
    if (ac == SC_ONE)                                                          \
      for (int64_t e = 0; e < N; e++)                                          \
        y[e] += Y[e];                                                          \
    else if (ac == SC_NEG_ONE)                                                 \
      for (int64_t e = 0; e < N; e++)                                          \
        y[e] -= Y[e];                                                          \
    else                                                                       \
      for (int64_t e = 0; e < N; e++)                                          \
        y[e] += alpha * Y[e];                                                  \
  }

DEFINE_ACCUM(double, accum_f64)
DEFINE_ACCUM(float, accum_f32)

#define DEFINE_BETA(T, NAME)                                                   \
  ORC_INLINE void NAME(int64_t N, enum scalar_case bc, T beta, T *y)           \
  {                                                                            \
    if (bc == SC_ZERO)                                                         \
      for (int64_t e = 0; e < N; e++)                                          \
        y[e] = 0;                                                              \
    else if (bc == SC_NEG_ONE)                                                 \
      for (int64_t e = 0; e < N; e++)                                          \
        y[e] = -y[e];                                                          \
    else if (bc == SC_OTHER)                                                   \
      for (int64_t e = 0; e < N; e++)                                          \
        y[e] *= beta;                                                          \
  }

DEFINE_BETA(double, beta_f64)
DEFINE_BETA(float, beta_f32)

/* Contraction of one (column, term): d stages, ping-ponging W/Y.
 * Returns the buffer holding the result. */
#define DEFINE_CONTRACT(T, NAME, STAGE)                                        \
  ORC_INLINE T *NAME(int d, int n, const T *vA, const int *ia, const T *x,     \
                     T *W, T *Y)                                               \
  {                                                                            \
    int64_t outer = 1, inner = ipow_i(n, d - 1);                               \
    const T *in = x;                                                           \
    T *out = (d % 2 == 1) ? Y : W;                                             \
    for (int k = 0; k < d; k++)                                                \
    {                                                                          \
      STAGE(n, outer, inner, vA + ia[k], in, out);                             \
      in = out;                                                                \
      out = (out == W) ? Y : W;                                                \
      outer *= n;                                                              \
      inner /= n;                                                              \
    }                                                                          \
    return (T *)in;                                                            \
  }

DEFINE_CONTRACT(double, contract_f64, stage_f64)
DEFINE_CONTRACT(float, contract_f32, stage_f32)

/* Sparse kronmult body for n > 1; d, n are compile-time constants at the
 * instantiation sites below. */
#define DEFINE_SPARSE_BODY(T, NAME, CONTRACT, BETA, ACCUM)                     \
  ORC_INLINE void NAME(int d, int n, int num_rows, const int64_t *pntr,        \
                       const int *indx, int nterms, const int *iA,             \
                       const T *vA, T alpha, const T *x, T beta, T *y)         \
  {                                                                            \
    const int64_t N = ipow_i(n, d);                                            \
    const enum scalar_case ac = classify((double)alpha);                       \
    const enum scalar_case bc = classify((double)beta);                        \
    _Pragma("omp parallel")                                                    \
    {                                                                          \
      T *W = (T *)malloc(sizeof(T) * (size_t)N);                               \
      T *Y = (T *)malloc(sizeof(T) * (size_t)N);                               \
      _Pragma("omp for")                                                       \
      for (int iy = 0; iy < num_rows; iy++)                                    \
      {                                                                        \
        T *yrow = y + (int64_t)iy * N;                                         \
        BETA(N, bc, beta, yrow);                                               \
        int64_t ma = pntr[iy] * (int64_t)nterms * d;                           \
        for (int64_t jx = pntr[iy]; jx < pntr[iy + 1]; jx++)                   \
        {                                                                      \
          const T *xc = x + (int64_t)indx[jx] * N;                             \
          for (int t = 0; t < nterms; t++)                                     \
          {                                                                    \
            const T *res = CONTRACT(d, n, vA, iA + ma, xc, W, Y);              \
            ma += d;                                                           \
            ACCUM(N, ac, alpha, res, yrow);                                    \
          }                                                                    \
        }                                                                      \
      }                                                                        \
      free(W);                                                                 \
      free(Y);                                                                 \
    }                                                                          \
  }

DEFINE_SPARSE_BODY(double, sparse_body_f64, contract_f64, beta_f64, accum_f64)
DEFINE_SPARSE_BODY(float, sparse_body_f32, contract_f32, beta_f32, accum_f32)

/* n == 1: cpu_n0, src/device/asgard_spkronmult_cpu.cpp:26-60 */
#define DEFINE_SPARSE_N1(T, NAME)                                              \
  static void NAME(int d, int num_rows, const int64_t *pntr, const int *indx,  \
                   int nterms, const int *iA, const T *vA, T alpha,            \
                   const T *x, T beta, T *y)                                   \
  {                                                                            \
    const enum scalar_case ac = classify((double)alpha);                       \
    const enum scalar_case bc = classify((double)beta);                        \
    _Pragma("omp parallel for")                                                \
    for (int iy = 0; iy < num_rows; iy++)                                      \
    {                                                                          \
      if (bc == SC_ZERO)                                                       \
        y[iy] = 0;                                                             \
      else if (bc == SC_NEG_ONE)                                               \
        y[iy] = -y[iy];                                                        \
      else if (bc == SC_OTHER)                                                 \
        y[iy] *= beta;                                                         \
      int64_t ma = pntr[iy] * (int64_t)nterms * d;                             \
      for (int64_t jx = pntr[iy]; jx < pntr[iy + 1]; jx++)                     \
        for (int t = 0; t < nterms; t++)                                       \
        {                                                                      \
          T totalA = 1;                                                        \
          for (int k = 0; k < d; k++)                                          \
            totalA *= vA[iA[ma++]];                                            \
          if (ac == SC_ONE)                                                    \
            y[iy] += totalA * x[indx[jx]];                                     \
          else if (ac == SC_NEG_ONE)                                           \
            y[iy] -= totalA * x[indx[jx]];                                     \
          else                                                                 \
            y[iy] += alpha * totalA * x[indx[jx]];                             \
        }                                                                      \
    }                                                                          \
  }

DEFINE_SPARSE_N1(double, sparse_n1_f64)
DEFINE_SPARSE_N1(float, sparse_n1_f32)

/* constant-size instantiations */
#define INST(T, BODY, D, NN)                                                   \
  static void BODY##_##D##_##NN(int num_rows, const int64_t *pntr,             \
                                const int *indx, int nterms, const int *iA,    \
                                const T *vA, T alpha, const T *x, T beta,      \
                                T *y)                                          \
  {                                                                            \
    BODY(D, NN, num_rows, pntr, indx, nterms, iA, vA, alpha, x, beta, y);      \
  }
#define INST_N(T, BODY, D)                                                     \
  INST(T, BODY, D, 2) INST(T, BODY, D, 3) INST(T, BODY, D, 4)
#define INST_ALL(T, BODY)                                                      \
  INST_N(T, BODY, 1)                                                           \
  INST_N(T, BODY, 2)                                                           \
  INST_N(T, BODY, 3)                                                           \
  INST_N(T, BODY, 4) INST_N(T, BODY, 5) INST_N(T, BODY, 6)

INST_ALL(double, sparse_body_f64)
INST_ALL(float, sparse_body_f32)

typedef void (*sparse_fn_f64)(int, const int64_t *, const int *, int,
                              const int *, const double *, double,
                              const double *, double, double *);
typedef void (*sparse_fn_f32)(int, const int64_t *, const int *, int,
                              const int *, const float *, float, const float *,
                              float, float *);

#define TABLE_ROW(BODY, D) {BODY##_##D##_2, BODY##_##D##_3, BODY##_##D##_4}
static const sparse_fn_f64 table_f64[6][3] = {
    TABLE_ROW(sparse_body_f64, 1), TABLE_ROW(sparse_body_f64, 2),
    TABLE_ROW(sparse_body_f64, 3), TABLE_ROW(sparse_body_f64, 4),
    TABLE_ROW(sparse_body_f64, 5), TABLE_ROW(sparse_body_f64, 6)};
static const sparse_fn_f32 table_f32[6][3] = {
    TABLE_ROW(sparse_body_f32, 1), TABLE_ROW(sparse_body_f32, 2),
    TABLE_ROW(sparse_body_f32, 3), TABLE_ROW(sparse_body_f32, 4),
    TABLE_ROW(sparse_body_f32, 5), TABLE_ROW(sparse_body_f32, 6)};

int orc_kronmult_sparse(int d, int n, int num_rows, const int64_t *pntr,
                        const int *indx, int T, const int *iA,
                        const double *vA, double alpha, const double *x,
                        double beta, double *y)
{
  if (d < 1 || d > 6 || n < 1)
    return -1;
  if (n == 1)
    sparse_n1_f64(d, num_rows, pntr, indx, T, iA, vA, alpha, x, beta, y);
  else if (n <= 4)
    table_f64[d - 1][n - 2](num_rows, pntr, indx, T, iA, vA, alpha, x, beta, y);
  else
    sparse_body_f64(d, n, num_rows, pntr, indx, T, iA, vA, alpha, x, beta, y);
  return 0;
}

int orc_kronmult_sparse_f32(int d, int n, int num_rows, const int64_t *pntr,
                            const int *indx, int T, const int *iA,
                            const float *vA, float alpha, const float *x,
                            float beta, float *y)
{
  if (d < 1 || d > 6 || n < 1)
    return -1;
  if (n == 1)
    sparse_n1_f32(d, num_rows, pntr, indx, T, iA, vA, alpha, x, beta, y);
  else if (n <= 4)
    table_f32[d - 1][n - 2](num_rows, pntr, indx, T, iA, vA, alpha, x, beta, y);
  else
    sparse_body_f32(d, n, num_rows, pntr, indx, T, iA, vA, alpha, x, beta, y);
  return 0;
}

/* kronmult::cpu_dense, src/device/asgard_kronmult_cpu.cpp:158-511 and cpu_n0
 * (:26-60).  The factor of dimension k for (row, col) is the block
 * terms[t][k*vstride + n*n*(elem[col*d+k]*nb + elem[row*d+k])]. */
int orc_kronmult_dense(int d, int n, int num_rows, int num_cols, int T,
                       const int *elem, int row_offset, int col_offset,
                       const double *const *terms, int nb, double alpha,
                       const double *x, double beta, double *y)
{
  if (d < 1 || d > 6 || n < 1)
    return -1;
  const int64_t N = ipow_i(n, d);
  const int64_t vstride = (int64_t)nb * nb * n * n;
  const enum scalar_case ac = classify(alpha);
  const enum scalar_case bc = classify(beta);
#pragma omp parallel
  {
    double *W = (double *)malloc(sizeof(double) * (size_t)N);
    double *Y = (double *)malloc(sizeof(double) * (size_t)N);
    const double *A[6];
#pragma omp for
    for (int rowy = 0; rowy < num_rows; rowy++)
    {
      double *yrow = y + (int64_t)rowy * N;
      beta_f64(N, bc, beta, yrow);
      const int *iy = elem + (int64_t)(rowy + row_offset) * d;
      for (int colx = 0; colx < num_cols; colx++)
      {
        const int *ix = elem + (int64_t)(colx + col_offset) * d;
        const double *xc = x + (int64_t)colx * N;
        for (int t = 0; t < T; t++)
        {
          if (n == 1)
          {
            double totalA = 1;
            for (int k = 0; k < d; k++)
              totalA *= terms[t][k * (int64_t)nb * nb + (int64_t)ix[k] * nb + iy[k]];
            if (ac == SC_ONE)
              yrow[0] += totalA * xc[0];
            else if (ac == SC_NEG_ONE)
              yrow[0] -= totalA * xc[0];
            else
              yrow[0] += alpha * totalA * xc[0];
            continue;
          }
          for (int k = 0; k < d; k++)
            A[k] = terms[t] + k * vstride +
                   (int64_t)n * n * ((int64_t)ix[k] * nb + iy[k]);
          /* stages as in contract_f64 but with per-dim base pointers */
          int64_t outer = 1, inner = ipow_i(n, d - 1);
          const double *in = xc;
          double *out = (d % 2 == 1) ? Y : W;
          for (int k = 0; k < d; k++)
          {
            stage_f64(n, outer, inner, A[k], in, out);
            in = out;
            out = (out == W) ? Y : W;
            outer *= n;
            inner /= n;
          }
          accum_f64(N, ac, alpha, in, yrow);
        }
      }
    }
    free(W);
    free(Y);
  }
  return 0;
}

void orc_kronecker(int m, const double *A, int n, const double *B,
                   double *result)
{
  for (int jm = 0; jm < m; jm++)
    for (int jn = 0; jn < n; jn++)
      for (int im = 0; im < m; im++)
        for (int in = 0; in < n; in++)
          result[(int64_t)(jm * n + jn) * (m * n) + im * n + in] =
              A[jm * m + im] * B[jn * n + in];
}

void orc_kronmult_one_bruteforce(int d, int n, const double *const *A,
                                 const double *x, double *y)
{
  /* kron = A_{d-1}; kron = A_i (x) kron for i = d-2 .. 0; y += kron * x */
  int64_t total = n;
  double *kron = (double *)malloc(sizeof(double) * (size_t)(n * n));
  memcpy(kron, A[d - 1], sizeof(double) * (size_t)(n * n));
  for (int i = d - 2; i >= 0; i--)
  {
    double *next = (double *)malloc(sizeof(double) * (size_t)(total * n * total * n));
    orc_kronecker(n, A[i], (int)total, kron, next);
    free(kron);
    kron = next;
    total *= n;
  }
  for (int64_t r = 0; r < total; r++)
  {
    double s = 0;
    for (int64_t c = 0; c < total; c++)
      s += kron[c * total + r] * x[c];
    y[r] += s;
  }
  free(kron);
}

int orc_num_threads(void)
{
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

void orc_set_num_threads(int n)
{
#ifdef _OPENMP
  omp_set_num_threads(n);
#else
  (void)n;
#endif
}
