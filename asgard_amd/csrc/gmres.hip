// Device-resident restarted GMRES for the implicit (backward-Euler) step
// (I - dt*A) x = b: the algorithm of solver::simple_gmres
// (src/solver.cpp:156-298) with every vector and Hessenberg operation on the
// GPU.  Per inner iteration: one fused apply (kronmult + identity), the
// classical Gram-Schmidt dot products V^T w, the update w -= V h fused with
// the partial sums of |w|^2, one single-wave kernel that finishes the norm,
// rotates the new Hessenberg column and the right-hand side (reference BLAS
// rot/rotg semantics) and writes the residual estimate, and the scaling of w.
// Reductions use fixed block partitions and orders, so results repeat
// bit for bit.  The host reads one scalar per iteration to test convergence.
#include "matrix.hpp"

#include <algorithm>
#include <cmath>
#include <vector>

namespace akm
{
constexpr int gm_threads = 256;
constexpr int gm_cols = 8; // Krylov columns per dot-product block

template<typename P>
struct gm_state
{
  P nrm;       // norm of the last orthogonalised vector
  P inner_res; // residual estimate |s_j g_j|
  P inv_nrm;   // 1 / nrm (0 if nrm == 0)
  P pad;
};

template<typename P>
__device__ __forceinline__ P block_sum(P v, P *sh)
{
  // fixed-order block reduction (deterministic)
  const int tid = threadIdx.x;
  sh[tid] = v;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1)
  {
    if (tid < s)
      sh[tid] += sh[tid + s];
    __syncthreads();
  }
  P r = sh[0];
  __syncthreads();
  return r;
}

//! partial[k*nblk + bx] = sum over rows of block bx of V_k[i] * w[i],
//! for k in [k0, k0 + gm_cols) intersect [0, ncols).
template<typename P>
__global__ void __launch_bounds__(gm_threads)
gm_dots(int64_t n, int64_t rows_per_block, const P *V, int64_t ldv, const P *w, int ncols,
        P *partial)
{
  __shared__ P sh[gm_threads];
  const int64_t r0 = blockIdx.x * rows_per_block;
  const int64_t r1 = min(n, r0 + rows_per_block);
  const int k0 = blockIdx.y * gm_cols;
  P acc[gm_cols];
#pragma unroll
  for (int c = 0; c < gm_cols; c++)
    acc[c] = 0;
  for (int64_t i = r0 + threadIdx.x; i < r1; i += gm_threads)
  {
    const P wi = w[i];
#pragma unroll
    for (int c = 0; c < gm_cols; c++)
      if (k0 + c < ncols)
        acc[c] = fma(V[(k0 + c) * ldv + i], wi, acc[c]);
  }
#pragma unroll
  for (int c = 0; c < gm_cols; c++)
  {
    P s = block_sum(acc[c], sh);
    if (threadIdx.x == 0 && k0 + c < ncols)
      partial[int64_t(k0 + c) * gridDim.x + blockIdx.x] = s;
  }
}

//! h[k] = sum_b partial[k*nblk + b] (one block per k).
template<typename P>
__global__ void __launch_bounds__(gm_threads) gm_reduce(int nblk, const P *partial, P *h)
{
  __shared__ P sh[gm_threads];
  P s = 0;
  for (int b = threadIdx.x; b < nblk; b += gm_threads)
    s += partial[int64_t(blockIdx.x) * nblk + b];
  s = block_sum(s, sh);
  if (threadIdx.x == 0)
    h[blockIdx.x] = s;
}

//! w -= V[:, 0:ncols] h; sq_partial[bx] = sum of w^2 over the block's rows.
template<typename P>
__global__ void __launch_bounds__(gm_threads)
gm_update(int64_t n, int64_t rows_per_block, const P *V, int64_t ldv, P *w, int ncols,
          const P *h, P *sq_partial)
{
  __shared__ P sh[gm_threads];
  extern __shared__ unsigned char hs_raw[];
  P *hs = reinterpret_cast<P *>(hs_raw);
  for (int k = threadIdx.x; k < ncols; k += gm_threads)
    hs[k] = h[k];
  __syncthreads();
  const int64_t r0 = blockIdx.x * rows_per_block;
  const int64_t r1 = min(n, r0 + rows_per_block);
  P sq = 0;
  for (int64_t i = r0 + threadIdx.x; i < r1; i += gm_threads)
  {
    P s = 0;
    for (int k = 0; k < ncols; k++)
      s = fma(V[k * ldv + i], hs[k], s);
    const P v = w[i] - s;
    w[i] = v;
    sq = fma(v, v, sq);
  }
  sq = block_sum(sq, sh);
  if (threadIdx.x == 0)
    sq_partial[blockIdx.x] = sq;
}

template<typename P>
__device__ __forceinline__ void dev_rot(P &x, P &y, P c, P s)
{
  const P t = c * x + s * y;
  y = c * y - s * x;
  x = t;
}

//! reference BLAS rotg: r = +-sqrt(a^2+b^2) with the sign of the larger
//! magnitude entry, c = a/r, s = b/r; a <- r.
template<typename P>
__device__ __forceinline__ void dev_rotg(P &a, P b, P &c, P &s)
{
  const P roe = (fabs(a) > fabs(b)) ? a : b;
  const P scale = fabs(a) + fabs(b);
  if (scale == P{0})
  {
    c = 1;
    s = 0;
    a = 0;
    return;
  }
  const P as = a / scale, bs = b / scale;
  P r = scale * sqrt(as * as + bs * bs);
  r = (roe >= P{0}) ? r : -r;
  c = a / r;
  s = b / r;
  a = r;
}

//! Finishes the norm from the block partials.  With j < 0 this starts a
//! restart cycle (g[0] = |r|); otherwise it processes Hessenberg column j:
//! stores h, applies the previous rotations, computes the new one, the
//! residual estimate and (when still above tol) rotates g.
template<typename P>
__global__ void __launch_bounds__(64)
gm_givens(int nblk, const P *sq_partial, const P *h, int j, P *H, P *cs, P *sn, P *g, P tol,
          gm_state<P> *st)
{
  __shared__ P sh[64];
  P s = 0;
  for (int b = threadIdx.x; b < nblk; b += 64)
    s += sq_partial[b];
  s = block_sum(s, sh);
  if (threadIdx.x != 0)
    return;
  const P nrm = sqrt(s);
  st->nrm = nrm;
  st->inv_nrm = (nrm > P{0}) ? P{1} / nrm : P{0};
  if (j < 0)
  {
    g[0] = nrm;
    st->inner_res = nrm;
    return;
  }
  P *col = H + (j * (j + 1)) / 2; // packed upper triangle, column j
  for (int k = 0; k <= j; k++)
    col[k] = h[k];
  for (int k = 0; k < j; k++)
    dev_rot(col[k], col[k + 1], cs[k], sn[k]);
  P c, sv;
  dev_rotg(col[j], nrm, c, sv);
  cs[j] = c;
  sn[j] = sv;
  const P res = fabs(sv * g[j]);
  st->inner_res = res;
  if (res > tol)
  {
    g[j + 1] = 0;
    dev_rot(g[j], g[j + 1], c, sv);
  }
}

template<typename P>
__global__ void __launch_bounds__(gm_threads) gm_scale(int64_t n, P *w, const gm_state<P> *st)
{
  const P inv = st->inv_nrm;
  if (inv == P{0})
    return;
  for (int64_t i = blockIdx.x * int64_t(gm_threads) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * gm_threads)
    w[i] *= inv;
}

//! Packed upper-triangular back-solve H y = g (tpsv 'U','N','N'), in g.
template<typename P>
__global__ void gm_backsolve(int m, const P *H, P *g)
{
  if (threadIdx.x != 0 || blockIdx.x != 0)
    return;
  for (int j = m - 1; j >= 0; j--)
  {
    g[j] /= H[j + (j * (j + 1)) / 2];
    const P gj = g[j];
    for (int i = 0; i < j; i++)
      g[i] -= gj * H[i + (j * (j + 1)) / 2];
  }
}

//! x += V[:, 0:m] y
template<typename P>
__global__ void __launch_bounds__(gm_threads)
gm_xupdate(int64_t n, const P *V, int64_t ldv, int m, const P *y, P *x)
{
  extern __shared__ unsigned char ys_raw[];
  P *ys = reinterpret_cast<P *>(ys_raw);
  for (int k = threadIdx.x; k < m; k += gm_threads)
    ys[k] = y[k];
  __syncthreads();
  for (int64_t i = blockIdx.x * int64_t(gm_threads) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * gm_threads)
  {
    P s = 0;
    for (int k = 0; k < m; k++)
      s = fma(V[k * ldv + i], ys[k], s);
    x[i] += s;
  }
}

} // namespace akm

//! Device workspace of the solver, kept with the matrix between solves.
struct akm_gmres_workspace
{
  int restart = 0;
  int64_t n = 0;
  size_t esize = 8;
  akm::dev_buffer<unsigned char> V, work, xb;
  akm::dev_buffer<unsigned char> state;
  void *host_state = nullptr; // pinned
  ~akm_gmres_workspace()
  {
    if (host_state != nullptr)
      (void)hipHostFree(host_state);
  }
};

static void delete_workspace(akm_gmres_workspace *w) { delete w; }

namespace akm
{
template<typename P>
static int gmres_typed(akm_matrix_s &m, P dt, P *x, const P *b, int restart, int max_outer,
                       P tol, double *err, int *iterations)
{
  const int64_t n = m.output_size();
  hipStream_t s = m.stream;
  auto &ws = *m.gmres;
  P *V = reinterpret_cast<P *>(ws.V.data());
  const int64_t ldv = n;
  // small work arrays: h, H, cs, sn, g, dot partials, sq partials
  const int64_t rows_per_block = std::max<int64_t>(4096, (n + 1023) / 1024);
  const int nblk = static_cast<int>((n + rows_per_block - 1) / rows_per_block);
  const int ncolblk_max = (restart + 1 + gm_cols - 1) / gm_cols;
  P *h = reinterpret_cast<P *>(ws.work.data());
  P *H = h + (restart + 1);
  P *cs = H + int64_t(restart) * (restart + 1) / 2 + 1;
  P *sn = cs + (restart + 1);
  P *g = sn + (restart + 1);
  P *dotp = g + (restart + 2);
  P *sqp = dotp + int64_t(restart + 1) * nblk;
  (void)ncolblk_max;
  auto *st = reinterpret_cast<gm_state<P> *>(ws.state.data());
  auto *hst = static_cast<gm_state<P> *>(ws.host_state);
  const int sblocks = static_cast<int>(std::min<int64_t>((n + gm_threads - 1) / gm_threads, 2048));

  auto read_state = [&]() -> int {
    AKM_HIP(hipMemcpyAsync(hst, st, sizeof(gm_state<P>), hipMemcpyDeviceToHost, s));
    AKM_HIP(hipStreamSynchronize(s));
    return AKM_SUCCESS;
  };

  int total = 0, outer = 0, inner = 0;
  P inner_res = 0, outer_res = tol + 1;
  while (outer_res > tol && outer < max_outer)
  {
    // r = b - (I - dt A) x into V_0
    AKM_HIP(hipMemcpyAsync(V, b, sizeof(P) * n, hipMemcpyDeviceToDevice, s));
    AKM_TRY(apply_device(m, -dt * P{-1}, x, P{1}, V, P{-1}, x));
    ++total;
    gm_update<P><<<nblk, gm_threads, 0, s>>>(n, rows_per_block, V, ldv, V, 0, h, sqp);
    gm_givens<P><<<1, 64, 0, s>>>(nblk, sqp, h, -1, H, cs, sn, g, tol, st);
    gm_scale<P><<<sblocks, gm_threads, 0, s>>>(n, V, st);
    AKM_TRY(check_launch("gmres restart"));
    AKM_TRY(read_state());
    inner_res = hst->inner_res;

    inner = 0;
    while (inner_res > tol && inner < restart)
    {
      const P *vj = V + int64_t(inner) * ldv;
      P *w = V + int64_t(inner + 1) * ldv;
      AKM_TRY(apply_device(m, -dt, vj, P{0}, w, P{1}, vj));
      ++total;
      const int ncols = inner + 1;
      dim3 grid(nblk, (ncols + gm_cols - 1) / gm_cols);
      gm_dots<P><<<grid, gm_threads, 0, s>>>(n, rows_per_block, V, ldv, w, ncols, dotp);
      gm_reduce<P><<<ncols, gm_threads, 0, s>>>(nblk, dotp, h);
      gm_update<P><<<nblk, gm_threads, sizeof(P) * ncols, s>>>(n, rows_per_block, V, ldv, w,
                                                               ncols, h, sqp);
      gm_givens<P><<<1, 64, 0, s>>>(nblk, sqp, h, inner, H, cs, sn, g, tol, st);
      gm_scale<P><<<sblocks, gm_threads, 0, s>>>(n, w, st);
      AKM_TRY(check_launch("gmres inner iteration"));
      AKM_TRY(read_state());
      inner_res = hst->inner_res;
      ++inner;
    }
    if (inner > 0)
    {
      gm_backsolve<P><<<1, 1, 0, s>>>(inner, H, g);
      gm_xupdate<P><<<sblocks, gm_threads, sizeof(P) * inner, s>>>(n, V, ldv, inner, g, x);
      AKM_TRY(check_launch("gmres update"));
    }
    ++outer;
    outer_res = inner_res;
  }
  AKM_HIP(hipStreamSynchronize(s));
  if (err)
    *err = static_cast<double>(outer_res);
  if (iterations)
    *iterations = total;
  return (outer_res > tol) ? AKM_ERR_NOT_CONVERGED : AKM_SUCCESS;
}

//! default_gmres_restarts (src/solver.cpp:142-152): clamp(512 MB / size, min(10, n), min(200, n)).
static int default_restarts(int64_t n, size_t esize)
{
  const int minimum = static_cast<int>(std::min<int64_t>(10, n));
  const int maximum = static_cast<int>(std::min<int64_t>(200, n));
  const double mb = double(n) * double(esize) / (1024.0 * 1024.0);
  int guess = static_cast<int>(512.0 / mb);
  return std::clamp(guess, minimum, maximum);
}

static int ensure_workspace(akm_matrix_s &m, int restart)
{
  const int64_t n = m.output_size();
  if (!m.gmres)
    m.gmres = std::unique_ptr<akm_gmres_workspace, void (*)(akm_gmres_workspace *)>(
        new akm_gmres_workspace(), delete_workspace);
  auto &ws = *m.gmres;
  if (ws.restart == restart && ws.n == n && ws.esize == m.esize)
    return AKM_SUCCESS;
  ws.restart = restart;
  ws.n = n;
  ws.esize = m.esize;
  AKM_TRY(ws.V.allocate(n * (restart + 1) * int64_t(m.esize)));
  const int64_t rows_per_block = std::max<int64_t>(4096, (n + 1023) / 1024);
  const int64_t nblk = (n + rows_per_block - 1) / rows_per_block;
  const int64_t small = 4 * int64_t(restart + 2) + int64_t(restart) * (restart + 1) / 2 + 2 +
                        int64_t(restart + 1) * nblk + nblk + 16;
  AKM_TRY(ws.work.allocate(small * int64_t(m.esize)));
  AKM_TRY(ws.state.allocate(sizeof(gm_state<double>)));
  AKM_TRY(ws.xb.allocate(2 * n * int64_t(m.esize)));
  if (ws.host_state == nullptr)
    AKM_HIP(hipHostMalloc(&ws.host_state, sizeof(gm_state<double>), hipHostMallocDefault));
  return AKM_SUCCESS;
}

} // namespace akm

using namespace akm;

extern "C" akm_status akm_gmres_euler(akm_matrix mat, double dt, void *x, const void *b,
                                      int restart, int max_iter, double tol, int on_device,
                                      double *err, int *iterations)
{
  if (mat == nullptr || x == nullptr || b == nullptr)
    return set_error(AKM_ERR_INVALID, "gmres: null argument");
  if (mat->rows != mat->cols)
    return set_error(AKM_ERR_INVALID, "gmres: the operator must be square");
  const int64_t n = mat->output_size();
  if (n <= 0)
    return set_error(AKM_ERR_INVALID, "gmres: empty system");
  if (tol <= 0)
    tol = (mat->prec == AKM_F64) ? 1e-12 : 1e-6;
  if (restart <= 0)
    restart = default_restarts(n, mat->esize);
  if (restart > n)
    return set_error(AKM_ERR_INVALID, "Number of inner iterations %d must be less than %lld!",
                     restart, static_cast<long long>(n));
  if (max_iter <= 0)
    max_iter = static_cast<int>(std::min<int64_t>(n, 1 << 30));
  AKM_TRY(ensure_workspace(*mat, restart));
  void *xd = x;
  const void *bd = b;
  auto &ws = *mat->gmres;
  const size_t bytes = size_t(n) * mat->esize;
  if (!on_device)
  {
    xd = ws.xb.data();
    bd = ws.xb.data() + bytes;
    AKM_HIP(hipMemcpyAsync(xd, x, bytes, hipMemcpyHostToDevice, mat->stream));
    AKM_HIP(hipMemcpyAsync(const_cast<void *>(bd), b, bytes, hipMemcpyHostToDevice,
                           mat->stream));
  }
  int rc;
  if (mat->prec == AKM_F64)
    rc = gmres_typed<double>(*mat, dt, static_cast<double *>(xd), static_cast<const double *>(bd),
                             restart, max_iter, tol, err, iterations);
  else
    rc = gmres_typed<float>(*mat, static_cast<float>(dt), static_cast<float *>(xd),
                            static_cast<const float *>(bd), restart, max_iter,
                            static_cast<float>(tol), err, iterations);
  if (rc < 0)
    return rc;
  if (!on_device)
  {
    AKM_HIP(hipMemcpyAsync(x, xd, bytes, hipMemcpyDeviceToHost, mat->stream));
    AKM_HIP(hipStreamSynchronize(mat->stream));
  }
  return rc;
}
