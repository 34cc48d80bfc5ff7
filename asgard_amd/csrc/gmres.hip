// Device-resident restarted GMRES for the implicit (backward-Euler) step
// (I - dt*A) x = b.
//
// akm_gmres_euler is the algorithm of solver::simple_gmres
// (src/solver.cpp:156-298) with every vector and Hessenberg operation on the
// GPU.  Per inner iteration: one fused apply (kronmult + identity), the
// classical Gram-Schmidt dot products V^T w, the update w -= V h (which first
// reduces the dot partials to h) fused with the partial sums of |w|^2, and
// one kernel that finishes the norm and scales w while its first workgroup
// rotates the new Hessenberg column and the right-hand side (reference BLAS
// rot/rotg semantics) and writes the residual estimate.
// Reductions use fixed block partitions and orders, so results repeat
// bit for bit.  The host reads one scalar per iteration to test convergence.
//
// akm_gmres_euler_mixed is an FP64 iterative refinement around flexible
// GMRES cycles: the residual b - (I - dt*A)x, the Krylov basis, the
// Gram-Schmidt and the Hessenberg least squares are FP64; the operator inside
// a cycle may be the FP32 kronmult (vectors rounded to FP32 for the apply
// only) and a right preconditioner may act on every basis vector.  After each
// cycle the true FP64 residual decides: it must at least halve, else the
// cycle is undone and plain FP64 GMRES finishes the solve.
#include "matrix.hpp"
#include "precond.hpp"

#include <algorithm>
#include <cmath>
#include <functional>
#include <limits>
#include <mutex>
#include <vector>

namespace akm
{
constexpr int gm_threads = 256;
constexpr int gm_cols = 8; // Krylov columns per dot-product block
// the Givens kernel stages 3 * restart values in LDS (64 KB)
constexpr int gm_max_restart = 2000;
// host copies of the step state: Arnoldi steps in flight at once (the step
// waited for plus the speculative ones) stay below this
constexpr int gm_slots = 4;
// speculative steps enqueued past the one waited for (two measured no faster
// than one at d=4: the step is bound by the GPU's kernel chain, not the host)
constexpr int gm_ahead = 1;

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
  // fixed-order block reduction (deterministic): a shuffle tree inside each
  // wavefront, then the wavefronts' sums in order; two barriers
#pragma unroll
  for (int off = wave_size / 2; off > 0; off >>= 1)
    v += __shfl_xor(v, off, wave_size);
  const int wave = threadIdx.x / wave_size;
  if (threadIdx.x % wave_size == 0)
    sh[wave] = v;
  __syncthreads();
  P r = 0;
  for (int w = 0; w < static_cast<int>(blockDim.x) / wave_size; w++)
    r += sh[w];
  __syncthreads();
  return r;
}

//! Fixed-order sum over one wavefront (the shuffle tree of block_sum).
template<typename P>
__device__ __forceinline__ P wave_sum(P v)
{
#pragma unroll
  for (int off = wave_size / 2; off > 0; off >>= 1)
    v += __shfl_xor(v, off, wave_size);
  return v;
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

//! w -= V[:, 0:ncols] h and sq_partial[bx] = sum of w^2 over the block's
//! rows, with h formed first by every block from the dot partials:
//! h[k] = sum_b partial[k*nblk + b] (one wavefront per column, lanes striding
//! over the blocks, then the shuffle tree; the same order in every block).
//! Block 0 stores h.
template<typename P>
__global__ void __launch_bounds__(gm_threads)
gm_reduce_update(int64_t n, int64_t rows_per_block, const P *V, int64_t ldv, P *w, int ncols,
                 const P *partial, P *h, P *sq_partial)
{
  __shared__ P sh[gm_threads];
  extern __shared__ unsigned char hs_raw[];
  P *hs = reinterpret_cast<P *>(hs_raw);
  const int nblk = gridDim.x;
  const int wave = threadIdx.x / wave_size, lane = threadIdx.x % wave_size;
  for (int k = wave; k < ncols; k += gm_threads / wave_size)
  {
    P v = 0;
    for (int b = lane; b < nblk; b += wave_size)
      v += partial[int64_t(k) * nblk + b];
    v = wave_sum(v);
    if (lane == 0)
    {
      hs[k] = v;
      if (blockIdx.x == 0)
        h[k] = v;
    }
  }
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

//! sq_partial[bx] = sum of x^2 over the block's rows (x is only read).
template<typename P>
__global__ void __launch_bounds__(gm_threads)
gm_sumsq(int64_t n, int64_t rows_per_block, const P *x, P *sq_partial)
{
  __shared__ P sh[gm_threads];
  const int64_t r0 = blockIdx.x * rows_per_block;
  const int64_t r1 = min(n, r0 + rows_per_block);
  P sq = 0;
  for (int64_t i = r0 + threadIdx.x; i < r1; i += gm_threads)
  {
    const P v = x[i];
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
//! residual estimate and rotates g.  With rotate_always == 0, g is rotated
//! only while the estimate is above tol, as simple_gmres does
//! (src/solver.cpp:261-267).
template<typename P>
__global__ void __launch_bounds__(64)
gm_givens(int nblk, const P *sq_partial, const P *h, int j, P *H, P *cs, P *sn, P *g, P tol,
          int rotate_always, gm_state<P> *st)
{
  // column j, then the previous rotations, staged in LDS so that the serial
  // part below runs on registers and LDS (dynamic LDS: 3j + 1 values)
  __shared__ P sh[64];
  extern __shared__ unsigned char gv_raw[];
  P *lc = reinterpret_cast<P *>(gv_raw);
  P *lcs = lc + (j + 1);
  P *lsn = lcs + j;
  for (int k = threadIdx.x; k <= j; k += 64)
    lc[k] = h[k];
  for (int k = threadIdx.x; k < j; k += 64)
  {
    lcs[k] = cs[k];
    lsn[k] = sn[k];
  }
  P s = 0;
  for (int b = threadIdx.x; b < nblk; b += 64)
    s += sq_partial[b];
  s = block_sum(s, sh); // its barriers also order the LDS writes above
  if (threadIdx.x == 0)
  {
    const P nrm = sqrt(s);
    st->nrm = nrm;
    st->inv_nrm = (nrm > P{0}) ? P{1} / nrm : P{0};
    if (j < 0)
    {
      g[0] = nrm;
      st->inner_res = nrm;
    }
    else
    {
      // apply the old rotations down the column, the running entry in a register
      P x = lc[0];
      for (int k = 0; k < j; k++)
      {
        P y = lc[k + 1];
        dev_rot(x, y, lcs[k], lsn[k]);
        lc[k] = x;
        x = y;
      }
      P c, sv;
      dev_rotg(x, nrm, c, sv);
      lc[j] = x;
      cs[j] = c;
      sn[j] = sv;
      const P res = fabs(sv * g[j]);
      st->inner_res = res;
      if (rotate_always != 0 || res > tol)
      {
        g[j + 1] = 0;
        dev_rot(g[j], g[j + 1], c, sv);
      }
    }
  }
  if (j < 0)
    return;
  __syncthreads();
  P *col = H + (j * (j + 1)) / 2; // packed upper triangle, column j
  for (int k = threadIdx.x; k <= j; k += 64)
    col[k] = lc[k];
}

//! gm_givens for column j >= 0 and gm_scale in one launch: every block
//! finishes the norm from the partials (the first wavefront, in gm_givens'
//! order) and scales its rows of w; block 0 also rotates the Hessenberg
//! column and the right-hand side and writes the state (dynamic LDS:
//! 3j + 1 values).
template<typename P>
__global__ void __launch_bounds__(gm_threads)
gm_givens_scale(int64_t n, P *w, int nblk, const P *sq_partial, const P *h, int j, P *H, P *cs,
                P *sn, P *g, P tol, int rotate_always, gm_state<P> *st)
{
  __shared__ P s_nrm;
  extern __shared__ unsigned char gs_raw[];
  if (threadIdx.x < wave_size)
  {
    P s = 0;
    for (int b = threadIdx.x; b < nblk; b += wave_size)
      s += sq_partial[b];
    s = wave_sum(s);
    if (threadIdx.x == 0)
      s_nrm = sqrt(s);
  }
  P *lc = reinterpret_cast<P *>(gs_raw);
  P *lcs = lc + (j + 1);
  P *lsn = lcs + j;
  if (blockIdx.x == 0)
  {
    for (int k = threadIdx.x; k <= j; k += gm_threads)
      lc[k] = h[k];
    for (int k = threadIdx.x; k < j; k += gm_threads)
    {
      lcs[k] = cs[k];
      lsn[k] = sn[k];
    }
  }
  __syncthreads();
  const P nrm = s_nrm;
  const P inv = (nrm > P{0}) ? P{1} / nrm : P{0};
  if (blockIdx.x == 0)
  {
    if (threadIdx.x == 0)
    {
      st->nrm = nrm;
      st->inv_nrm = inv;
      P x = lc[0];
      for (int k = 0; k < j; k++)
      {
        P y = lc[k + 1];
        dev_rot(x, y, lcs[k], lsn[k]);
        lc[k] = x;
        x = y;
      }
      P c, sv;
      dev_rotg(x, nrm, c, sv);
      lc[j] = x;
      cs[j] = c;
      sn[j] = sv;
      const P res = fabs(sv * g[j]);
      st->inner_res = res;
      if (rotate_always != 0 || res > tol)
      {
        g[j + 1] = 0;
        dev_rot(g[j], g[j + 1], c, sv);
      }
    }
    __syncthreads();
    P *col = H + (j * (j + 1)) / 2; // packed upper triangle, column j
    for (int k = threadIdx.x; k <= j; k += gm_threads)
      col[k] = lc[k];
  }
  if (inv == P{0})
    return;
  for (int64_t i = blockIdx.x * int64_t(gm_threads) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * gm_threads)
    w[i] *= inv;
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
//! One workgroup; g (and H when it fits, `h_in_lds`) staged in LDS; column
//! by column, the updates of one column spread over the threads (the same
//! operations in the same order per entry as a serial solve).
template<typename P>
__global__ void __launch_bounds__(256) gm_backsolve(int m, const P *H, P *g, int h_in_lds)
{
  extern __shared__ unsigned char bs_raw[];
  P *lg = reinterpret_cast<P *>(bs_raw);
  P *lh = lg + m;
  const int tid = threadIdx.x;
  for (int i = tid; i < m; i += 256)
    lg[i] = g[i];
  if (h_in_lds)
    for (int i = tid; i < m * (m + 1) / 2; i += 256)
      lh[i] = H[i];
  __syncthreads();
  const P *Hs = h_in_lds ? lh : H;
  for (int j = m - 1; j >= 0; j--)
  {
    if (tid == 0)
    {
      const P hjj = Hs[j + (j * (j + 1)) / 2];
      // a zero diagonal only follows an exact breakdown (residual already
      // zero); that column then takes no part in the solution
      lg[j] = (hjj != P{0}) ? lg[j] / hjj : P{0};
    }
    __syncthreads();
    const P gj = lg[j];
    for (int i = tid; i < j; i += 256)
      lg[i] -= gj * Hs[i + (j * (j + 1)) / 2];
    __syncthreads();
  }
  for (int i = tid; i < m; i += 256)
    g[i] = lg[i];
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

//! r = b - r
template<typename P>
__global__ void __launch_bounds__(gm_threads) gm_rsub(int64_t n, const P *b, P *r)
{
  for (int64_t i = blockIdx.x * int64_t(gm_threads) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * gm_threads)
    r[i] = b[i] - r[i];
}

//! y = fp32(x)
__global__ void __launch_bounds__(gm_threads) gm_to_f32(int64_t n, const double *x, float *y)
{
  for (int64_t i = blockIdx.x * int64_t(gm_threads) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * gm_threads)
    y[i] = static_cast<float>(x[i]);
}

//! w = v + c * y, y in FP32 (the identity part of I - dt*A stays FP64).
__global__ void __launch_bounds__(gm_threads)
gm_axpy_f32(int64_t n, const double *v, double c, const float *y, double *w)
{
  for (int64_t i = blockIdx.x * int64_t(gm_threads) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * gm_threads)
    w[i] = fma(c, static_cast<double>(y[i]), v[i]);
}

//! y = alpha*A*x + beta*y for a dense n x n column-major A, one thread per
//! row (consecutive threads read consecutive entries of a column); beta == 0
//! never reads y.
template<typename P>
__global__ void __launch_bounds__(gm_threads)
gm_dense_gemv(int n, const P *A, const P *x, P alpha, P beta, P *y)
{
  const int i = blockIdx.x * gm_threads + threadIdx.x;
  if (i >= n)
    return;
  P s = 0;
  for (int j = 0; j < n; j++)
    s = fma(A[int64_t(j) * n + i], x[j], s);
  y[i] = (beta == P{0}) ? alpha * s : fma(alpha, s, beta * y[i]);
}

//! v = M^-1 v from the LU factors of M with partial pivoting (getrs after
//! getrf): the row interchanges piv in order, then the unit lower and the
//! upper triangular solves, column by column.  One workgroup, v in LDS.
template<typename P>
__global__ void __launch_bounds__(256) gm_lu_solve(int n, const P *LU, const int *piv, P *v)
{
  extern __shared__ unsigned char lu_raw[];
  P *lv = reinterpret_cast<P *>(lu_raw);
  const int tid = threadIdx.x;
  for (int i = tid; i < n; i += 256)
    lv[i] = v[i];
  __syncthreads();
  if (tid == 0)
    for (int k = 0; k < n; k++)
    {
      const int p = piv[k];
      if (p != k)
      {
        const P t = lv[k];
        lv[k] = lv[p];
        lv[p] = t;
      }
    }
  __syncthreads();
  for (int k = 0; k < n; k++)
  {
    const P vk = lv[k];
    for (int i = k + 1 + tid; i < n; i += 256)
      lv[i] -= LU[int64_t(k) * n + i] * vk;
    __syncthreads();
  }
  for (int k = n - 1; k >= 0; k--)
  {
    if (tid == 0)
      lv[k] /= LU[int64_t(k) * n + k];
    __syncthreads();
    const P vk = lv[k];
    for (int i = tid; i < k; i += 256)
      lv[i] -= LU[int64_t(k) * n + i] * vk;
    __syncthreads();
  }
  for (int i = tid; i < n; i += 256)
    v[i] = lv[i];
}

} // namespace akm

//! Device workspace of the solvers, kept with the matrix between solves.
struct akm_gmres_workspace
{
  int restart = 0;
  int64_t n = 0;
  size_t esize = 8;
  akm::dev_buffer<unsigned char> V, work, xb;
  akm::dev_buffer<unsigned char> state;
  // flexible / mixed-precision solver only
  akm::dev_buffer<double> Z, xprev;
  akm::dev_buffer<float> v32, y32;
  void *host_state = nullptr; // pinned, gm_slots slots
  hipEvent_t state_ready[akm::gm_slots] = {};
  ~akm_gmres_workspace()
  {
    if (host_state != nullptr)
      (void)hipHostFree(host_state);
    for (auto &e : state_ready)
      if (e != nullptr)
        (void)hipEventDestroy(e);
  }
};

static void delete_workspace(akm_gmres_workspace *w) { delete w; }

namespace akm
{
//! Rows of a vector handled by one reduction block: at least 256 (one row per
//! thread), otherwise enough blocks to cover the chip (about 1024) and no
//! more, so the per-column partial sums stay short.
static int64_t rows_per_block_for(int64_t n)
{
  return std::max<int64_t>(256, (n + 1023) / 1024);
}

//! Arnoldi steps are enqueued one ahead of the host's convergence check when
//! the operator apply takes less than this (milliseconds).
constexpr double speculate_below_ms = 0.25;

//! Runs f (which enqueues work on s) and measures it with HIP events.
template<typename F>
static int timed_on_stream(hipStream_t s, F &&f, double &ms)
{
  hipEvent_t e0, e1;
  AKM_HIP(hipEventCreate(&e0));
  AKM_HIP(hipEventCreate(&e1));
  int rc = AKM_SUCCESS;
  if (hipEventRecord(e0, s) != hipSuccess)
    rc = set_error(AKM_ERR_HIP, "gmres: event record failed");
  if (rc == AKM_SUCCESS)
    rc = f();
  float t = 0;
  if (rc == AKM_SUCCESS &&
      (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
       hipEventElapsedTime(&t, e0, e1) != hipSuccess))
    rc = set_error(AKM_ERR_HIP, "gmres: event timing failed");
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  ms = t;
  return rc;
}

//! Memory of one solve: the Krylov basis V (n x (restart + 1)), the small
//! work array (h, packed H, cs, sn, g, dot and norm partials:
//! gm_work_values of them), the device step state, gm_slots host copies of
//! it, and gm_slots events marking those copies (or null: the host then
//! waits for the whole stream).  Every value is written by a kernel before
//! any kernel or the host reads it, so none needs clearing between solves.
struct gm_memory
{
  void *V = nullptr;
  void *work = nullptr;
  void *state = nullptr;
  void *host_state = nullptr;
  hipEvent_t *ready = nullptr;
};

//! Values in the small work array of a solve.
static int64_t gm_work_values(int64_t n, int restart)
{
  const int64_t rpb = rows_per_block_for(n);
  const int64_t nblk = (n + rpb - 1) / rpb;
  return 4 * int64_t(restart + 2) + int64_t(restart) * (restart + 1) / 2 + 2 +
         int64_t(restart + 1) * nblk + nblk + 16;
}

//! The memory of a solve in a matrix's workspace.
static gm_memory memory_of(akm_gmres_workspace &ws)
{
  gm_memory mem;
  mem.V = ws.V.data();
  mem.work = ws.work.data();
  mem.state = ws.state.data();
  mem.host_state = ws.host_state;
  mem.ready = ws.state_ready;
  return mem;
}

//! One restart cycle's device state and the kernel sequence of its steps.
template<typename P>
struct gm_cycle
{
  int64_t n = 0;
  hipStream_t s = nullptr;
  int restart = 0;
  P *V = nullptr;
  P *Z = nullptr; // preconditioned directions (flexible GMRES), else null
  int64_t ldv = 0;
  int64_t rpb = 0;
  int nblk = 0, sblocks = 0;
  P *h = nullptr, *H = nullptr, *cs = nullptr, *sn = nullptr, *g = nullptr;
  P *dotp = nullptr, *sqp = nullptr;
  gm_state<P> *st = nullptr, *hst = nullptr;
  hipEvent_t *ready = nullptr;
  int rotate_always = 0;

  gm_cycle(gm_memory const &mem, int64_t n_, hipStream_t s_, int restart_, P *Z_,
           int rotate_always_)
      : n(n_), s(s_), restart(restart_), Z(Z_), rotate_always(rotate_always_)
  {
    V = static_cast<P *>(mem.V);
    ldv = n;
    rpb = rows_per_block_for(n);
    nblk = static_cast<int>((n + rpb - 1) / rpb);
    sblocks = static_cast<int>(std::min<int64_t>((n + gm_threads - 1) / gm_threads, 2048));
    // small work arrays: h, H (packed), cs, sn, g, dot partials, sq partials
    h = static_cast<P *>(mem.work);
    H = h + (restart + 1);
    cs = H + int64_t(restart) * (restart + 1) / 2 + 1;
    sn = cs + (restart + 1);
    g = sn + (restart + 1);
    dotp = g + (restart + 2);
    sqp = dotp + int64_t(restart + 1) * nblk;
    st = static_cast<gm_state<P> *>(mem.state);
    hst = static_cast<gm_state<P> *>(mem.host_state);
    ready = mem.ready;
  }

  int sync_state()
  {
    AKM_HIP(hipMemcpyAsync(hst, st, sizeof(gm_state<P>), hipMemcpyDeviceToHost, s));
    AKM_HIP(hipStreamSynchronize(s));
    return AKM_SUCCESS;
  }

  //! |x| of a vector that is only read.
  int norm(const P *x, P &out)
  {
    gm_sumsq<P><<<nblk, gm_threads, 0, s>>>(n, rpb, x, sqp);
    gm_givens<P><<<1, 64, 0, s>>>(nblk, sqp, h, -1, H, cs, sn, g, P{0}, rotate_always, st);
    AKM_TRY(check_launch("gmres norm"));
    AKM_TRY(sync_state());
    out = hst->nrm;
    return AKM_SUCCESS;
  }

  //! V_0 holds a residual r: g_0 = |r|, V_0 = r / |r|; beta = |r|.
  int start(P tol, P &beta)
  {
    gm_sumsq<P><<<nblk, gm_threads, 0, s>>>(n, rpb, V, sqp);
    gm_givens<P><<<1, 64, 0, s>>>(nblk, sqp, h, -1, H, cs, sn, g, tol, rotate_always, st);
    gm_scale<P><<<sblocks, gm_threads, 0, s>>>(n, V, st);
    AKM_TRY(check_launch("gmres restart"));
    AKM_TRY(sync_state());
    beta = hst->inner_res;
    return AKM_SUCCESS;
  }

  //! start() without the host's wait: the state goes to the last host slot,
  //! which run_steps reads (with beta_pending) after enqueueing its first
  //! steps, so the GPU runs on through the round trip.  Needs events.
  int start_async(P tol)
  {
    static_assert(gm_ahead + 1 < gm_slots, "the first steps in flight reach the start's slot");
    gm_sumsq<P><<<nblk, gm_threads, 0, s>>>(n, rpb, V, sqp);
    gm_givens<P><<<1, 64, 0, s>>>(nblk, sqp, h, -1, H, cs, sn, g, tol, rotate_always, st);
    gm_scale<P><<<sblocks, gm_threads, 0, s>>>(n, V, st);
    AKM_TRY(check_launch("gmres restart"));
    AKM_HIP(hipMemcpyAsync(hst + (gm_slots - 1), st, sizeof(gm_state<P>),
                           hipMemcpyDeviceToHost, s));
    AKM_HIP(hipEventRecord(ready[gm_slots - 1], s));
    return AKM_SUCCESS;
  }

  //! Arnoldi step j: z_j = pre(v_j) (flexible) or v_j, w = op(z_j), CGS
  //! against V_0..V_j, Givens; est = new residual estimate.
  int step(int j, std::function<int(const P *, P *)> const &op,
           std::function<int(const P *, P *)> const *pre, P tol, P &est)
  {
    AKM_TRY(launch_step(j, op, pre, tol));
    return wait_step(j, est);
  }

  //! Enqueues Arnoldi step j and the copy of its state into host slot j % 2.
  int launch_step(int j, std::function<int(const P *, P *)> const &op,
                  std::function<int(const P *, P *)> const *pre, P tol)
  {
    const P *vj = V + int64_t(j) * ldv;
    P *w = V + int64_t(j + 1) * ldv;
    const P *zj = vj;
    if (pre != nullptr)
    {
      P *zz = Z + int64_t(j) * ldv;
      AKM_TRY((*pre)(vj, zz));
      zj = zz;
    }
    AKM_TRY(op(zj, w));
    const int ncols = j + 1;
    dim3 grid(nblk, (ncols + gm_cols - 1) / gm_cols);
    gm_dots<P><<<grid, gm_threads, 0, s>>>(n, rpb, V, ldv, w, ncols, dotp);
    gm_reduce_update<P><<<nblk, gm_threads, sizeof(P) * ncols, s>>>(n, rpb, V, ldv, w, ncols,
                                                                      dotp, h, sqp);
    gm_givens_scale<P><<<sblocks, gm_threads, sizeof(P) * (3 * j + 1), s>>>(
        n, w, nblk, sqp, h, j, H, cs, sn, g, tol, rotate_always, st);
    AKM_TRY(check_launch("gmres inner iteration"));
    AKM_HIP(hipMemcpyAsync(hst + (j % gm_slots), st, sizeof(gm_state<P>),
                           hipMemcpyDeviceToHost, s));
    if (ready != nullptr)
      AKM_HIP(hipEventRecord(ready[j % gm_slots], s));
    return AKM_SUCCESS;
  }

  //! Arnoldi steps 0, 1, ... until the estimate reaches `stop` or `limit`
  //! steps ran; count = the steps taken.  With `speculate`, steps j + 1 ..
  //! j + gm_ahead are enqueued before step j's estimate is read, which keeps
  //! the GPU busy through the host's round trips.  When step j ends the
  //! cycle the speculative steps' columns, Hessenberg columns, rotations and
  //! g entries are never used (the update reads columns and g entries below
  //! count only), and the stream runs them before any later work.
  int run_steps(int limit, std::function<int(const P *, P *)> const &op,
                std::function<int(const P *, P *)> const *pre, P tol, P stop, P &est,
                int &count, bool speculate, bool beta_pending = false)
  {
    static_assert(gm_ahead + 1 <= gm_slots, "steps in flight exceed the state slots");
    count = 0;
    int launched = 0;
    const int depth = speculate ? gm_ahead + 1 : 1;
    if (beta_pending)
    {
      // the first steps go out before the start's residual norm is known;
      // when that norm already meets `stop` they are never used (count 0)
      while (launched < depth && launched < limit)
        AKM_TRY(launch_step(launched++, op, pre, tol));
      AKM_HIP(hipEventSynchronize(ready[gm_slots - 1]));
      est = hst[gm_slots - 1].inner_res;
    }
    while (est > stop && count < limit)
    {
      while (launched < count + depth && launched < limit)
        AKM_TRY(launch_step(launched++, op, pre, tol));
      AKM_TRY(wait_step(count, est));
      ++count;
    }
    return AKM_SUCCESS;
  }

  //! Waits for step j's state; est = its residual estimate.
  int wait_step(int j, P &est)
  {
    if (ready != nullptr)
      AKM_HIP(hipEventSynchronize(ready[j % gm_slots]));
    else
      AKM_HIP(hipStreamSynchronize(s));
    est = hst[j % gm_slots].inner_res;
    return AKM_SUCCESS;
  }

  //! x += B[:, 0:m] H^-1 g with B = Z (flexible) or V.
  int finish(int m, P *x, bool flexible)
  {
    if (m <= 0)
      return AKM_SUCCESS;
    const size_t hbytes = sizeof(P) * (size_t(m) * (m + 1) / 2);
    const int h_in_lds = (hbytes + sizeof(P) * m <= 48 * 1024) ? 1 : 0;
    gm_backsolve<P><<<1, 256, sizeof(P) * m + (h_in_lds ? hbytes : 0), s>>>(m, H, g, h_in_lds);
    gm_xupdate<P><<<sblocks, gm_threads, sizeof(P) * m, s>>>(n, flexible ? Z : V, ldv, m, g, x);
    return check_launch("gmres update");
  }
};

//! The restart loop of simple_gmres (src/solver.cpp:201-295) on a cycle:
//! `residual(outer, speculate)` writes b - op(x) (preconditioned, for a left
//! preconditioner) into V_0 and may set `speculate`; then the Arnoldi steps,
//! the update of x and the absolute test of the residual estimate.
template<typename P, typename R>
static int gmres_outer(gm_cycle<P> &cyc, R &&residual,
                       std::function<int(const P *, P *)> const &op, P *x, int restart,
                       int max_outer, P tol, double *err, int *iterations)
{
  int total = 0, outer = 0;
  P inner_res = 0, outer_res = tol + 1;
  bool speculate = false;
  while (outer_res > tol && outer < max_outer)
  {
    AKM_TRY(residual(outer, speculate));
    ++total;
    // with events the host reads the start's norm only once the first
    // steps are enqueued
    const bool pending = cyc.ready != nullptr;
    if (pending)
      AKM_TRY(cyc.start_async(tol));
    else
      AKM_TRY(cyc.start(tol, inner_res));
    int inner = 0;
    AKM_TRY(cyc.run_steps(restart, op, nullptr, tol, tol, inner_res, inner, speculate, pending));
    total += inner;
    AKM_TRY(cyc.finish(inner, x, false));
    ++outer;
    outer_res = inner_res;
  }
  AKM_HIP(hipStreamSynchronize(cyc.s));
  if (err)
    *err = static_cast<double>(outer_res);
  if (iterations)
    *iterations = total;
  return (outer_res > tol) ? AKM_ERR_NOT_CONVERGED : AKM_SUCCESS;
}

//! The solve of akm_gmres_euler on device vectors in the memory `mem`.
//! Without events in `mem` the host waits for the whole stream after every
//! step, and no step is enqueued ahead.
template<typename P>
static int gmres_typed(akm_matrix_s &m, gm_memory const &mem, P dt, P *x, const P *b,
                       int restart, int max_outer, P tol, double *err, int *iterations)
{
  const int64_t n = m.output_size();
  hipStream_t s = m.stream;
  gm_cycle<P> cyc(mem, n, s, restart, nullptr, 0);
  P *V = cyc.V;
  // the global operator's solve is left-preconditioned by its Jacobi
  // diagonal, fused into the apply (src/solver.cpp:94-118)
  const bool jacobi = (m.mode == mode_global);
  auto apply_op = [&](P alpha, const P *v, P beta, P *w, P gamma, const P *g) -> int {
    if (jacobi)
      return global_apply(m, alpha, v, beta, w, gamma, g, true, static_cast<double>(dt));
    return apply_device(m, alpha, v, beta, w, gamma, g);
  };
  const std::function<int(const P *, P *)> op = [&](const P *v, P *w) -> int {
    return apply_op(-dt, v, P{0}, w, P{1}, v);
  };
  // r = b - (I - dt A) x into V_0; the first one is timed (when events are
  // at hand) to decide whether steps are enqueued ahead
  auto residual = [&](int outer, bool &speculate) -> int {
    AKM_HIP(hipMemcpyAsync(V, b, sizeof(P) * n, hipMemcpyDeviceToDevice, s));
    if (outer == 0 && mem.ready != nullptr)
    {
      // the first solve on the handle times its first apply (a host wait);
      // later solves reuse that time
      if (m.gmres_apply_ms < 0)
      {
        double ms = 0;
        AKM_TRY(timed_on_stream(
            s, [&]() { return apply_op(-dt * P{-1}, x, P{1}, V, P{-1}, x); }, ms));
        m.gmres_apply_ms = ms;
        speculate = ms < speculate_below_ms;
        return AKM_SUCCESS;
      }
      speculate = m.gmres_apply_ms < speculate_below_ms;
    }
    return apply_op(-dt * P{-1}, x, P{1}, V, P{-1}, x);
  };
  return gmres_outer<P>(cyc, residual, op, x, restart, max_outer, tol, err, iterations);
}

//! The mixed-precision / flexible solver (FP64 matrix m64).
static int gmres_mixed(akm_matrix_s &m64, akm_matrix_s *m32, akm_precond_s *pre, double dt,
                       double *x, const double *b, int restart, int max_cycles, double tol_rel,
                       double inner_rel, akm_solve_stats &stats)
{
  const int64_t n = m64.output_size();
  hipStream_t s = m64.stream;
  auto &ws = *m64.gmres;
  double *Z = (pre != nullptr) ? ws.Z.data() : nullptr;
  gm_cycle<double> cyc(memory_of(ws), n, s, restart, Z, 1);
  double *V = cyc.V;
  const int sblocks = cyc.sblocks;

  std::function<int(const double *, double *)> op64 = [&](const double *v, double *w) -> int {
    ++stats.applies_f64;
    return apply_device(m64, -dt, v, 0.0, w, 1.0, v);
  };
  std::function<int(const double *, double *)> op32 = [&](const double *v, double *w) -> int {
    ++stats.applies_f32;
    gm_to_f32<<<sblocks, gm_threads, 0, s>>>(n, v, ws.v32.data());
    AKM_TRY(check_launch("gmres fp32 cast"));
    AKM_TRY(apply_device(*m32, 1.0, ws.v32.data(), 0.0, ws.y32.data(), 0.0, nullptr));
    gm_axpy_f32<<<sblocks, gm_threads, 0, s>>>(n, v, -dt, ws.y32.data(), w);
    return check_launch("gmres fp32 axpy");
  };
  std::function<int(const double *, double *)> precond = [&](const double *v,
                                                             double *z) -> int {
    ++stats.precond_applies;
    return pre->apply(v, z, s);
  };
  // r = b - (I - dt A) x in FP64 into V_0, normalised; beta = |r|
  auto true_residual = [&](double &beta) -> int {
    AKM_HIP(hipMemcpyAsync(V, b, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    ++stats.applies_f64;
    AKM_TRY(apply_device(m64, dt, x, 1.0, V, -1.0, x));
    return cyc.start(0.0, beta);
  };

  double bnorm = 0;
  AKM_TRY(cyc.norm(b, bnorm));
  stats.rhs_norm = bnorm;
  if (bnorm == 0.0)
  {
    // the solution of (I - dt A) x = 0 is x = 0
    AKM_HIP(hipMemsetAsync(x, 0, sizeof(double) * n, s));
    AKM_HIP(hipStreamSynchronize(s));
    stats.converged = 1;
    return AKM_SUCCESS;
  }
  const double tol_abs = tol_rel * bnorm;
  const auto &inner_op = (m32 != nullptr) ? op32 : op64;
  // with an FP32 operator a cycle cannot resolve the FP64 residual below
  // ~FP32 epsilon, so it stops at inner_rel of its starting residual
  const double cycle_rel = (m32 != nullptr) ? inner_rel : 0.0;

  double beta = 0;
  double first_ms = 0;
  if (m32 != nullptr)
  {
    // the first cycle only needs its residual to the accuracy of its own
    // target (inner_rel of it): b - x in FP64 and dt*A*x with the FP32 operator
    AKM_TRY(timed_on_stream(s, [&]() { return op32(x, V); }, first_ms));
    gm_rsub<double><<<sblocks, gm_threads, 0, s>>>(n, b, V);
    AKM_TRY(check_launch("gmres residual"));
    AKM_TRY(cyc.start(0.0, beta));
  }
  else
  {
    AKM_HIP(hipMemcpyAsync(V, b, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    ++stats.applies_f64;
    AKM_TRY(timed_on_stream(s, [&]() { return apply_device(m64, dt, x, 1.0, V, -1.0, x); },
                            first_ms));
    AKM_TRY(cyc.start(0.0, beta));
  }
  // speculative steps pay off when a step's operator is cheap next to the
  // host round trip; the preconditioner (if any) runs in every step too
  const bool speculate = (pre == nullptr) && first_ms < speculate_below_ms;
  bool fallback = false;
  while (beta > tol_abs && stats.cycles < max_cycles)
  {
    const double stop = std::max(0.5 * tol_abs, cycle_rel * beta);
    if (!fallback)
      AKM_HIP(hipMemcpyAsync(ws.xprev.data(), x, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    const auto &op = fallback ? op64 : inner_op;
    const auto *pc = (fallback || pre == nullptr) ? nullptr : &precond;
    double est = beta;
    int inner = 0;
    AKM_TRY(cyc.run_steps(restart, op, pc, stop, stop, est, inner, speculate && !fallback));
    AKM_TRY(cyc.finish(inner, x, pc != nullptr));
    stats.iterations += inner;
    if (fallback)
      stats.fallback_iterations += inner;
    ++stats.cycles;
    double beta_new = 0;
    AKM_TRY(true_residual(beta_new));
    if (!fallback && !(beta_new <= 0.5 * beta))
    {
      // the cycle did not halve the FP64 residual: undo it, go plain FP64
      AKM_HIP(hipMemcpyAsync(x, ws.xprev.data(), sizeof(double) * n, hipMemcpyDeviceToDevice, s));
      fallback = true;
      stats.fallback = 1;
      AKM_TRY(true_residual(beta));
      continue;
    }
    beta = beta_new;
  }
  AKM_HIP(hipStreamSynchronize(s));
  stats.abs_residual = beta;
  stats.rel_residual = beta / bnorm;
  stats.converged = (beta <= tol_abs) ? 1 : 0;
  return stats.converged ? AKM_SUCCESS : AKM_ERR_NOT_CONVERGED;
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

//! Sizes the workspace for n unknowns, `restart` and elements of `esize`
//! bytes (no work when it already has that shape).
static int size_workspace(akm_gmres_workspace &ws, int64_t n, int restart, size_t esize)
{
  if (ws.restart == restart && ws.n == n && ws.esize == esize)
    return AKM_SUCCESS;
  ws.restart = restart;
  ws.n = n;
  ws.esize = esize;
  AKM_TRY(ws.V.allocate(n * (restart + 1) * int64_t(esize)));
  AKM_TRY(ws.work.allocate(gm_work_values(n, restart) * int64_t(esize)));
  AKM_TRY(ws.state.allocate(sizeof(gm_state<double>)));
  AKM_TRY(ws.xb.allocate(2 * n * int64_t(esize)));
  if (ws.host_state == nullptr)
    AKM_HIP(hipHostMalloc(&ws.host_state, gm_slots * sizeof(gm_state<double>),
                          hipHostMallocDefault));
  for (auto &e : ws.state_ready)
    if (e == nullptr)
      AKM_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  // buffers of the flexible / mixed solver are sized on first use
  ws.Z.reset();
  ws.xprev.reset();
  ws.v32.reset();
  ws.y32.reset();
  return AKM_SUCCESS;
}

static int ensure_workspace(akm_matrix_s &m, int restart)
{
  if (!m.gmres)
    m.gmres = std::unique_ptr<akm_gmres_workspace, void (*)(akm_gmres_workspace *)>(
        new akm_gmres_workspace(), delete_workspace);
  return size_workspace(*m.gmres, m.output_size(), restart, m.esize);
}

//! LU factorisation with partial pivoting of the n x n column-major a, in
//! place (getrf): piv[k] is the row swapped with row k at step k.  Returns
//! false for an exactly singular matrix.
template<typename P>
static bool lu_factor(int n, std::vector<P> &a, std::vector<int> &piv)
{
  piv.assign(n, 0);
  for (int k = 0; k < n; k++)
  {
    int p = k;
    for (int i = k + 1; i < n; i++)
      if (std::abs(a[int64_t(k) * n + i]) > std::abs(a[int64_t(k) * n + p]))
        p = i;
    piv[k] = p;
    if (a[int64_t(k) * n + p] == P{0})
      return false;
    if (p != k)
      for (int j = 0; j < n; j++)
        std::swap(a[int64_t(j) * n + k], a[int64_t(j) * n + p]);
    const P inv = P{1} / a[int64_t(k) * n + k];
    for (int i = k + 1; i < n; i++)
      a[int64_t(k) * n + i] *= inv;
    for (int j = k + 1; j < n; j++)
    {
      const P akj = a[int64_t(j) * n + k];
      if (akj != P{0})
        for (int i = k + 1; i < n; i++)
          a[int64_t(j) * n + i] -= a[int64_t(k) * n + i] * akj;
    }
  }
  return true;
}

//! simple_gmres on an explicit matrix (src/solver.cpp:54-74): host A, x, b
//! and optional M; the solve on the null stream with its own workspace.
template<typename P>
static int gmres_dense_typed(int n, const P *A_h, P *x_h, const P *b_h, const P *M_h,
                             int restart, int max_outer, P tol, double *err, int *iterations)
{
  constexpr int max_lu = 4096; // v in one workgroup's LDS
  if (M_h != nullptr && n > max_lu)
    return set_error(AKM_ERR_UNSUPPORTED, "gmres dense: preconditioner of order %d above %d", n,
                     max_lu);
  hipStream_t s = nullptr;
  dev_buffer<P> A, xb, LU;
  dev_buffer<int> piv;
  AKM_TRY(A.upload(A_h, int64_t(n) * n, s));
  AKM_TRY(xb.allocate(2 * int64_t(n)));
  P *x = xb.data();
  P *b = xb.data() + n;
  AKM_HIP(hipMemcpyAsync(x, x_h, sizeof(P) * n, hipMemcpyHostToDevice, s));
  AKM_HIP(hipMemcpyAsync(b, b_h, sizeof(P) * n, hipMemcpyHostToDevice, s));
  if (M_h != nullptr)
  {
    std::vector<P> lu(M_h, M_h + int64_t(n) * n);
    std::vector<int> p;
    if (!lu_factor<P>(n, lu, p))
      return set_error(AKM_ERR_INVALID, "gmres dense: the preconditioner is singular");
    AKM_TRY(LU.upload(lu.data(), int64_t(n) * n, s));
    AKM_TRY(piv.upload(p.data(), n, s));
  }
  akm_gmres_workspace ws;
  AKM_TRY(size_workspace(ws, n, restart, sizeof(P)));
  gm_cycle<P> cyc(memory_of(ws), n, s, restart, nullptr, 0);
  const int gblocks = (n + gm_threads - 1) / gm_threads;
  auto precondition = [&](P *v) -> int {
    if (M_h == nullptr)
      return AKM_SUCCESS;
    gm_lu_solve<P><<<1, 256, sizeof(P) * n, s>>>(n, LU.data(), piv.data(), v);
    return check_launch("gmres dense preconditioner");
  };
  const std::function<int(const P *, P *)> op = [&](const P *v, P *w) -> int {
    gm_dense_gemv<P><<<gblocks, gm_threads, 0, s>>>(n, A.data(), v, P{1}, P{0}, w);
    AKM_TRY(check_launch("gmres dense gemv"));
    return precondition(w);
  };
  auto residual = [&](int, bool &) -> int {
    AKM_HIP(hipMemcpyAsync(cyc.V, b, sizeof(P) * n, hipMemcpyDeviceToDevice, s));
    gm_dense_gemv<P><<<gblocks, gm_threads, 0, s>>>(n, A.data(), x, P{-1}, P{1}, cyc.V);
    AKM_TRY(check_launch("gmres dense gemv"));
    return precondition(cyc.V);
  };
  const int rc = gmres_outer<P>(cyc, residual, op, x, restart, max_outer, tol, err, iterations);
  if (rc < 0)
    return rc;
  AKM_HIP(hipMemcpy(x_h, x, sizeof(P) * n, hipMemcpyDeviceToHost));
  return rc;
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
  if (restart > gm_max_restart)
    return set_error(AKM_ERR_UNSUPPORTED, "gmres: restart %d above %d", restart, gm_max_restart);
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
  const gm_memory mem = memory_of(ws);
  if (mat->prec == AKM_F64)
    rc = gmres_typed<double>(*mat, mem, dt, static_cast<double *>(xd),
                             static_cast<const double *>(bd), restart, max_iter, tol, err,
                             iterations);
  else
    rc = gmres_typed<float>(*mat, mem, static_cast<float>(dt), static_cast<float *>(xd),
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

// ------------------------------------------------- allocation-free solve ---

namespace akm
{
//! Device memory of akm_gmres_euler_noalloc: the Krylov basis, the small
//! work array and the step state of one solve.  Every GPU the library runs on
//! holds its own copy until the process ends.  Host code never passes this
//! name to a kernel or a copy: it resolves the array's device address once
//! per GPU (noalloc_memory) and uses that.
static __device__ __align__(256) unsigned char scratch[AKM_NOALLOC_SCRATCH_BYTES];

constexpr int noalloc_max_devices = 64;

//! Per GPU: the device address of `scratch` and the host copies of the step
//! state (plain host memory: the host waits for the stream before reading).
struct noalloc_slot
{
  void *dev = nullptr;
  alignas(64) unsigned char host[gm_slots * sizeof(gm_state<double>)];
};
static noalloc_slot noalloc_slots[noalloc_max_devices];
static std::mutex noalloc_mtx;

static int64_t round_up_256(int64_t b) { return (b + 255) / 256 * 256; }

//! Scratch bytes of a solve of n unknowns with `restart` of `esize` bytes.
static int64_t noalloc_bytes(int64_t n, int restart, size_t esize)
{
  return round_up_256(n * (restart + 1) * int64_t(esize)) +
         round_up_256(gm_work_values(n, restart) * int64_t(esize)) +
         round_up_256(int64_t(sizeof(gm_state<double>)));
}

//! The solve's memory carved from the current GPU's scratch array, resolving
//! its address on the first call on that GPU.  No events: the host waits for
//! the whole stream after every step and no step is enqueued ahead.
static int noalloc_memory(int64_t n, int restart, size_t esize, gm_memory &mem)
{
  int dev = 0;
  AKM_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= noalloc_max_devices)
    return set_error(AKM_ERR_UNSUPPORTED, "gmres noalloc: device %d above %d", dev,
                     noalloc_max_devices - 1);
  noalloc_slot &slot = noalloc_slots[dev];
  {
    std::lock_guard<std::mutex> lock(noalloc_mtx);
    if (slot.dev == nullptr)
    {
      void *p = nullptr;
      AKM_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
      if (p == nullptr)
        return set_error(AKM_ERR_HIP, "gmres noalloc: no device address for the scratch array");
      slot.dev = p;
    }
  }
  unsigned char *base = static_cast<unsigned char *>(slot.dev);
  const int64_t vbytes = round_up_256(n * (restart + 1) * int64_t(esize));
  const int64_t wbytes = round_up_256(gm_work_values(n, restart) * int64_t(esize));
  mem.V = base;
  mem.work = base + vbytes;
  mem.state = base + vbytes + wbytes;
  mem.host_state = slot.host;
  mem.ready = nullptr;
  return AKM_SUCCESS;
}
} // namespace akm

extern "C" akm_status akm_gmres_euler_noalloc(akm_matrix mat, double dt, void *x, const void *b,
                                              int restart, int max_iter, double tol, double *err,
                                              int *iterations)
{
  if (mat == nullptr || x == nullptr || b == nullptr)
    return set_error(AKM_ERR_INVALID, "gmres noalloc: null argument");
  if (mat->rows != mat->cols)
    return set_error(AKM_ERR_INVALID, "gmres noalloc: the operator must be square");
  const int64_t n = mat->output_size();
  if (n <= 0)
    return set_error(AKM_ERR_INVALID, "gmres noalloc: empty system");
  const int64_t cap = AKM_NOALLOC_SCRATCH_BYTES;
  if (tol <= 0)
    tol = (mat->prec == AKM_F64) ? 1e-12 : 1e-6;
  if (restart <= 0)
  {
    // the reference default, lowered until the basis fits the scratch array
    restart = default_restarts(n, mat->esize);
    while (restart > 1 && noalloc_bytes(n, restart, mat->esize) > cap)
      --restart;
  }
  if (restart > n)
    return set_error(AKM_ERR_INVALID, "Number of inner iterations %d must be less than %lld!",
                     restart, static_cast<long long>(n));
  if (restart > gm_max_restart)
    return set_error(AKM_ERR_UNSUPPORTED, "gmres: restart %d above %d", restart, gm_max_restart);
  const int64_t need = noalloc_bytes(n, restart, mat->esize);
  if (need > cap)
    return set_error(AKM_ERR_INVALID,
                     "gmres noalloc: %lld unknowns with restart %d need %lld bytes of scratch, "
                     "above the fixed %lld",
                     static_cast<long long>(n), restart, static_cast<long long>(need),
                     static_cast<long long>(cap));
  if (max_iter <= 0)
    max_iter = static_cast<int>(std::min<int64_t>(n, 1 << 30));
  gm_memory mem;
  AKM_TRY(noalloc_memory(n, restart, mat->esize, mem));
  if (mat->prec == AKM_F64)
    return gmres_typed<double>(*mat, mem, dt, static_cast<double *>(x),
                               static_cast<const double *>(b), restart, max_iter, tol, err,
                               iterations);
  return gmres_typed<float>(*mat, mem, static_cast<float>(dt), static_cast<float *>(x),
                            static_cast<const float *>(b), restart, max_iter,
                            static_cast<float>(tol), err, iterations);
}

extern "C" akm_status akm_gmres_euler_mixed(akm_matrix mat, akm_matrix mat32, akm_precond pre,
                                            double dt, double *x, const double *b, int restart,
                                            int max_cycles, double tol, double inner_tol,
                                            akm_solve_stats *stats)
{
  if (mat == nullptr || x == nullptr || b == nullptr)
    return set_error(AKM_ERR_INVALID, "gmres: null argument");
  if (mat->prec != AKM_F64)
    return set_error(AKM_ERR_INVALID, "gmres mixed: the operator must be FP64");
  if (mat->rows != mat->cols)
    return set_error(AKM_ERR_INVALID, "gmres: the operator must be square");
  const int64_t n = mat->output_size();
  if (n <= 0)
    return set_error(AKM_ERR_INVALID, "gmres: empty system");
  if (mat32 != nullptr &&
      (mat32->prec != AKM_F32 || mat32->output_size() != n || mat32->input_size() != n))
    return set_error(AKM_ERR_INVALID,
                     "gmres mixed: the inner operator must be FP32 and of the same size");
  if (pre != nullptr && pre->size() != n)
    return set_error(AKM_ERR_INVALID, "gmres mixed: preconditioner size %lld != system size %lld",
                     static_cast<long long>(pre->size()), static_cast<long long>(n));
  if (tol <= 0)
    tol = 1e-10;
  if (inner_tol <= 0)
    inner_tol = 1e-6;
  if (restart <= 0)
    restart = default_restarts(n, mat->esize);
  if (restart > n)
    return set_error(AKM_ERR_INVALID, "Number of inner iterations %d must be less than %lld!",
                     restart, static_cast<long long>(n));
  if (restart > gm_max_restart)
    return set_error(AKM_ERR_UNSUPPORTED, "gmres: restart %d above %d", restart, gm_max_restart);
  if (max_cycles <= 0)
    max_cycles = static_cast<int>(std::min<int64_t>(n, 1 << 30));
  AKM_TRY(ensure_workspace(*mat, restart));
  auto &ws = *mat->gmres;
  if (pre != nullptr && ws.Z.size() != n * restart)
    AKM_TRY(ws.Z.allocate(n * restart));
  AKM_TRY(ws.xprev.allocate(n));
  if (mat32 != nullptr)
  {
    AKM_TRY(ws.v32.allocate(n));
    AKM_TRY(ws.y32.allocate(n));
  }
  akm_solve_stats local{};
  hipStream_t saved = nullptr;
  if (mat32 != nullptr)
  {
    saved = mat32->stream;
    mat32->stream = mat->stream; // one stream for the whole solve
  }
  int rc = gmres_mixed(*mat, mat32, pre, dt, x, b, restart, max_cycles, tol, inner_tol, local);
  if (mat32 != nullptr)
    mat32->stream = saved;
  if (stats != nullptr)
    *stats = local;
  return rc;
}

extern "C" akm_status akm_gmres_dense(int n, const void *A, void *x, const void *b,
                                      const void *M, int restart, int max_iter, double tol,
                                      akm_prec prec, double *err, int *iterations)
{
  if (A == nullptr || x == nullptr || b == nullptr)
    return set_error(AKM_ERR_INVALID, "gmres dense: null argument");
  if (n <= 0)
    return set_error(AKM_ERR_INVALID, "gmres dense: empty system");
  if (prec != AKM_F64 && prec != AKM_F32)
    return set_error(AKM_ERR_INVALID, "gmres dense: unknown precision %d", int(prec));
  const size_t esize = (prec == AKM_F64) ? sizeof(double) : sizeof(float);
  if (tol <= 0)
    tol = (prec == AKM_F64) ? 1e-12 : 1e-6;
  if (restart <= 0)
    restart = default_restarts(n, esize);
  if (restart > n)
    return set_error(AKM_ERR_INVALID, "Number of inner iterations %d must be less than %d!",
                     restart, n);
  if (restart > gm_max_restart)
    return set_error(AKM_ERR_UNSUPPORTED, "gmres: restart %d above %d", restart, gm_max_restart);
  if (max_iter <= 0)
    max_iter = n;
  if (prec == AKM_F64)
    return gmres_dense_typed<double>(n, static_cast<const double *>(A), static_cast<double *>(x),
                                     static_cast<const double *>(b),
                                     static_cast<const double *>(M), restart, max_iter, tol,
                                     err, iterations);
  return gmres_dense_typed<float>(n, static_cast<const float *>(A), static_cast<float *>(x),
                                  static_cast<const float *>(b), static_cast<const float *>(M),
                                  restart, max_iter, static_cast<float>(tol), err, iterations);
}
