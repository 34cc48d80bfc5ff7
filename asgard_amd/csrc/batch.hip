// Batched implicit-step solves: many independent systems
// (I - dt_s A_s) x_s = b_s, of any sizes and (d, n), iterate restarted GMRES
// in lockstep on the GPU.
//
// * The operators are applied by a persistent kernel per (d, n) group that
//   walks a (system, row) work list sorted longest row first across all the
//   group's systems, so small and large grids share one launch.
// * Every vector kernel runs over a global list of row blocks covering all
//   systems; per-system scalars (Hessenberg column, rotations, right-hand
//   side, norms) live in device arrays, and a system that converges or runs
//   out of its cycle is masked out by a device flag set in the Givens kernel.
// * A restart cycle is three stream captures replayed as hipGraphs:
//   the cycle start (b - (I - dt A) x for every unconverged system, its norm
//   and the convergence test), one Arnoldi iteration (replayed up to
//   `restart` times; the current column index is a device counter advanced by
//   the graph itself), and the update x += V H^-1 g.  The host reads the
//   per-system flags after every replay to stop early.
// * No memset is recorded in the graphs; all initialisation is by kernels.
//
// The arithmetic per system is the reference GMRES (src/solver.cpp:156-298)
// with classical Gram-Schmidt and Givens rotations, the residual estimate
// tested against tol*|b_s|, and the true residual recomputed at every cycle
// start (a system is converged when that true residual is below tolerance).
#include "matrix.hpp"

#include <algorithm>
#include <map>
#include <memory>
#include <numeric>
#include <vector>

namespace akm
{
//! Arnoldi-iteration graph replays between two host reads of the batch state.
constexpr int bt_chunk = 8;

constexpr int bt_threads = 256;
constexpr int bt_cols = 8;
constexpr int64_t bt_rows_per_block = 2048;

//! Static per-system data of the vector kernels.
struct bt_sys
{
  double *V;       // Krylov basis, (restart + 1) columns of n
  double *X;       // solution
  const double *B; // right-hand side
  int64_t n;
  int nblk; // row blocks of this system
  int blk0; // first global row block
  double *h, *H, *cs, *sn, *g, *dotp, *sqp;
};

//! Mutable per-system state.
struct bt_state
{
  double tol; // absolute tolerance tol_rel * |b|
  double nrm; // last norm
  double inv; // 1 / nrm (0 when nrm == 0)
  double res; // residual estimate / true residual at cycle start
  int converged;
  int m;     // Krylov columns of the current cycle
  int iters; // operator applications
  int pad;
};

__device__ __forceinline__ double bt_block_sum(double v, double *sh)
{
  const int tid = threadIdx.x;
  sh[tid] = v;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1)
  {
    if (tid < s)
      sh[tid] += sh[tid + s];
    __syncthreads();
  }
  const double r = sh[0];
  __syncthreads();
  return r;
}

__device__ __forceinline__ void bt_rows(bt_sys const &q, int lb, int64_t &r0, int64_t &r1)
{
  r0 = int64_t(lb) * bt_rows_per_block;
  r1 = min(q.n, r0 + bt_rows_per_block);
}

//! Start of a cycle: active = !converged, m = 0; column counter = 0.
__global__ void bt_begin(int nsys, bt_state *st, int *active, int *jcol)
{
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s == 0)
    *jcol = 0;
  if (s < nsys)
  {
    active[s] = st[s].converged ? 0 : 1;
    st[s].m = 0;
  }
}

//! V_s[0] = b_s for the active systems.
__global__ void __launch_bounds__(bt_threads)
bt_copy_rhs(const int *blk_sys, const bt_sys *sys, const int *active)
{
  const int s = blk_sys[blockIdx.x];
  if (!active[s])
    return;
  const bt_sys q = sys[s];
  int64_t r0, r1;
  bt_rows(q, blockIdx.x - q.blk0, r0, r1);
  for (int64_t i = r0 + threadIdx.x; i < r1; i += bt_threads)
    q.V[i] = q.B[i];
}

//! sq partial of |V_s[col]|^2 (col = 0) or |B_s|^2 (use_rhs).
__global__ void __launch_bounds__(bt_threads)
bt_sumsq(const int *blk_sys, const bt_sys *sys, const int *active, int use_rhs)
{
  __shared__ double sh[bt_threads];
  const int s = blk_sys[blockIdx.x];
  if (!use_rhs && !active[s])
    return;
  const bt_sys q = sys[s];
  const int lb = blockIdx.x - q.blk0;
  int64_t r0, r1;
  bt_rows(q, lb, r0, r1);
  const double *v = use_rhs ? q.B : q.V;
  double sq = 0;
  for (int64_t i = r0 + threadIdx.x; i < r1; i += bt_threads)
    sq = fma(v[i], v[i], sq);
  sq = bt_block_sum(sq, sh);
  if (threadIdx.x == 0)
    q.sqp[lb] = sq;
}

//! Per system: tol = tol_rel * |b|, converged = 0, iters = 0.
__global__ void __launch_bounds__(64) bt_set_tol(const bt_sys *sys, bt_state *st, double tol_rel)
{
  __shared__ double sh[64];
  const int s = blockIdx.x;
  const bt_sys q = sys[s];
  double v = 0;
  for (int b = threadIdx.x; b < q.nblk; b += 64)
    v += q.sqp[b];
  v = bt_block_sum(v, sh);
  if (threadIdx.x == 0)
  {
    st[s].tol = tol_rel * sqrt(v);
    st[s].converged = 0;
    st[s].iters = 0;
    st[s].m = 0;
  }
}

__device__ __forceinline__ void bt_rot(double &x, double &y, double c, double s)
{
  const double t = c * x + s * y;
  y = c * y - s * x;
  x = t;
}

//! reference BLAS rotg semantics
__device__ __forceinline__ void bt_rotg(double &a, double b, double &c, double &s)
{
  const double roe = (fabs(a) > fabs(b)) ? a : b;
  const double scale = fabs(a) + fabs(b);
  if (scale == 0.0)
  {
    c = 1;
    s = 0;
    a = 0;
    return;
  }
  const double as = a / scale, bs = b / scale;
  double r = scale * sqrt(as * as + bs * bs);
  r = (roe >= 0.0) ? r : -r;
  c = a / r;
  s = b / r;
  a = r;
}

//! One wavefront per system: finishes |w| from the block partials.  At the
//! cycle start (start != 0): g_0 = |r|, the convergence test on the true
//! residual.  Otherwise: Hessenberg column j = *jcol, old rotations, the new
//! one, the estimate; m = j + 1; a system whose estimate is below tolerance
//! leaves the cycle.
__global__ void __launch_bounds__(64)
bt_givens(const bt_sys *sys, bt_state *st, int *active, const int *jcol, int start)
{
  __shared__ double sh[64];
  const int s = blockIdx.x;
  if (!active[s])
    return;
  const bt_sys q = sys[s];
  double v = 0;
  for (int b = threadIdx.x; b < q.nblk; b += 64)
    v += q.sqp[b];
  v = bt_block_sum(v, sh);
  if (threadIdx.x != 0)
    return;
  bt_state &t = st[s];
  const double nrm = sqrt(v);
  t.nrm = nrm;
  t.inv = (nrm > 0.0) ? 1.0 / nrm : 0.0;
  t.iters += 1;
  if (start)
  {
    q.g[0] = nrm;
    t.res = nrm;
    if (nrm <= t.tol)
    {
      t.converged = 1;
      active[s] = 0;
    }
    return;
  }
  const int j = *jcol;
  double *col = q.H + (j * (j + 1)) / 2;
  for (int k = 0; k <= j; k++)
    col[k] = q.h[k];
  for (int k = 0; k < j; k++)
    bt_rot(col[k], col[k + 1], q.cs[k], q.sn[k]);
  double c, sv;
  bt_rotg(col[j], nrm, c, sv);
  q.cs[j] = c;
  q.sn[j] = sv;
  q.g[j + 1] = 0;
  bt_rot(q.g[j], q.g[j + 1], c, sv);
  t.res = fabs(q.g[j + 1]);
  t.m = j + 1;
  if (t.res <= t.tol)
    active[s] = 0;
}

//! V_s[col] *= inv_s, col = 0 (start) or *jcol + 1.
__global__ void __launch_bounds__(bt_threads)
bt_scale(const int *blk_sys, const bt_sys *sys, const bt_state *st, const int *active,
         const int *jcol, int start)
{
  const int s = blk_sys[blockIdx.x];
  if (!active[s])
    return;
  const bt_sys q = sys[s];
  const double inv = st[s].inv;
  if (inv == 0.0)
    return;
  int64_t r0, r1;
  bt_rows(q, blockIdx.x - q.blk0, r0, r1);
  double *w = q.V + (start ? 0 : int64_t(*jcol + 1) * q.n);
  for (int64_t i = r0 + threadIdx.x; i < r1; i += bt_threads)
    w[i] *= inv;
}

//! partial dots of V_s[k], k <= j, with w = V_s[j+1].
__global__ void __launch_bounds__(bt_threads)
bt_dots(const int *blk_sys, const bt_sys *sys, const int *active, const int *jcol)
{
  __shared__ double sh[bt_threads];
  const int s = blk_sys[blockIdx.x];
  if (!active[s])
    return;
  const int ncols = *jcol + 1;
  const int k0 = blockIdx.y * bt_cols;
  if (k0 >= ncols)
    return;
  const bt_sys q = sys[s];
  const int lb = blockIdx.x - q.blk0;
  int64_t r0, r1;
  bt_rows(q, lb, r0, r1);
  const double *w = q.V + int64_t(ncols) * q.n;
  double acc[bt_cols];
#pragma unroll
  for (int c = 0; c < bt_cols; c++)
    acc[c] = 0;
  for (int64_t i = r0 + threadIdx.x; i < r1; i += bt_threads)
  {
    const double wi = w[i];
#pragma unroll
    for (int c = 0; c < bt_cols; c++)
      if (k0 + c < ncols)
        acc[c] = fma(q.V[int64_t(k0 + c) * q.n + i], wi, acc[c]);
  }
#pragma unroll
  for (int c = 0; c < bt_cols; c++)
  {
    const double v = bt_block_sum(acc[c], sh);
    if (threadIdx.x == 0 && k0 + c < ncols)
      q.dotp[int64_t(k0 + c) * q.nblk + lb] = v;
  }
}

//! h_s[k] = sum of the partials (block (s, k)).
__global__ void __launch_bounds__(64)
bt_reduce(const bt_sys *sys, const int *active, const int *jcol)
{
  __shared__ double sh[64];
  const int s = blockIdx.x, k = blockIdx.y;
  if (!active[s] || k > *jcol)
    return;
  const bt_sys q = sys[s];
  double v = 0;
  for (int b = threadIdx.x; b < q.nblk; b += 64)
    v += q.dotp[int64_t(k) * q.nblk + b];
  v = bt_block_sum(v, sh);
  if (threadIdx.x == 0)
    q.h[k] = v;
}

//! w -= V_s[0..j] h_s and the partials of |w|^2.
__global__ void __launch_bounds__(bt_threads)
bt_update(const int *blk_sys, const bt_sys *sys, const int *active, const int *jcol)
{
  __shared__ double sh[bt_threads];
  __shared__ double hs[256];
  const int s = blk_sys[blockIdx.x];
  if (!active[s])
    return;
  const bt_sys q = sys[s];
  const int ncols = *jcol + 1;
  for (int k = threadIdx.x; k < ncols; k += bt_threads)
    hs[k] = q.h[k];
  __syncthreads();
  const int lb = blockIdx.x - q.blk0;
  int64_t r0, r1;
  bt_rows(q, lb, r0, r1);
  double *w = q.V + int64_t(ncols) * q.n;
  double sq = 0;
  for (int64_t i = r0 + threadIdx.x; i < r1; i += bt_threads)
  {
    double acc = 0;
    for (int k = 0; k < ncols; k++)
      acc = fma(q.V[int64_t(k) * q.n + i], hs[k], acc);
    const double v = w[i] - acc;
    w[i] = v;
    sq = fma(v, v, sq);
  }
  sq = bt_block_sum(sq, sh);
  if (threadIdx.x == 0)
    q.sqp[lb] = sq;
}

__global__ void bt_advance(int *jcol)
{
  if (threadIdx.x == 0 && blockIdx.x == 0)
    *jcol += 1;
}

//! y = H^-1 g (packed upper triangle) for every system with m > 0.
__global__ void bt_backsolve(int nsys, const bt_sys *sys, const bt_state *st)
{
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nsys)
    return;
  const int m = st[s].m;
  if (m <= 0)
    return;
  const bt_sys q = sys[s];
  for (int j = m - 1; j >= 0; j--)
  {
    const double hjj = q.H[j + (j * (j + 1)) / 2];
    q.g[j] = (hjj != 0.0) ? q.g[j] / hjj : 0.0;
    const double gj = q.g[j];
    for (int i = 0; i < j; i++)
      q.g[i] -= gj * q.H[i + (j * (j + 1)) / 2];
  }
}

//! x_s += V_s[0..m) y_s.
__global__ void __launch_bounds__(bt_threads)
bt_xupdate(const int *blk_sys, const bt_sys *sys, const bt_state *st)
{
  __shared__ double ys[256];
  const int s = blk_sys[blockIdx.x];
  const int m = st[s].m;
  if (m <= 0)
    return;
  const bt_sys q = sys[s];
  for (int k = threadIdx.x; k < m; k += bt_threads)
    ys[k] = q.g[k];
  __syncthreads();
  int64_t r0, r1;
  bt_rows(q, blockIdx.x - q.blk0, r0, r1);
  for (int64_t i = r0 + threadIdx.x; i < r1; i += bt_threads)
  {
    double acc = 0;
    for (int k = 0; k < m; k++)
      acc = fma(q.V[int64_t(k) * q.n + i], ys[k], acc);
    q.X[i] += acc;
  }
}

//! dst = value on n entries (kernel fill: no memset nodes in the graphs).
__global__ void __launch_bounds__(bt_threads) bt_fill(int64_t n, double *dst, double value)
{
  for (int64_t i = blockIdx.x * int64_t(bt_threads) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * bt_threads)
    dst[i] = value;
}

} // namespace akm

using namespace akm;

struct akm_batch_s
{
  struct group
  {
    int d = 0, n = 0, waves = 4, grid = 1;
    dev_buffer<int2> work;
    int nwork = 0;
  };
  int nsys = 0, restart = 0;
  int64_t total_n = 0;
  int total_blocks = 0;
  std::vector<akm_matrix> mats;
  std::vector<int64_t> n_of, x_off;
  std::vector<group> groups;
  dev_buffer<kron_args<double>> args;
  dev_buffer<batch_system> vsys;
  dev_buffer<bt_sys> sys;
  dev_buffer<bt_state> state;
  dev_buffer<int> active, jcol, blk_sys;
  dev_buffer<double> V, X, B, small;
  std::vector<bt_state> host_state;
  std::vector<int> host_active;
  hipStream_t stream = nullptr; // own stream (the null stream cannot be captured)
  hipGraphExec_t g_start = nullptr, g_iter = nullptr, g_finish = nullptr;
  hipGraph_t graphs[3] = {nullptr, nullptr, nullptr};

  ~akm_batch_s()
  {
    for (auto e : {g_start, g_iter, g_finish})
      if (e != nullptr)
        (void)hipGraphExecDestroy(e);
    for (auto g : graphs)
      if (g != nullptr)
        (void)hipGraphDestroy(g);
    if (stream != nullptr)
      (void)hipStreamDestroy(stream);
  }

  int apply_groups(int phase)
  {
    for (auto &gr : groups)
      AKM_TRY(launch_rows_batch(gr.d, gr.n, gr.waves, args.data(), vsys.data(), gr.work.data(),
                                gr.nwork, gr.grid, jcol.data(), active.data(), phase, stream));
    return AKM_SUCCESS;
  }

  int enqueue_start()
  {
    bt_begin<<<(nsys + 255) / 256, 256, 0, stream>>>(nsys, state.data(), active.data(),
                                                   jcol.data());
    bt_copy_rhs<<<total_blocks, bt_threads, 0, stream>>>(blk_sys.data(), sys.data(),
                                                         active.data());
    AKM_TRY(check_launch("batch start"));
    AKM_TRY(apply_groups(0));
    bt_sumsq<<<total_blocks, bt_threads, 0, stream>>>(blk_sys.data(), sys.data(), active.data(),
                                                      0);
    bt_givens<<<nsys, 64, 0, stream>>>(sys.data(), state.data(), active.data(), jcol.data(), 1);
    bt_scale<<<total_blocks, bt_threads, 0, stream>>>(blk_sys.data(), sys.data(), state.data(),
                                                      active.data(), jcol.data(), 1);
    return check_launch("batch residual");
  }

  int enqueue_iteration()
  {
    AKM_TRY(apply_groups(1));
    dim3 dgrid(total_blocks, (restart + bt_cols) / bt_cols);
    bt_dots<<<dgrid, bt_threads, 0, stream>>>(blk_sys.data(), sys.data(), active.data(),
                                             jcol.data());
    bt_reduce<<<dim3(nsys, restart), 64, 0, stream>>>(sys.data(), active.data(), jcol.data());
    bt_update<<<total_blocks, bt_threads, 0, stream>>>(blk_sys.data(), sys.data(), active.data(),
                                                       jcol.data());
    bt_givens<<<nsys, 64, 0, stream>>>(sys.data(), state.data(), active.data(), jcol.data(), 0);
    bt_scale<<<total_blocks, bt_threads, 0, stream>>>(blk_sys.data(), sys.data(), state.data(),
                                                      active.data(), jcol.data(), 0);
    bt_advance<<<1, 64, 0, stream>>>(jcol.data());
    return check_launch("batch iteration");
  }

  int enqueue_finish()
  {
    bt_backsolve<<<(nsys + 63) / 64, 64, 0, stream>>>(nsys, sys.data(), state.data());
    bt_xupdate<<<total_blocks, bt_threads, 0, stream>>>(blk_sys.data(), sys.data(),
                                                        state.data());
    return check_launch("batch update");
  }

  int capture(int which, hipGraphExec_t &exec)
  {
    AKM_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
    int rc = (which == 0) ? enqueue_start() : (which == 1) ? enqueue_iteration() : enqueue_finish();
    hipGraph_t graph = nullptr;
    hipError_t e = hipStreamEndCapture(stream, &graph);
    if (rc != AKM_SUCCESS)
    {
      if (graph != nullptr)
        (void)hipGraphDestroy(graph);
      return rc;
    }
    if (e != hipSuccess)
      return set_error(AKM_ERR_HIP, "batch: graph capture failed: %s", hipGetErrorString(e));
    graphs[which] = graph;
    AKM_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
    return AKM_SUCCESS;
  }

  int read_state()
  {
    AKM_HIP(hipMemcpyAsync(host_state.data(), state.data(), sizeof(bt_state) * nsys,
                           hipMemcpyDeviceToHost, stream));
    AKM_HIP(hipMemcpyAsync(host_active.data(), active.data(), sizeof(int) * nsys,
                           hipMemcpyDeviceToHost, stream));
    AKM_HIP(hipStreamSynchronize(stream));
    return AKM_SUCCESS;
  }
};

static int time_group(akm_batch_s &b, akm_batch_s::group &gr, int waves, double &ms)
{
  hipEvent_t e0, e1;
  AKM_HIP(hipEventCreate(&e0));
  AKM_HIP(hipEventCreate(&e1));
  auto run = [&]() {
    return launch_rows_batch(gr.d, gr.n, waves, b.args.data(), b.vsys.data(), gr.work.data(),
                             gr.nwork, gr.grid, b.jcol.data(), b.active.data(), 0, b.stream);
  };
  int rc = run();
  if (rc == AKM_SUCCESS)
  {
    (void)hipEventRecord(e0, b.stream);
    for (int r = 0; r < 3 && rc == AKM_SUCCESS; r++)
      rc = run();
    (void)hipEventRecord(e1, b.stream);
    (void)hipEventSynchronize(e1);
    float f = 0;
    (void)hipEventElapsedTime(&f, e0, e1);
    ms = f / 3.0;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return rc;
}

extern "C" {

akm_status akm_batch_create(int count, const akm_matrix *mats, const double *dts, int restart,
                            akm_batch *out)
{
  if (out == nullptr || mats == nullptr || dts == nullptr || count < 1)
    return set_error(AKM_ERR_INVALID, "batch: invalid arguments");
  *out = nullptr;
  auto b = std::make_unique<akm_batch_s>();
  b->nsys = count;
  AKM_HIP(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
  for (int s = 0; s < count; s++)
  {
    akm_matrix m = mats[s];
    if (m == nullptr || m->prec != AKM_F64 || m->mode != mode_grid || m->rows != m->cols ||
        m->rows < 1)
      return set_error(AKM_ERR_INVALID,
                       "batch: system %d must be a non-empty square FP64 sparse-grid matrix", s);
    if (!(m->n >= 2 && m->n <= 4 && m->d >= 1 && m->d <= 6) || (m->d == 6 && m->n > 2) ||
        (m->d >= 4 && m->n > 3))
      return set_error(AKM_ERR_UNSUPPORTED, "batch: (d, n) = (%d, %d) of system %d", m->d, m->n,
                       s);
  }
  int64_t min_n = mats[0]->output_size();
  for (int s = 0; s < count; s++)
    min_n = std::min<int64_t>(min_n, mats[s]->output_size());
  if (restart <= 0)
    restart = 50;
  restart = static_cast<int>(std::min<int64_t>({int64_t(restart), min_n, int64_t(255)}));
  b->restart = restart;
  b->mats.assign(mats, mats + count);

  // vector and block layout
  std::vector<bt_sys> hs(count);
  std::vector<int> blk_sys;
  const int64_t small_per = [&]() {
    int64_t m = restart;
    return 4 * (m + 2) + m * (m + 1) / 2 + 2;
  }();
  int64_t vtot = 0, small_tot = 0;
  for (int s = 0; s < count; s++)
  {
    const int64_t n = mats[s]->output_size();
    const int nblk = static_cast<int>((n + bt_rows_per_block - 1) / bt_rows_per_block);
    b->n_of.push_back(n);
    b->x_off.push_back(b->total_n);
    hs[s].n = n;
    hs[s].nblk = nblk;
    hs[s].blk0 = static_cast<int>(blk_sys.size());
    for (int q = 0; q < nblk; q++)
      blk_sys.push_back(s);
    b->total_n += n;
    vtot += n * (restart + 1);
    small_tot += small_per + int64_t(restart + 1) * nblk + nblk;
  }
  b->total_blocks = static_cast<int>(blk_sys.size());
  AKM_TRY(b->V.allocate(vtot));
  AKM_TRY(b->X.allocate(b->total_n));
  AKM_TRY(b->B.allocate(b->total_n));
  AKM_TRY(b->small.allocate(small_tot));
  {
    int64_t vo = 0, so = 0;
    for (int s = 0; s < count; s++)
    {
      bt_sys &q = hs[s];
      q.V = b->V.data() + vo;
      q.X = b->X.data() + b->x_off[s];
      q.B = b->B.data() + b->x_off[s];
      double *p = b->small.data() + so;
      const int m = restart;
      q.h = p;
      q.H = q.h + (m + 1);
      q.cs = q.H + int64_t(m) * (m + 1) / 2 + 1;
      q.sn = q.cs + (m + 1);
      q.g = q.sn + (m + 1);
      q.dotp = q.g + (m + 2);
      q.sqp = q.dotp + int64_t(m + 1) * q.nblk;
      vo += q.n * (m + 1);
      so += small_per + int64_t(m + 1) * q.nblk + q.nblk;
    }
  }
  AKM_TRY(b->sys.upload(hs.data(), count));
  AKM_TRY(b->blk_sys.upload(blk_sys.data(), int64_t(blk_sys.size())));
  std::vector<kron_args<double>> ha(count);
  std::vector<batch_system> hv(count);
  for (int s = 0; s < count; s++)
  {
    ha[s] = kron_args_f64(*mats[s]);
    hv[s] = batch_system{hs[s].V, hs[s].X, hs[s].n, dts[s]};
  }
  AKM_TRY(b->args.upload(ha.data(), count));
  AKM_TRY(b->vsys.upload(hv.data(), count));
  std::vector<bt_state> st0(count);
  for (auto &t : st0)
    t = bt_state{0, 0, 0, 0, 0, 0, 0, 0};
  AKM_TRY(b->state.upload(st0.data(), count));
  std::vector<int> ones(count, 1);
  AKM_TRY(b->active.upload(ones.data(), count));
  int zero = 0;
  AKM_TRY(b->jcol.upload(&zero, 1));
  b->host_state.resize(count);
  b->host_active.resize(count);

  // (d, n) groups with a longest-row-first work list over all their systems
  std::map<std::pair<int, int>, std::vector<int>> by_dn;
  for (int s = 0; s < count; s++)
    by_dn[{mats[s]->d, mats[s]->n}].push_back(s);
  for (auto &kv : by_dn)
  {
    akm_batch_s::group gr;
    gr.d = kv.first.first;
    gr.n = kv.first.second;
    std::vector<std::pair<int64_t, int2>> items;
    for (int s : kv.second)
    {
      const auto &pn = mats[s]->pntr;
      for (int r = 0; r < mats[s]->rows; r++)
        items.push_back({pn[r + 1] - pn[r], make_int2(s, r)});
    }
    std::stable_sort(items.begin(), items.end(),
                     [](auto const &x, auto const &y) { return x.first > y.first; });
    std::vector<int2> w(items.size());
    for (size_t i = 0; i < items.size(); i++)
      w[i] = items[i].second;
    gr.nwork = static_cast<int>(w.size());
    gr.grid = std::min(gr.nwork, 4096);
    AKM_TRY(gr.work.upload(w.data(), int64_t(w.size())));
    b->groups.push_back(std::move(gr));
  }
  // autotune the workgroup shape of every group on its own work list
  {
    const int64_t nx = b->total_n;
    bt_fill<<<256, bt_threads, 0, b->stream>>>(nx, b->X.data(), 0.5);
    bt_fill<<<256, bt_threads, 0, b->stream>>>(vtot, b->V.data(), 0.0);
    AKM_TRY(check_launch("batch init"));
    for (auto &gr : b->groups)
    {
      double best = 1e300;
      int bw = 4;
      for (int w : {1, 2, 4, 8})
      {
        double ms = 0;
        AKM_TRY(time_group(*b, gr, w, ms));
        if (ms < best)
        {
          best = ms;
          bw = w;
        }
      }
      gr.waves = bw;
    }
    AKM_HIP(hipStreamSynchronize(b->stream));
  }
  AKM_TRY(b->capture(0, b->g_start));
  AKM_TRY(b->capture(1, b->g_iter));
  AKM_TRY(b->capture(2, b->g_finish));
  *out = b.release();
  return AKM_SUCCESS;
}

akm_status akm_batch_info(akm_batch b, int *restart, int *num_groups, int *group_waves)
{
  if (b == nullptr)
    return set_error(AKM_ERR_INVALID, "batch: null handle");
  if (restart)
    *restart = b->restart;
  if (num_groups)
    *num_groups = static_cast<int>(b->groups.size());
  if (group_waves)
    for (size_t i = 0; i < b->groups.size(); i++)
      group_waves[i] = b->groups[i].waves;
  return AKM_SUCCESS;
}

akm_status akm_batch_solve(akm_batch b, double *const *x, const double *const *rhs,
                           int max_cycles, double tol, double *rel_residual, int *iterations,
                           int *converged)
{
  if (b == nullptr || x == nullptr || rhs == nullptr)
    return set_error(AKM_ERR_INVALID, "batch: null argument");
  if (tol <= 0)
    tol = 1e-10;
  if (max_cycles <= 0)
    max_cycles = 1000;
  hipStream_t s = b->stream;
  // x and b may come from any stream of the caller
  AKM_HIP(hipDeviceSynchronize());
  for (int q = 0; q < b->nsys; q++)
  {
    if (x[q] == nullptr || rhs[q] == nullptr)
      return set_error(AKM_ERR_INVALID, "batch: null vector of system %d", q);
    const size_t bytes = sizeof(double) * b->n_of[q];
    AKM_HIP(hipMemcpyAsync(b->X.data() + b->x_off[q], x[q], bytes, hipMemcpyDeviceToDevice, s));
    AKM_HIP(hipMemcpyAsync(b->B.data() + b->x_off[q], rhs[q], bytes, hipMemcpyDeviceToDevice, s));
  }
  // tolerances from |b|
  bt_sumsq<<<b->total_blocks, bt_threads, 0, s>>>(b->blk_sys.data(), b->sys.data(),
                                                  b->active.data(), 1);
  bt_set_tol<<<b->nsys, 64, 0, s>>>(b->sys.data(), b->state.data(), tol);
  AKM_TRY(check_launch("batch tolerance"));
  int cycles = 0;
  bool all_done = false;
  while (true)
  {
    AKM_HIP(hipGraphLaunch(b->g_start, s));
    AKM_TRY(b->read_state());
    int open = 0;
    for (int q = 0; q < b->nsys; q++)
      open += b->host_active[q];
    if (open == 0)
    {
      all_done = true;
      break;
    }
    if (cycles >= max_cycles)
      break;
    // Arnoldi iterations in chunks of bt_chunk graph replays between host
    // reads of the per-system flags: a system that converges inside a chunk
    // has cleared its own flag, and the kernels skip it from then on
    for (int j = 0; j < b->restart;)
    {
      const int k = std::min(bt_chunk, b->restart - j);
      for (int q = 0; q < k; q++)
        AKM_HIP(hipGraphLaunch(b->g_iter, s));
      j += k;
      AKM_TRY(b->read_state());
      int still = 0;
      for (int q = 0; q < b->nsys; q++)
        still += b->host_active[q];
      if (still == 0)
        break;
    }
    AKM_HIP(hipGraphLaunch(b->g_finish, s));
    ++cycles;
  }
  for (int q = 0; q < b->nsys; q++)
    AKM_HIP(hipMemcpyAsync(x[q], b->X.data() + b->x_off[q], sizeof(double) * b->n_of[q],
                           hipMemcpyDeviceToDevice, s));
  AKM_TRY(b->read_state()); // also waits for the copies into x
  for (int q = 0; q < b->nsys; q++)
  {
    const bt_state &t = b->host_state[q];
    const double bn = (tol > 0) ? t.tol / tol : 0.0;
    if (rel_residual)
      rel_residual[q] = (bn > 0) ? t.res / bn : t.res;
    if (iterations)
      iterations[q] = t.iters;
    if (converged)
      converged[q] = t.converged;
  }
  return all_done ? AKM_SUCCESS : AKM_ERR_NOT_CONVERGED;
}

akm_status akm_batch_destroy(akm_batch b)
{
  if (b != nullptr)
    (void)hipStreamSynchronize(b->stream);
  delete b;
  return AKM_SUCCESS;
}

} // extern "C"
