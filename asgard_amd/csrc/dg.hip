// 1D DG operators generated on the GPU: the coefficient matrices of
// generate_coefficients (src/coefficients.cpp:18-245) for terms whose
// coefficient functions change in time, so that an operator update never
// leaves the device.
//
// A 1D term is the product of its partial terms (rechain_dimension), each a
// mass, div or grad operator with its coefficient g given at the
// Gauss-Legendre points of every cell and at the cell edges, on a domain
// [a, b] of 2^level cells with periodic, Dirichlet or Neumann boundaries:
//   mass: int g phi_j psi_i
//   div: -int g phi_j psi_i' + numerical flux {{g u}} + s |g|/2 [[u]] on every edge
//   grad: the negative transpose of a div with the flux direction reversed
//         and the Dirichlet and Neumann roles swapped
//   penalty: the jump part of the div flux only
// in the cell basis of orthonormal Legendre polynomials (rows = test
// functions), each multiplied by the inverse of the unit mass matrix, and the
// product rotated to the multiwavelet basis: out = F (R_1 ... R_p) F^T with F
// the caller's transform (operator_two_scale).
//
// Kernels: dg_assemble builds one partial term (one thread per row of the
// cell basis: its volume integral and its flux couplings to the two
// neighbouring cells), dg_finish applies the mass inversion (and the grad
// transpose), dg_gemm multiplies the dense N x N matrices (LDS tiles).
#include "akm_internal.hpp"

#include <algorithm>
#include <cmath>
#include <vector>

namespace akm
{
constexpr int dg_max_n = 8;
constexpr int dg_max_q = 32;
constexpr int dg_threads = 256;
constexpr int dg_tile = 16;
constexpr int64_t dg_max_size = 8192; // largest N = n 2^level (3 N x N work matrices)
constexpr double pi = 3.14159265358979323846;

//! Basis values on the reference cell [-1, 1], scaled to a cell of width h.
struct dg_basis
{
  double P[dg_max_q * dg_max_n];  // P[k*n + p]: phi_p at node k / sqrt(h)
  double dP[dg_max_q * dg_max_n]; // d/dx phi_p at node k
  double w[dg_max_q];             // weights times the Jacobian h/2
  double PL[dg_max_n], PR[dg_max_n];
  double mdiag[dg_max_n]; // diagonal of the unit mass matrix
};

//! Gauss-Legendre nodes (ascending) and weights on [-1, 1], Newton on P_q.
static void gauss_legendre(int q, std::vector<double> &x, std::vector<double> &w)
{
  x.assign(q, 0.0);
  w.assign(q, 0.0);
  for (int i = 0; i < (q + 1) / 2; i++)
  {
    double z = std::cos(pi * (i + 0.75) / (q + 0.5)), dp = 1.0;
    for (int it = 0; it < 100; it++)
    {
      double p0 = 1.0, p1 = z;
      for (int k = 2; k <= q; k++)
      {
        const double p2 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
        p0 = p1;
        p1 = p2;
      }
      dp = q * (z * p1 - p0) / (z * z - 1.0);
      const double dz = p1 / dp;
      z -= dz;
      if (std::fabs(dz) < 1e-15)
        break;
    }
    x[i] = -z;
    x[q - 1 - i] = z;
    w[i] = w[q - 1 - i] = 2.0 / ((1.0 - z * z) * dp * dp);
  }
}

//! Orthonormal Legendre values and derivatives at xi: the three-term
//! recurrence (k+1) P_{k+1} = (2k+1) x P_k - k P_{k-1} and
//! P'_{k+1} = P'_{k-1} + (2k+1) P_k, scaled by sqrt((2k+1)/2).
static void legendre_onb(double xi, int n, double *P, double *dP)
{
  double pm = 0.0, p = 1.0, dpm = 0.0, dp = 0.0;
  for (int k = 0; k < n; k++)
  {
    const double s = std::sqrt((2.0 * k + 1.0) / 2.0);
    P[k] = p * s;
    dP[k] = dp * s;
    const double pn = ((2.0 * k + 1.0) * xi * p - k * pm) / (k + 1.0);
    const double dpn = dpm + (2.0 * k + 1.0) * p;
    pm = p;
    p = pn;
    dpm = dp;
    dp = dpn;
  }
}

static dg_basis make_basis(int n, int nq, double h)
{
  dg_basis B{};
  std::vector<double> x, w;
  gauss_legendre(nq, x, w);
  double p[dg_max_n], dp[dg_max_n];
  const double rs = 1.0 / std::sqrt(h), jac = h / 2.0;
  for (int k = 0; k < nq; k++)
  {
    legendre_onb(x[k], n, p, dp);
    for (int i = 0; i < n; i++)
    {
      B.P[k * n + i] = p[i] * rs;
      B.dP[k * n + i] = dp[i] * (2.0 / h) * rs;
    }
    B.w[k] = w[k] * jac;
  }
  legendre_onb(-1.0, n, p, dp);
  for (int i = 0; i < n; i++)
    B.PL[i] = p[i] * rs;
  legendre_onb(1.0, n, p, dp);
  for (int i = 0; i < n; i++)
    B.PR[i] = p[i] * rs;
  for (int i = 0; i < n; i++)
  {
    double m = 0;
    for (int k = 0; k < nq; k++)
      m += B.P[k * n + i] * B.w[k] * B.P[k * n + i];
    B.mdiag[i] = m;
  }
  return B;
}

__global__ void __launch_bounds__(dg_threads) dg_fill_zero(int64_t count, double *v)
{
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < count;
       i += int64_t(gridDim.x) * blockDim.x)
    v[i] = 0.0;
}

//! One partial term in the cell basis, column-major N x N (C zeroed): thread
//! (cell i, row p) writes row n*i + p of the blocks (i, i-1), (i, i), (i, i+1)
//! (neighbours wrap around on a periodic domain).  kind 0 mass, 1 div,
//! 3 penalty (the jump part of the div flux only); lbc / rbc the boundary
//! conditions at a and b (AKM_DG_PERIODIC, _DIRICHLET, _NEUMANN).
__global__ void __launch_bounds__(dg_threads)
    dg_assemble(const dg_basis *B, int n, int nq, int cells, int kind, double fscale, int lbc,
                int rbc, const double *gq, const double *ge, double *C)
{
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= cells * n)
    return;
  const int i = t / n, p = t - i * n;
  const int64_t N = int64_t(cells) * n;
  double left[dg_max_n], self[dg_max_n], right[dg_max_n];
  const bool penalty = (kind == 3);
  for (int q = 0; q < n; q++)
  {
    double v = 0;
    if (!penalty)
      for (int k = 0; k < nq; k++)
      {
        const double c = gq[int64_t(i) * nq + k] * B->w[k];
        v += (kind == 0 ? B->P[k * n + p] : -B->dP[k * n + p]) * c * B->P[k * n + q];
      }
    self[q] = v;
    left[q] = right[q] = 0;
  }
  if (kind != 0)
  {
    double fl = 0.5 * ge[i], fr = 0.5 * ge[i + 1];
    const double fla = fscale * fabs(fl), fra = fscale * fabs(fr);
    if (penalty)
      fl = fr = 0;
    const bool first = (i == 0), last = (i == cells - 1);
    for (int q = 0; q < n; q++)
    {
      // left edge: an interior or periodic edge couples to the left neighbour
      if (!first || lbc == AKM_DG_PERIODIC)
      {
        left[q] += (-fl - fla) * B->PL[p] * B->PR[q];
        self[q] += (-fl + fla) * B->PL[p] * B->PL[q];
      }
      else if (lbc == AKM_DG_DIRICHLET)
      {
        if (penalty)
          self[q] += fla * B->PL[p] * B->PL[q];
      }
      else if (!penalty)
        self[q] += -2.0 * fl * B->PL[p] * B->PL[q];
      // right edge
      if (!last || rbc == AKM_DG_PERIODIC)
      {
        self[q] += (fr + fra) * B->PR[p] * B->PR[q];
        right[q] += (fr - fra) * B->PR[p] * B->PL[q];
      }
      else if (rbc == AKM_DG_DIRICHLET)
      {
        if (penalty)
          self[q] += fra * B->PR[p] * B->PR[q];
      }
      else if (!penalty)
        self[q] += 2.0 * fr * B->PR[p] * B->PR[q];
    }
  }
  const int il = (i + cells - 1) % cells, ir = (i + 1) % cells;
  const int64_t row = int64_t(i) * n + p;
  for (int q = 0; q < n; q++)
  {
    double vs = self[q], vl = left[q], vr = right[q];
    if (il == i)
      vs += vl;
    if (ir == i)
      vs += vr;
    if (il == ir && il != i)
      vl += vr;
    C[(int64_t(i) * n + q) * N + row] = vs;
    if (il != i)
      C[(int64_t(il) * n + q) * N + row] = vl;
    if (ir != i && ir != il)
      C[(int64_t(ir) * n + q) * N + row] = vr;
  }
}

//! R = M^-1 C (rows scaled by the unit mass diagonal), or for grad
//! R = M^-1 (-C^T).
__global__ void __launch_bounds__(dg_threads)
    dg_finish(const dg_basis *B, int n, int64_t N, bool grad, const double *C, double *R)
{
  for (int64_t idx = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; idx < N * N;
       idx += int64_t(gridDim.x) * blockDim.x)
  {
    const int64_t c = idx / N, r = idx - c * N;
    const double v = grad ? -C[r * N + c] : C[idx];
    R[idx] = v / B->mdiag[r % n];
  }
}

//! C = op(A) op(B), N x N column-major, op = transpose when the flag is set;
//! 16 x 16 tiles through LDS.
__global__ void __launch_bounds__(dg_tile *dg_tile)
    dg_gemm(int N, const double *A, bool ta, const double *Bm, bool tb, double *C)
{
  __shared__ double sa[dg_tile][dg_tile + 1], sb[dg_tile][dg_tile + 1];
  const int tx = threadIdx.x % dg_tile, ty = threadIdx.x / dg_tile;
  const int row = blockIdx.x * dg_tile + tx, col = blockIdx.y * dg_tile + ty;
  double acc = 0;
  for (int k0 = 0; k0 < N; k0 += dg_tile)
  {
    // sa[kk][rr] = op(A)(row0 + rr, k0 + kk), sb[cc][kk] = op(B)(k0 + kk, col0 + cc)
    {
      const int r = blockIdx.x * dg_tile + tx, k = k0 + ty;
      double v = 0;
      if (r < N && k < N)
        v = ta ? A[int64_t(r) * N + k] : A[int64_t(k) * N + r];
      sa[ty][tx] = v;
    }
    {
      const int k = k0 + tx, c = blockIdx.y * dg_tile + ty;
      double v = 0;
      if (k < N && c < N)
        v = tb ? Bm[int64_t(k) * N + c] : Bm[int64_t(c) * N + k];
      sb[ty][tx] = v;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < dg_tile; kk++)
      acc += sa[kk][tx] * sb[ty][kk];
    __syncthreads();
  }
  if (row < N && col < N)
    C[int64_t(col) * N + row] = acc;
}

static int gemm(hipStream_t s, int N, const double *A, bool ta, const double *B, bool tb,
                double *C)
{
  const dim3 grid((N + dg_tile - 1) / dg_tile, (N + dg_tile - 1) / dg_tile);
  dg_gemm<<<grid, dg_tile * dg_tile, 0, s>>>(N, A, ta, B, tb, C);
  return check_launch("dg operator product");
}
} // namespace akm

using namespace akm;

akm_status akm_dg_term(int n, int level, double a, double b, int nq,
                       const akm_dg_partial *partials, int num_partials, const double *F,
                       double *out, void *stream)
{
  if (n < 1 || n > dg_max_n || level < 0 || level > 12 || nq < 1 || nq > dg_max_q || !(b > a))
    return set_error(AKM_ERR_INVALID,
                     "dg term: need 1 <= n <= %d, 0 <= level <= 12, 1 <= nq <= %d, b > a",
                     dg_max_n, dg_max_q);
  if (partials == nullptr || num_partials < 1 || F == nullptr || out == nullptr)
    return set_error(AKM_ERR_INVALID, "dg term: null argument or no partial terms");
  for (int i = 0; i < num_partials; i++)
  {
    const akm_dg_partial &pt = partials[i];
    const bool bc_ok = pt.left >= AKM_DG_PERIODIC && pt.left <= AKM_DG_NEUMANN &&
                       pt.right >= AKM_DG_PERIODIC && pt.right <= AKM_DG_NEUMANN;
    if (pt.kind < AKM_DG_MASS || pt.kind > AKM_DG_PENALTY || !bc_ok ||
        (pt.kind != AKM_DG_PENALTY && pt.gq == nullptr) ||
        (pt.kind != AKM_DG_MASS && pt.ge == nullptr))
      return set_error(AKM_ERR_INVALID, "dg term: partial term %d is malformed", i);
  }
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int cells = 1 << level;
  const int64_t N = int64_t(cells) * n, NN = N * N;
  if (N > dg_max_size)
    return set_error(AKM_ERR_INVALID, "dg term: operator size %lld above %lld",
                     static_cast<long long>(N), static_cast<long long>(dg_max_size));
  const dg_basis host = make_basis(n, nq, (b - a) / cells);
  dev_buffer<dg_basis> basis;
  dev_buffer<double> work;
  AKM_TRY(basis.upload(&host, 1, s));
  AKM_TRY(work.allocate(3 * NN));
  double *C = work.data(), *R = C + NN, *P = R + NN;
  const int blocks = static_cast<int>(std::min<int64_t>((NN + dg_threads - 1) / dg_threads, 4096));
  for (int i = 0; i < num_partials; i++)
  {
    const akm_dg_partial &pt = partials[i];
    dg_fill_zero<<<blocks, dg_threads, 0, s>>>(NN, C);
    AKM_TRY(check_launch("dg operator clear"));
    // grad: a div with the flux direction reversed and Dirichlet and Neumann
    // swapped, transposed and negated
    const bool grad = (pt.kind == AKM_DG_GRAD);
    const double fs = grad ? -pt.flux_scale : pt.flux_scale;
    auto swap_bc = [grad](int bc) {
      if (!grad || bc == AKM_DG_PERIODIC)
        return bc;
      return bc == AKM_DG_DIRICHLET ? AKM_DG_NEUMANN : AKM_DG_DIRICHLET;
    };
    const int kind = (pt.kind == AKM_DG_MASS) ? 0 : (pt.kind == AKM_DG_PENALTY ? 3 : 1);
    dg_assemble<<<(cells * n + dg_threads - 1) / dg_threads, dg_threads, 0, s>>>(
        basis.data(), n, nq, cells, kind, fs, swap_bc(pt.left), swap_bc(pt.right), pt.gq, pt.ge,
        C);
    AKM_TRY(check_launch("dg operator assembly"));
    // the first partial term goes straight to P, later ones are chained on
    double *dst = (i == 0) ? P : R;
    dg_finish<<<blocks, dg_threads, 0, s>>>(basis.data(), n, N, pt.kind == AKM_DG_GRAD, C, dst);
    AKM_TRY(check_launch("dg mass inversion"));
    if (i > 0)
    {
      AKM_TRY(gemm(s, int(N), P, false, R, false, C));
      AKM_HIP(hipMemcpyAsync(P, C, NN * sizeof(double), hipMemcpyDeviceToDevice, s));
    }
  }
  // out = F P F^T
  AKM_TRY(gemm(s, int(N), F, false, P, false, C));
  AKM_TRY(gemm(s, int(N), C, false, F, true, out));
  AKM_HIP(hipStreamSynchronize(s));
  return AKM_SUCCESS;
}
