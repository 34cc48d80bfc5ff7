// Global Kronecker operator: the reference's default single-node kronmult
// mode (global_kron_matrix, src/asgard_kronmult_matrix.hpp:780-989,
// src/asgard_kronmult_matrix.cpp:1008-1590), on the GPU.
//
// The operator sum_t (A_t0 x ... x A_t,d-1) restricted to the grid's DOFs is
// applied one dimension at a time.  Every term with k non-identity factors
// becomes 2^(k-1) chains of k one-dimensional products over the cell set
// completed with its missing ancestors ("padded"); in each chain the upper
// parts (diagonal included) of some factors act first, one factor acts whole,
// and the strictly lower parts of the others act last
// (src/device/asgard_kronmult.hpp:152-220).  On a set closed under taking
// ancestors in every dimension this reproduces the pair-by-pair operator
// exactly for hierarchical (volume) couplings, and for every coupling on
// standard sparse grids.
//
// GPU layout: DOF vectors keep the pair operator's layout (cell-major, n^d
// tensor per cell, dimension 0 slowest), padded cells appended after the
// active ones.  For every dimension k and cell the host lists the connected
// cells of its fibre (cells equal in all other 1D indices) with their
// connect_1d positions.  A one-dimensional product in dimension k runs in
// one of three layouts, which akm_autotune times: one lane per (chain, DOF),
// the lanes of a cell reading each connected cell's tensor together; a group
// of 8 lanes per DOF that split the list and sum by a fixed shuffle tree;
// or one wavefront per (chain, cell) whose lane groups split the list.
// Cells are taken fibre by fibre, so the cells reading the same tensors run
// together.  The factor values are the same connect_1d blocks (vA) the pair
// operator uses.  All chains advance together, one launch per stage, each
// chain in its own ping-pong slot; one more launch sums the chains in chain
// order and applies alpha/beta/gamma (and optionally the reference's Jacobi
// scaling).
#include "akm_internal.hpp"
#include "grid.hpp"
#include "matrix.hpp"

#include <algorithm>
#include <map>
#include <memory>
#include <vector>

namespace akm
{
constexpr int gk_max_dims = 6;
constexpr int gk_threads = 256;
constexpr int gk_group_max = 4; // chains sharing one first-stage input read

enum gk_fill : int
{
  gk_upper = 0, // from the diagonal on (diagonal included)
  gk_full = 1,
  gk_lower = 2, // strictly before the diagonal
};

//! Device view of the global operator.
struct gk_args
{
  int d, n, N, T;
  int64_t active_cells, cells; // active first, then padding
  const int *elem;             // cells x d 1D indices
  // per dimension: every cell's connected cells of its fibre (cells equal in
  // all other 1D indices), 1D index ascending, with their connect_1d position
  const int *nb_ptr[gk_max_dims];  // cells + 1
  const int *nb_cell[gk_max_dims];
  const int *nb_q[gk_max_dims];
  const int *nb_diag[gk_max_dims]; // position of the cell itself
  const int *order[gk_max_dims];   // cells fibre by fibre (processing order)
  const int *ch_term;  // chain -> term
  const int *ch_nst;   // chain -> number of stages
  const int *ch_dim;   // chain -> dims of its stages (gk_max_dims each)
  const int *ch_fill;  // chain -> fills of its stages
  const int *ch_slot;  // chain -> work-buffer slot (-1 for identity-only terms)
  int nchains;
  int nslots;          // chains with at least one stage
  int64_t stride;      // work-buffer slot size (padded DOFs)
  bool xcd_order;      // cell layouts: each XCD takes contiguous items
};

//! Lane g's share (connections g, g + LANES, ...) of one stage's product for
//! DOF i of a chain: (fill part of A_{t,k}) applied along dimension k to
//! `in` (padded vector, or the active-only x when `in_active_only`).  With
//! one lane per DOF the lanes of a cell read each connected cell's tensor
//! together (coalesced); with several, long connection lists are split.
template<typename P, int LANES>
__device__ __forceinline__ P gk_row_part(gk_args const &a, const P *__restrict__ vA, int t,
                                         int k, int fill, const P *__restrict__ in,
                                         bool in_active_only, int64_t i, int g)
{
  const int N = a.N, n = a.n, d = a.d;
  const int64_t cell = i / N;
  const int e = static_cast<int>(i - cell * N);
  int stride_k = 1;
  for (int j = k + 1; j < d; j++)
    stride_k *= n;
  const int s = (e / stride_k) % n;
  const int e0 = e - s * stride_k;
  const int dg = a.nb_diag[k][cell];
  const int lo = (fill == gk_upper) ? dg : a.nb_ptr[k][cell];
  const int hi = (fill == gk_lower) ? dg + 1 : a.nb_ptr[k][cell + 1];
  const int64_t nn = int64_t(n) * n;
  const int64_t tstride = int64_t(d) * nn;
  const int64_t cstride = tstride * a.T;
  const P *Ab = vA + int64_t(t) * tstride + k * nn + s;
  const int *__restrict__ ncell = a.nb_cell[k];
  const int *__restrict__ nq = a.nb_q[k];
  P acc = 0;
#pragma unroll 4
  for (int p = lo + g; p < hi; p += LANES)
  {
    const int cj = ncell[p];
    if (in_active_only && cj >= a.active_cells)
      continue;
    int jlo = 0, jhi = n;
    if (p == dg)
    {
      if (fill == gk_upper)
        jlo = s;
      else if (fill == gk_lower)
        jhi = s;
    }
    const P *A = Ab + int64_t(nq[p]) * cstride; // A[j*n] = A(s, j)
    const P *xin = in + int64_t(cj) * N + e0;
    for (int j = jlo; j < jhi; j++)
      acc = fma(A[j * n], xin[j * stride_k], acc);
  }
  return acc;
}

//! Sum over the LANES lanes of a group, the same fixed tree on every run.
template<typename P, int LANES>
__device__ __forceinline__ P gk_group_sum(P v)
{
#pragma unroll
  for (int off = LANES / 2; off > 0; off >>= 1)
    v += __shfl_xor(v, off, LANES);
  return v;
}

//! Stage `st` of every chain with more than st stages; one group of LANES
//! lanes per (chain, padded DOF), cells taken fibre by fibre so that the
//! cells reading the same connected tensors run together.
template<typename P, int LANES>
__global__ void __launch_bounds__(gk_threads)
    gk_stage(gk_args a, const P *__restrict__ vA, const int *__restrict__ list, int nlist, int st,
             const P *__restrict__ x, P *work)
{
  const int64_t dofs = a.cells * a.N;
  const int64_t groups = dofs * nlist;
  const int g = threadIdx.x % LANES;
  for (int64_t gi = (blockIdx.x * int64_t(gk_threads) + threadIdx.x) / LANES; gi < groups;
       gi += int64_t(gk_threads / LANES) * gridDim.x)
  {
    const int li = static_cast<int>(gi / dofs);
    const int64_t pos = gi - int64_t(li) * dofs;
    const int c = list[li];
    const int slot = a.ch_slot[c];
    const int k = a.ch_dim[c * gk_max_dims + st];
    const int64_t pcell = pos / a.N;
    const int64_t i = int64_t(a.order[k][pcell]) * a.N + (pos - pcell * a.N);
    const int fill = a.ch_fill[c * gk_max_dims + st];
    const P *in = (st == 0) ? x : work + (int64_t((st - 1) & 1) * a.nslots + slot) * a.stride;
    const P v = gk_group_sum<P, LANES>(
        gk_row_part<P, LANES>(a, vA, a.ch_term[c], k, fill, in, st == 0, i, g));
    if (g == 0)
      work[(int64_t(st & 1) * a.nslots + slot) * a.stride + i] = v;
  }
}

//! Cell layout: one wavefront per (chain, cell).  Lane (g, e), e < N, g <
//! 64/N, takes every (64/N)-th connection of the cell's list starting at g;
//! the lanes of one g read a connected cell's tensor together, and the list
//! itself is loaded 64 entries at a time and passed between lanes.  Two
//! connections are fetched before either is used (compile-time n = NT; NT = 0
//! reads n at run time; two keeps the kernel at 5+ wavefronts per SIMD, four
//! measured slower).  Returns this lane's partial sum for entry e (0 on
//! idle lanes).
template<typename P, int NT>
__device__ __forceinline__ P gk_cell_part(gk_args const &a, const P *__restrict__ vA, int t, int k,
                                          int fill, const P *__restrict__ in, bool in_active_only,
                                          int64_t cell, int lane)
{
  constexpr int U = 2;
  const int N = a.N, d = a.d;
  const int n = (NT > 0) ? NT : a.n;
  const int G = wave_size / N;
  const int g = lane / N;
  const int e = lane - g * N;
  int stride_k = 1;
  for (int j = k + 1; j < d; j++)
    stride_k *= n;
  const int s = (e / stride_k) % n;
  const int e0 = e - s * stride_k;
  const int dg = a.nb_diag[k][cell];
  const int lo = (fill == gk_upper) ? dg : a.nb_ptr[k][cell];
  const int hi = (fill == gk_lower) ? dg + 1 : a.nb_ptr[k][cell + 1];
  const int64_t nn = int64_t(n) * n;
  const int64_t tstride = int64_t(d) * nn;
  const int64_t cstride = tstride * a.T;
  const P *Ab = vA + int64_t(t) * tstride + k * nn + s;
  P acc = 0;
  for (int base = lo; base < hi; base += wave_size)
  {
    const int pl = base + lane;
    const int my_cj = (pl < hi) ? a.nb_cell[k][pl] : 0;
    const int my_q = (pl < hi) ? a.nb_q[k][pl] : 0;
    const int cnt = min(wave_size, hi - base);
    // the same trip count on every lane: the shuffles run with the whole
    // wavefront active (a shuffle from an inactive lane reads zero)
    const int iters = (cnt + G - 1) / G;
    for (int it = 0; it < iters; it += U)
    {
      int cjv[U], qv[U], jlov[U], jhiv[U];
#pragma unroll
      for (int u = 0; u < U; u++)
      {
        const int r = (it + u) * G + g;
        const int rr = min(r, wave_size - 1);
        const int cj = __shfl(my_cj, rr, wave_size);
        const int q = __shfl(my_q, rr, wave_size);
        const bool ok =
            (it + u < iters) && g < G && r < cnt && !(in_active_only && cj >= a.active_cells);
        const bool diag = ok && (base + r == dg);
        // idle entries read cell 0 / connection 0 (valid addresses) and add nothing
        cjv[u] = ok ? cj : 0;
        qv[u] = ok ? q : 0;
        jlov[u] = (diag && fill == gk_upper) ? s : 0;
        jhiv[u] = !ok ? 0 : ((diag && fill == gk_lower) ? s : n);
      }
      if constexpr (NT > 0)
      {
        P av[U][NT], xv[U][NT];
#pragma unroll
        for (int u = 0; u < U; u++)
#pragma unroll
          for (int j = 0; j < NT; j++)
          {
            av[u][j] = Ab[int64_t(qv[u]) * cstride + j * NT];
            xv[u][j] = in[int64_t(cjv[u]) * N + e0 + j * stride_k];
          }
#pragma unroll
        for (int u = 0; u < U; u++)
#pragma unroll
          for (int j = 0; j < NT; j++)
          {
            const P v = fma(av[u][j], xv[u][j], acc);
            acc = (j >= jlov[u] && j < jhiv[u]) ? v : acc;
          }
      }
      else
      {
        for (int u = 0; u < U; u++)
        {
          const P *A = Ab + int64_t(qv[u]) * cstride;
          const P *xin = in + int64_t(cjv[u]) * N + e0;
          for (int j = jlov[u]; j < jhiv[u]; j++)
            acc = fma(A[j * n], xin[j * stride_k], acc);
        }
      }
    }
  }
  return acc;
}

//! Orders the LDS accesses of one wavefront (no workgroup barrier needed).
__device__ __forceinline__ void gk_wave_sync()
{
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

//! Sums the wavefront's partials over g in order into lanes g = 0 (LDS
//! slice of wave_size values owned by this wavefront).
template<typename P>
__device__ __forceinline__ P gk_cell_reduce(P v, P *red, int lane, int N)
{
  if (N == wave_size)
    return v; // one lane group: nothing to add
  red[lane] = v;
  gk_wave_sync();
  P sum = 0;
  if (lane < N)
    for (int g = 0; g < wave_size / N; g++)
      sum += red[g * N + lane];
  gk_wave_sync();
  return sum;
}

//! Workgroup index with the workgroups of one XCD made consecutive: the
//! hardware deals workgroups round-robin over the 8 XCDs, so workgroups
//! b, b + 8, b + 16, ... share an L2.  With the grid a multiple of 8 this is
//! a permutation, and the waves of one XCD take contiguous items (cells of
//! the same fibres), whose neighbours then hit that XCD's L2.
__device__ __forceinline__ int64_t gk_xcd_block(bool on)
{
  constexpr int xcds = 8;
  const int per = gridDim.x / xcds;
  return (on && per * xcds == static_cast<int>(gridDim.x))
             ? int64_t(blockIdx.x % xcds) * per + blockIdx.x / xcds
             : int64_t(blockIdx.x);
}

//! Stage `st` in the cell layout.
template<typename P, int NT>
__global__ void __launch_bounds__(gk_threads)
    gk_cell_stage(gk_args a, const P *__restrict__ vA, const int *__restrict__ list, int nlist,
                  int st, const P *__restrict__ x, P *work)
{
  __shared__ P red[gk_threads];
  const int lane = threadIdx.x % wave_size;
  P *wred = red + (threadIdx.x / wave_size) * wave_size;
  const int64_t items = a.cells * nlist;
  for (int64_t w = (gk_xcd_block(a.xcd_order) * gk_threads + threadIdx.x) / wave_size; w < items;
       w += int64_t(gk_threads / wave_size) * gridDim.x)
  {
    const int li = static_cast<int>(w / a.cells);
    const int c = list[li];
    const int slot = a.ch_slot[c];
    const int k = a.ch_dim[c * gk_max_dims + st];
    const int64_t cell = a.order[k][w - int64_t(li) * a.cells];
    const P *in = (st == 0) ? x : work + (int64_t((st - 1) & 1) * a.nslots + slot) * a.stride;
    const P v = gk_cell_reduce(gk_cell_part<P, NT>(a, vA, a.ch_term[c], k,
                                               a.ch_fill[c * gk_max_dims + st], in, st == 0,
                                               cell, lane),
                               wred, lane, a.N);
    if (lane < a.N)
      work[(int64_t(st & 1) * a.nslots + slot) * a.stride + cell * a.N + lane] = v;
  }
}

//! gk_cell_part for a group of chains whose stage reads the same input
//! along the same dimension with the same fill (the first stage of chains of
//! different terms): the input values are loaded once per connection and
//! each chain's factor block is applied to them.  Every chain's sum is
//! formed in the same order as gk_cell_part forms it.
template<typename P, int NT, int GM>
__device__ __forceinline__ void gk_cell_part_group(gk_args const &a, const P *__restrict__ vA,
                                                   const int (&tg)[GM], int gm, int k, int fill,
                                                   const P *__restrict__ in, int64_t cell,
                                                   int lane, P (&acc)[GM])
{
  constexpr int U = 2;
  const int N = a.N, d = a.d;
  const int G = wave_size / N;
  const int g = lane / N;
  const int e = lane - g * N;
  int stride_k = 1;
  for (int j = k + 1; j < d; j++)
    stride_k *= NT;
  const int s = (e / stride_k) % NT;
  const int e0 = e - s * stride_k;
  const int dg = a.nb_diag[k][cell];
  const int lo = (fill == gk_upper) ? dg : a.nb_ptr[k][cell];
  const int hi = (fill == gk_lower) ? dg + 1 : a.nb_ptr[k][cell + 1];
  const int64_t nn = int64_t(NT) * NT;
  const int64_t tstride = int64_t(d) * nn;
  const int64_t cstride = tstride * a.T;
  const P *Ab[GM];
#pragma unroll
  for (int m = 0; m < GM; m++)
  {
    acc[m] = 0;
    // members past gm repeat the first chain's (valid) blocks; their sums are dropped
    Ab[m] = vA + int64_t(tg[m < gm ? m : 0]) * tstride + k * nn + s;
  }
  for (int base = lo; base < hi; base += wave_size)
  {
    const int pl = base + lane;
    const int my_cj = (pl < hi) ? a.nb_cell[k][pl] : 0;
    const int my_q = (pl < hi) ? a.nb_q[k][pl] : 0;
    const int cnt = min(wave_size, hi - base);
    const int iters = (cnt + G - 1) / G;
    for (int it = 0; it < iters; it += U)
    {
      int cjv[U], qv[U], jlov[U], jhiv[U];
#pragma unroll
      for (int u = 0; u < U; u++)
      {
        const int r = (it + u) * G + g;
        const int rr = min(r, wave_size - 1);
        const int cj = __shfl(my_cj, rr, wave_size);
        const int q = __shfl(my_q, rr, wave_size);
        const bool ok = (it + u < iters) && g < G && r < cnt && cj < a.active_cells;
        const bool diag = ok && (base + r == dg);
        cjv[u] = ok ? cj : 0;
        qv[u] = ok ? q : 0;
        jlov[u] = (diag && fill == gk_upper) ? s : 0;
        jhiv[u] = !ok ? 0 : ((diag && fill == gk_lower) ? s : NT);
      }
      P xv[U][NT];
#pragma unroll
      for (int u = 0; u < U; u++)
#pragma unroll
        for (int j = 0; j < NT; j++)
          xv[u][j] = in[int64_t(cjv[u]) * N + e0 + j * stride_k];
#pragma unroll
      for (int m = 0; m < GM; m++)
      {
        if (m >= gm) // the same on every lane
          break;
        P av[U][NT];
#pragma unroll
        for (int u = 0; u < U; u++)
#pragma unroll
          for (int j = 0; j < NT; j++)
            av[u][j] = Ab[m][int64_t(qv[u]) * cstride + j * NT];
#pragma unroll
        for (int u = 0; u < U; u++)
#pragma unroll
          for (int j = 0; j < NT; j++)
          {
            const P v = fma(av[u][j], xv[u][j], acc[m]);
            acc[m] = (j >= jlov[u] && j < jhiv[u]) ? v : acc[m];
          }
      }
    }
  }
}

//! The first stage in the cell layout with chains grouped by (dimension,
//! fill): one wavefront per (group, cell) reads the neighbours' x values once
//! for up to GM chains.  Group gi holds chains gchain[gptr[gi] .. gptr[gi+1]).
template<typename P, int NT, int GM>
__global__ void __launch_bounds__(gk_threads)
    gk_cell_first(gk_args a, const P *__restrict__ vA, const int *__restrict__ gptr,
                  const int *__restrict__ gchain, int ngroups, const P *__restrict__ x, P *work)
{
  __shared__ P red[gk_threads];
  const int lane = threadIdx.x % wave_size;
  P *wred = red + (threadIdx.x / wave_size) * wave_size;
  const int64_t items = a.cells * ngroups;
  for (int64_t w = (gk_xcd_block(a.xcd_order) * gk_threads + threadIdx.x) / wave_size; w < items;
       w += int64_t(gk_threads / wave_size) * gridDim.x)
  {
    const int gi = static_cast<int>(w / a.cells);
    const int g0 = gptr[gi], gm = gptr[gi + 1] - g0;
    const int c0 = gchain[g0];
    const int k = a.ch_dim[c0 * gk_max_dims];
    const int fill = a.ch_fill[c0 * gk_max_dims];
    const int64_t cell = a.order[k][w - int64_t(gi) * a.cells];
    int tg[GM];
#pragma unroll
    for (int m = 0; m < GM; m++)
      tg[m] = (m < gm) ? a.ch_term[gchain[g0 + m]] : 0;
    P acc[GM];
    gk_cell_part_group<P, NT, GM>(a, vA, tg, gm, k, fill, x, cell, lane, acc);
#pragma unroll
    for (int m = 0; m < GM; m++)
    {
      if (m < gm)
      {
        const P v = gk_cell_reduce(acc[m], wred, lane, a.N);
        if (lane < a.N)
          work[int64_t(a.ch_slot[gchain[g0 + m]]) * a.stride + cell * a.N + lane] = v;
      }
    }
  }
}

//! y = (alpha * sum over chains of the chain's last stage (x for a term of
//! identity factors), in chain order, + beta*y + gamma*xg) / (1 - jdt * pc)
//! (no division when pc is null).
template<typename P>
__global__ void __launch_bounds__(gk_threads)
    gk_sum(gk_args a, const P *__restrict__ x, const P *__restrict__ work, P alpha, P beta, P *y,
           P gamma, const P *xg, const double *pc, double jdt)
{
  const int64_t dofs = a.active_cells * a.N;
  for (int64_t i = blockIdx.x * int64_t(gk_threads) + threadIdx.x; i < dofs;
       i += int64_t(gk_threads) * gridDim.x)
  {
    P sum = 0;
    for (int c = 0; c < a.nchains; c++)
    {
      const int ns = a.ch_nst[c];
      sum += (ns == 0) ? x[i]
                       : work[(int64_t((ns - 1) & 1) * a.nslots + a.ch_slot[c]) * a.stride + i];
    }
    P out = alpha * sum;
    if (beta != P{0})
      out += beta * y[i];
    if (gamma != P{0})
      out += gamma * xg[i];
    if (pc != nullptr)
      out = static_cast<P>(static_cast<double>(out) / (1.0 - jdt * pc[i]));
    y[i] = out;
  }
}

//! y = beta*y + gamma*xg (no terms), with the optional Jacobi division.
template<typename P>
__global__ void __launch_bounds__(gk_threads)
    gk_scale(int64_t dofs, P beta, P *y, P gamma, const P *xg, const double *pc, double jdt)
{
  for (int64_t i = blockIdx.x * int64_t(gk_threads) + threadIdx.x; i < dofs;
       i += int64_t(gk_threads) * gridDim.x)
  {
    P out = (beta != P{0}) ? beta * y[i] : P{0};
    if (gamma != P{0})
      out += gamma * xg[i];
    if (pc != nullptr)
      out = static_cast<P>(static_cast<double>(out) / (1.0 - jdt * pc[i]));
    y[i] = out;
  }
}

static int gk_blocks(int64_t work)
{
  return static_cast<int>(std::max<int64_t>(1, std::min<int64_t>((work + gk_threads - 1) /
                                                                     gk_threads,
                                                                 65536)));
}

//! gk_blocks rounded up to a multiple of the 8 XCDs (for gk_xcd_block).
static int gk_blocks_xcd(int64_t work)
{
  const int b = gk_blocks(work);
  return (b + 7) / 8 * 8;
}
} // namespace akm

//! Host and device data of a global operator.
struct akm_global_s
{
  int64_t active_cells = 0, cells = 0;
  std::vector<int> elem_host; // padded cells x d
  akm::dev_buffer<int> elem;
  akm::dev_buffer<int> nb_ptr[akm::gk_max_dims], nb_cell[akm::gk_max_dims],
      nb_q[akm::gk_max_dims], nb_diag[akm::gk_max_dims], order[akm::gk_max_dims];
  // chains
  std::vector<int> ch_nst_host;
  akm::dev_buffer<int> ch_term, ch_nst, ch_dim, ch_fill, ch_slot;
  std::vector<akm::dev_buffer<int>> stage_list; // per stage: the chains that have it
  std::vector<int> stage_count;
  // first stage, cell layout: chains grouped by (dimension, fill), at most
  // gk_group_max per group
  akm::dev_buffer<int> first_ptr, first_chain;
  int first_groups = 0, first_largest = 0;
  bool first_grouped = false; // chosen by global_autotune
  bool xcd_order = false;     // chosen by global_autotune (cell layouts)
  int nchains = 0, nslots = 0, max_stages = 0;
  // lanes per DOF row: 1 (coalesced tensors), 8 (split lists), or 0: one
  // wavefront per cell (n^d <= 64)
  int lanes = 1;
  akm::dev_buffer<unsigned char> work;
  akm::dev_buffer<double> pc; // Jacobi diagonal (active DOFs)
  std::vector<int> c1_pntr_h, c1_indx_h, c1_diag_h;
  akm::dev_buffer<int> c1_diag; // device copy, made by the first device update
  int64_t flops = 0;
};

namespace akm
{
void destroy_global(akm_global_s *g) { delete g; }

static gk_args global_args(akm_matrix_s const &m)
{
  const akm_global_s &g = *m.gk;
  gk_args a{};
  a.d = m.d;
  a.n = m.n;
  a.N = static_cast<int>(m.N);
  a.T = m.T;
  a.active_cells = g.active_cells;
  a.cells = g.cells;
  a.elem = g.elem.data();
  for (int k = 0; k < m.d; k++)
  {
    a.nb_ptr[k] = g.nb_ptr[k].data();
    a.nb_cell[k] = g.nb_cell[k].data();
    a.nb_q[k] = g.nb_q[k].data();
    a.nb_diag[k] = g.nb_diag[k].data();
    a.order[k] = g.order[k].data();
  }
  a.ch_term = g.ch_term.data();
  a.ch_nst = g.ch_nst.data();
  a.ch_dim = g.ch_dim.data();
  a.ch_fill = g.ch_fill.data();
  a.ch_slot = g.ch_slot.data();
  a.nchains = g.nchains;
  a.nslots = g.nslots;
  a.stride = g.cells * m.N;
  a.xcd_order = g.xcd_order;
  return a;
}

template<typename P>
static int global_apply_typed(akm_matrix_s &m, P alpha, const P *x, P beta, P *y, P gamma,
                              const P *xg, bool jacobi, double jdt)
{
  akm_global_s &g = *m.gk;
  const int64_t dofs = g.active_cells * m.N;
  const double *pc = jacobi ? g.pc.data() : nullptr;
  if (g.nchains == 0 || alpha == P{0} || dofs == 0)
  {
    if (dofs > 0)
      gk_scale<P><<<gk_blocks(dofs), gk_threads, 0, m.stream>>>(dofs, beta, y, gamma, xg, pc,
                                                                  jdt);
    return check_launch("global kron scale");
  }
  const gk_args a = global_args(m);
  const P *vA = reinterpret_cast<const P *>(m.vA.data());
  P *work = reinterpret_cast<P *>(g.work.data());
  const int lanes = g.lanes;
  for (int st = 0; st < g.max_stages; st++)
  {
    const int nl = g.stage_count[st];
    if (nl == 0)
      continue;
    const int blocks = (lanes == 0) ? gk_blocks_xcd(g.cells * nl * wave_size)
                                    : gk_blocks(g.cells * m.N * nl * lanes);
    if (lanes == 0 && st == 0 && g.first_grouped && g.first_largest > 1 && m.n >= 2 &&
        m.n <= 4)
    {
      // chains sharing their first-stage input read it once
      const int fb = gk_blocks_xcd(g.cells * g.first_groups * wave_size);
      const int *gp = g.first_ptr.data(), *gc = g.first_chain.data();
      const int ng = g.first_groups;
      const bool two = g.first_largest <= 2;
      if (m.n == 2 && two)
        gk_cell_first<P, 2, 2><<<fb, gk_threads, 0, m.stream>>>(a, vA, gp, gc, ng, x, work);
      else if (m.n == 2)
        gk_cell_first<P, 2, 4><<<fb, gk_threads, 0, m.stream>>>(a, vA, gp, gc, ng, x, work);
      else if (m.n == 3 && two)
        gk_cell_first<P, 3, 2><<<fb, gk_threads, 0, m.stream>>>(a, vA, gp, gc, ng, x, work);
      else if (m.n == 3)
        gk_cell_first<P, 3, 4><<<fb, gk_threads, 0, m.stream>>>(a, vA, gp, gc, ng, x, work);
      else if (two)
        gk_cell_first<P, 4, 2><<<fb, gk_threads, 0, m.stream>>>(a, vA, gp, gc, ng, x, work);
      else
        gk_cell_first<P, 4, 4><<<fb, gk_threads, 0, m.stream>>>(a, vA, gp, gc, ng, x, work);
    }
    else if (lanes == 0)
    {
      const int *list = g.stage_list[st].data();
      switch (m.n)
      {
      case 2:
        gk_cell_stage<P, 2><<<blocks, gk_threads, 0, m.stream>>>(a, vA, list, nl, st, x, work);
        break;
      case 3:
        gk_cell_stage<P, 3><<<blocks, gk_threads, 0, m.stream>>>(a, vA, list, nl, st, x, work);
        break;
      case 4:
        gk_cell_stage<P, 4><<<blocks, gk_threads, 0, m.stream>>>(a, vA, list, nl, st, x, work);
        break;
      default:
        gk_cell_stage<P, 0><<<blocks, gk_threads, 0, m.stream>>>(a, vA, list, nl, st, x, work);
      }
    }
    else if (lanes == 8)
      gk_stage<P, 8><<<blocks, gk_threads, 0, m.stream>>>(a, vA, g.stage_list[st].data(), nl, st,
                                                           x, work);
    else
      gk_stage<P, 1><<<blocks, gk_threads, 0, m.stream>>>(a, vA, g.stage_list[st].data(), nl, st,
                                                           x, work);
    AKM_TRY(check_launch("global kron stage"));
  }
  gk_sum<P><<<gk_blocks(dofs), gk_threads, 0, m.stream>>>(a, x, work, alpha, beta, y, gamma, xg,
                                                          pc, jdt);
  return check_launch("global kron sum");
}

int global_apply(akm_matrix_s &m, double alpha, const void *x, double beta, void *y,
                 double gamma, const void *xg, bool jacobi, double jdt)
{
  if (m.prec == AKM_F64)
    return global_apply_typed<double>(m, alpha, static_cast<const double *>(x), beta,
                                      static_cast<double *>(y), gamma,
                                      static_cast<const double *>(xg), jacobi, jdt);
  return global_apply_typed<float>(m, static_cast<float>(alpha), static_cast<const float *>(x),
                                   static_cast<float>(beta), static_cast<float *>(y),
                                   static_cast<float>(gamma), static_cast<const float *>(xg),
                                   jacobi, jdt);
}

//! The Jacobi diagonal of the reference's implicit step
//! (build_preconditioner, src/asgard_kronmult_matrix.cpp:1016-1067): for
//! every active DOF, the sum over terms of the product over dimensions of the
//! 1D factors' diagonal entries (every dimension, identity factors included).
template<typename P>
static void jacobi_diagonal(akm_matrix_s const &m, const P *vA, std::vector<double> &pc)
{
  const akm_global_s &g = *m.gk;
  const int d = m.d, n = m.n, T = m.T;
  const int64_t N = m.N, nn = int64_t(n) * n;
  pc.assign(g.active_cells * N, 0.0);
  for (int64_t c = 0; c < g.active_cells; c++)
    for (int64_t e = 0; e < N; e++)
    {
      double sum = 0;
      for (int t = 0; t < T; t++)
      {
        double prod = 1;
        int64_t rest = e;
        for (int k = d - 1; k >= 0; k--)
        {
          const int s = static_cast<int>(rest % n);
          rest /= n;
          const int r1 = g.elem_host[c * d + k];
          const int64_t q = g.c1_diag_h[r1];
          prod *= static_cast<double>(vA[((q * T + t) * d + k) * nn + int64_t(s) * n + s]);
        }
        sum += prod;
      }
      pc[c * N + e] = sum;
    }
}

int global_update(akm_matrix_s &m, const void *vA)
{
  AKM_HIP(hipMemcpyAsync(m.vA.data(), vA, m.num_vA * m.esize, hipMemcpyHostToDevice, m.stream));
  std::vector<double> pc;
  if (m.prec == AKM_F64)
    jacobi_diagonal(m, static_cast<const double *>(vA), pc);
  else
    jacobi_diagonal(m, static_cast<const float *>(vA), pc);
  AKM_TRY(m.gk->pc.upload(pc.data(), int64_t(pc.size()), m.stream));
  AKM_HIP(hipStreamSynchronize(m.stream));
  return AKM_SUCCESS;
}

//! jacobi_diagonal on the device, one thread per active DOF, with the same
//! operations in the same order (no contraction into fma), so the two agree
//! bit for bit.
template<typename P>
__global__ void __launch_bounds__(gk_threads)
    gk_jacobi(int64_t active_cells, int d, int n, int T, int64_t N, const int *elem,
              const int *c1_diag, const P *vA, double *pc)
{
  const int64_t nn = int64_t(n) * n;
  for (int64_t idx = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; idx < active_cells * N;
       idx += int64_t(gridDim.x) * blockDim.x)
  {
    const int64_t c = idx / N, e = idx - c * N;
    double sum = 0;
    for (int t = 0; t < T; t++)
    {
      double prod = 1;
      int64_t rest = e;
      for (int k = d - 1; k >= 0; k--)
      {
        const int s = static_cast<int>(rest % n);
        rest /= n;
        const int64_t q = c1_diag[elem[c * d + k]];
        prod = __dmul_rn(prod, static_cast<double>(vA[((q * T + t) * d + k) * nn + s * n + s]));
      }
      sum = __dadd_rn(sum, prod);
    }
    pc[idx] = sum;
  }
}

int global_device_jacobi(akm_matrix_s &m)
{
  akm_global_s &g = *m.gk;
  if (g.c1_diag.size() == 0)
    AKM_TRY(g.c1_diag.upload(g.c1_diag_h.data(), int64_t(g.c1_diag_h.size()), m.stream));
  const int64_t work = g.active_cells * m.N;
  if (work == 0)
    return AKM_SUCCESS;
  if (m.prec == AKM_F64)
    gk_jacobi<double><<<gk_blocks(work), gk_threads, 0, m.stream>>>(
        g.active_cells, m.d, m.n, m.T, m.N, g.elem.data(), g.c1_diag.data(),
        reinterpret_cast<const double *>(m.vA.data()), g.pc.data());
  else
    gk_jacobi<float><<<gk_blocks(work), gk_threads, 0, m.stream>>>(
        g.active_cells, m.d, m.n, m.T, m.N, g.elem.data(), g.c1_diag.data(),
        reinterpret_cast<const float *>(m.vA.data()), g.pc.data());
  return check_launch("global kron Jacobi diagonal");
}

int64_t global_flops(akm_matrix_s const &m) { return m.gk ? m.gk->flops : 0; }

int global_layout(akm_matrix_s const &m)
{
  if (!m.gk)
    return 0;
  if (m.gk->lanes == 0)
    return (m.gk->first_grouped ? 3 : 2) + (m.gk->xcd_order ? 2 : 0);
  return m.gk->lanes;
}

int global_set_layout(akm_matrix_s &m, int layout)
{
  if (!m.gk)
    return set_error(AKM_ERR_INVALID, "not a global kron matrix");
  if (layout == 0)
    return AKM_SUCCESS;
  if (layout != 1 && layout != 8 && (layout < 2 || layout > 5))
    return set_error(AKM_ERR_INVALID, "global kron layout %d is not 1, 2, 3, 4, 5 or 8", layout);
  const bool cell = (layout >= 2 && layout <= 5);
  if (cell && m.N > wave_size)
    return set_error(AKM_ERR_UNSUPPORTED, "global kron: the cell layout needs n^d <= 64");
  m.gk->lanes = cell ? 0 : layout;
  m.gk->first_grouped = (layout == 3 || layout == 5);
  m.gk->xcd_order = (layout == 4 || layout == 5);
  return AKM_SUCCESS;
}

//! Times the row layouts on the operator itself (y is overwritten) and
//! keeps the fastest.
int global_autotune(akm_matrix_s &m, const void *x, void *y)
{
  akm_global_s &g = *m.gk;
  hipEvent_t e0, e1;
  AKM_HIP(hipEventCreate(&e0));
  AKM_HIP(hipEventCreate(&e1));
  float best = -1;
  int best_lanes = 1;
  bool best_grouped = false;
  int rc = AKM_SUCCESS;
  bool best_xcd = false;
  // layouts 1 and 8, then the cell layout with and without the grouped
  // first stage (shared input reads against fewer, longer wavefronts), each
  // cell layout with and without the XCD-contiguous workgroup order
  const int shapes[6][3] = {{1, 0, 0}, {8, 0, 0}, {0, 0, 0}, {0, 0, 1}, {0, 1, 0}, {0, 1, 1}};
  for (auto const &shape : shapes)
  {
    const int lanes = shape[0];
    const bool grouped = shape[1] != 0;
    const bool xcd = shape[2] != 0;
    if (lanes == 0 && m.N > wave_size)
      continue;
    if (grouped && g.first_largest < 2)
      continue;
    g.lanes = lanes;
    g.first_grouped = grouped;
    g.xcd_order = xcd;
    rc = global_apply(m, 1.0, x, 0.0, y, 0.0, nullptr, false, 0.0);
    if (rc != AKM_SUCCESS)
      break;
    (void)hipEventRecord(e0, m.stream);
    for (int r = 0; r < 5 && rc == AKM_SUCCESS; r++)
      rc = global_apply(m, 1.0, x, 0.0, y, 0.0, nullptr, false, 0.0);
    (void)hipEventRecord(e1, m.stream);
    float ms = 0;
    if (rc == AKM_SUCCESS && hipEventSynchronize(e1) == hipSuccess &&
        hipEventElapsedTime(&ms, e0, e1) == hipSuccess && (best < 0 || ms < best))
    {
      best = ms;
      best_lanes = lanes;
      best_grouped = grouped;
      best_xcd = xcd;
    }
  }
  g.lanes = best_lanes;
  g.first_grouped = best_grouped;
  g.xcd_order = best_xcd;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return rc;
}

//! Ancestors of a 1D cell in the volume hierarchy: its parent chain down to
//! cell 0 (the part of a connect_1d volume row before the diagonal,
//! src/asgard_grid_1d.hpp:159-260).
static void parents_1d(int i, std::vector<int> &out)
{
  out.clear();
  while (i > 0)
  {
    i = (i == 1) ? 0 : i / 2;
    out.push_back(i);
  }
}

//! Builds the padded cell set, the fibres, the chains and the work buffers
//! (make_global_kron_matrix, src/asgard_kronmult_matrix.cpp:1198-1334).
static int build_global(akm_matrix_s &m, const std::vector<int> &elem_active,
                        const int *active_mask, const int *flux_dim)
{
  auto g = std::unique_ptr<akm_global_s, void (*)(akm_global_s *)>(new akm_global_s(),
                                                                    destroy_global);
  const int d = m.d;
  const int64_t nact = m.rows;
  g->active_cells = nact;

  // ancestry completion (compute_ancestry_completion over the volume
  // hierarchy, src/asgard_indexset.cpp:243-318): add the missing ancestors
  // in every dimension until none is missing; padding sorted
  std::map<std::vector<int>, int64_t> index;
  std::vector<int> key(d), anc;
  for (int64_t c = 0; c < nact; c++)
  {
    key.assign(elem_active.begin() + c * d, elem_active.begin() + (c + 1) * d);
    index.emplace(key, c);
  }
  std::vector<std::vector<int>> frontier, pad;
  for (int64_t c = 0; c < nact; c++)
    frontier.emplace_back(elem_active.begin() + c * d, elem_active.begin() + (c + 1) * d);
  std::map<std::vector<int>, int> padset;
  while (!frontier.empty())
  {
    std::vector<std::vector<int>> next;
    for (auto const &cell : frontier)
      for (int k = 0; k < d; k++)
      {
        parents_1d(cell[k], anc);
        for (int a : anc)
        {
          key = cell;
          key[k] = a;
          if (index.count(key) == 0 && padset.count(key) == 0)
          {
            padset.emplace(key, 0);
            next.push_back(key);
          }
        }
      }
    frontier.swap(next);
  }
  g->elem_host = elem_active;
  for (auto const &kv : padset)
  {
    index.emplace(kv.first, nact + static_cast<int64_t>(g->elem_host.size() / d - nact));
    g->elem_host.insert(g->elem_host.end(), kv.first.begin(), kv.first.end());
  }
  g->cells = static_cast<int64_t>(g->elem_host.size() / d);
  if (g->cells * m.N > (int64_t(1) << 40))
    return set_error(AKM_ERR_UNSUPPORTED, "global kron: too many padded DOFs");
  AKM_TRY(g->elem.upload(g->elem_host.data(), int64_t(g->elem_host.size())));

  // connect_1d (full hierarchy), the pattern of every 1D factor
  AKM_TRY(build_connect_1d(m.max_level, true, g->c1_pntr_h, g->c1_indx_h, g->c1_diag_h));

  // per dimension: the fibres (cells equal in every other 1D index) and for
  // every cell its connected cells of the fibre, 1D index ascending
  std::vector<int64_t> nnz_dim(d, 0); // DOF-level pattern entries per dimension
  for (int k = 0; k < d; k++)
  {
    std::map<std::vector<int>, int> fib;
    std::vector<int> fib_of(g->cells);
    std::vector<std::map<int, int>> members; // 1D index -> cell
    std::vector<int> other(d - 1);
    for (int64_t c = 0; c < g->cells; c++)
    {
      for (int j = 0, o = 0; j < d; j++)
        if (j != k)
          other[o++] = g->elem_host[c * d + j];
      auto it = fib.find(other);
      if (it == fib.end())
      {
        it = fib.emplace(other, static_cast<int>(members.size())).first;
        members.emplace_back();
      }
      fib_of[c] = it->second;
      members[it->second].emplace(g->elem_host[c * d + k], static_cast<int>(c));
    }
    std::vector<int> ptr(g->cells + 1, 0), ncell, nq, ndiag(g->cells, -1);
    for (int64_t c = 0; c < g->cells; c++)
    {
      const int r1 = g->elem_host[c * d + k];
      const auto &mem = members[fib_of[c]];
      for (int q = g->c1_pntr_h[r1]; q < g->c1_pntr_h[r1 + 1]; q++)
      {
        auto it = mem.find(g->c1_indx_h[q]);
        if (it == mem.end())
          continue;
        if (q == g->c1_diag_h[r1])
          ndiag[c] = static_cast<int>(ncell.size());
        ncell.push_back(it->second);
        nq.push_back(q);
      }
      if (ndiag[c] < 0 || ncell[ndiag[c]] != c)
        return set_error(AKM_ERR_INVALID, "global kron: connect_1d row without its diagonal");
      if (ncell.size() > (size_t(1) << 31) - 1)
        return set_error(AKM_ERR_UNSUPPORTED, "global kron: connection lists too large");
      ptr[c + 1] = static_cast<int>(ncell.size());
      nnz_dim[k] += int64_t(ptr[c + 1] - ptr[c]) * m.n * m.N;
    }
    AKM_TRY(g->nb_ptr[k].upload(ptr.data(), int64_t(ptr.size())));
    AKM_TRY(g->nb_cell[k].upload(ncell.data(), int64_t(ncell.size())));
    AKM_TRY(g->nb_q[k].upload(nq.data(), int64_t(nq.size())));
    AKM_TRY(g->nb_diag[k].upload(ndiag.data(), int64_t(ndiag.size())));
    // processing order: fibre by fibre (fibres in key order), 1D index ascending
    std::vector<int> ord;
    ord.reserve(g->cells);
    for (auto const &kv : fib)
      for (auto const &mc : members[kv.second])
        ord.push_back(mc.second);
    AKM_TRY(g->order[k].upload(ord.data(), int64_t(ord.size())));
  }

  // chains (kronmult::permutes, src/device/asgard_kronmult.hpp:152-220, and
  // the active-direction order of src/asgard_kronmult_matrix.cpp:1313-1331)
  std::vector<int> ch_term, ch_nst, ch_dim, ch_fill, ch_slot;
  for (int t = 0; t < m.T; t++)
  {
    std::vector<int> act;
    for (int k = 0; k < d; k++)
      if (active_mask == nullptr || active_mask[t * d + k] != 0)
        act.push_back(k);
    const int na = static_cast<int>(act.size());
    if (na > 1 && flux_dim != nullptr)
    {
      // the reference swaps positions 0 and flux_dim of the active list
      const int fd = flux_dim[t];
      if (fd >= 0 && fd < na && fd != act[0])
        std::swap(act[0], act[fd]);
    }
    if (na == 0)
    {
      ch_term.push_back(t);
      ch_nst.push_back(0);
      ch_dim.insert(ch_dim.end(), gk_max_dims, 0);
      ch_fill.insert(ch_fill.end(), gk_max_dims, 0);
      continue;
    }
    for (int p = 0; p < (1 << (na - 1)); p++)
    {
      std::vector<int> sgn(1, 0);
      int bits = p;
      for (int dd = 1; dd < na; dd++)
      {
        sgn.push_back((bits % 2 == 0) ? dd : -dd);
        bits /= 2;
      }
      std::sort(sgn.begin(), sgn.end());
      ch_term.push_back(t);
      ch_nst.push_back(na);
      for (int st = 0; st < gk_max_dims; st++)
      {
        if (st < na)
        {
          ch_dim.push_back(act[std::abs(sgn[st])]);
          ch_fill.push_back(sgn[st] < 0 ? gk_upper : (sgn[st] > 0 ? gk_lower : gk_full));
        }
        else
        {
          ch_dim.push_back(0);
          ch_fill.push_back(0);
        }
      }
    }
  }
  g->nchains = static_cast<int>(ch_term.size());
  g->max_stages = 0;
  int nslots = 0;
  for (int c = 0; c < g->nchains; c++)
  {
    g->max_stages = std::max(g->max_stages, ch_nst[c]);
    ch_slot.push_back(ch_nst[c] >= 1 ? nslots++ : -1);
  }
  g->nslots = nslots;
  g->ch_nst_host = ch_nst;
  if (g->nchains > 0)
  {
    AKM_TRY(g->ch_term.upload(ch_term.data(), int64_t(ch_term.size())));
    AKM_TRY(g->ch_nst.upload(ch_nst.data(), int64_t(ch_nst.size())));
    AKM_TRY(g->ch_dim.upload(ch_dim.data(), int64_t(ch_dim.size())));
    AKM_TRY(g->ch_fill.upload(ch_fill.data(), int64_t(ch_fill.size())));
    AKM_TRY(g->ch_slot.upload(ch_slot.data(), int64_t(ch_slot.size())));
  }
  // chains per stage (those with more than st stages)
  g->stage_list.resize(g->max_stages);
  g->stage_count.assign(g->max_stages, 0);
  for (int st = 0; st < g->max_stages; st++)
  {
    std::vector<int> list;
    for (int c = 0; c < g->nchains; c++)
      if (ch_nst[c] >= st + 1)
        list.push_back(c);
    g->stage_count[st] = static_cast<int>(list.size());
    if (!list.empty())
      AKM_TRY(g->stage_list[st].upload(list.data(), int64_t(list.size())));
  }
  if (g->max_stages > 0)
  {
    // first-stage groups: chains in order, grouped by (dimension, fill)
    std::vector<std::pair<int, int>> keys;
    std::vector<std::vector<int>> members;
    for (int c = 0; c < g->nchains; c++)
    {
      if (ch_nst[c] < 1)
        continue;
      const std::pair<int, int> key(ch_dim[c * gk_max_dims], ch_fill[c * gk_max_dims]);
      size_t i = 0;
      while (i < keys.size() && keys[i] != key)
        i++;
      if (i == keys.size())
      {
        keys.push_back(key);
        members.emplace_back();
      }
      members[i].push_back(c);
    }
    std::vector<int> gptr(1, 0), gchain;
    for (auto const &mem : members)
      for (size_t b = 0; b < mem.size(); b += gk_group_max)
      {
        for (size_t j = b; j < std::min(mem.size(), b + gk_group_max); j++)
          gchain.push_back(mem[j]);
        gptr.push_back(static_cast<int>(gchain.size()));
      }
    g->first_groups = static_cast<int>(gptr.size()) - 1;
    for (int i = 0; i < g->first_groups; i++)
      g->first_largest = std::max(g->first_largest, gptr[i + 1] - gptr[i]);
    AKM_TRY(g->first_ptr.upload(gptr.data(), int64_t(gptr.size())));
    AKM_TRY(g->first_chain.upload(gchain.data(), int64_t(gchain.size())));
  }
  if (nslots > 0)
    AKM_TRY(g->work.allocate(int64_t(2) * nslots * g->cells * m.N * int64_t(m.esize)));
  // flops as the reference counts them (src/asgard_kronmult_matrix.cpp:
  // 1389-1395): two per entry of the full DOF-level pattern of every
  // non-identity (term, dimension) on the padded set
  for (int t = 0; t < m.T; t++)
    for (int k = 0; k < d; k++)
      if (active_mask == nullptr || active_mask[t * d + k] != 0)
        g->flops += 2 * nnz_dim[k];
  m.gk = std::move(g);
  return AKM_SUCCESS;
}
} // namespace akm

using namespace akm;

extern "C" {

akm_status akm_create_global_grid(int d, int n, int64_t num_cells, const int *table,
                                  int max_level, int T, const void *vA, int64_t num_vA,
                                  const int *active, const int *flux_dim, akm_prec prec,
                                  akm_matrix *mat)
{
  if (mat == nullptr)
    return set_error(AKM_ERR_INVALID, "null matrix handle pointer");
  *mat = nullptr;
  if (d < 1 || d > gk_max_dims)
    return set_error(AKM_ERR_UNSUPPORTED, "global kron: d = %d not in [1, 6]", d);
  if (n < 1 || n > 16)
    return set_error(AKM_ERR_UNSUPPORTED, "global kron: n = %d not in [1, 16]", n);
  if (num_cells < 0 || num_cells > (int64_t(1) << 31) - 2 || T < 0)
    return set_error(AKM_ERR_INVALID, "global kron: invalid sizes");
  if (prec != AKM_F64 && prec != AKM_F32)
    return set_error(AKM_ERR_INVALID, "global kron: unknown precision %d", int(prec));
  if (table == nullptr && num_cells > 0)
    return set_error(AKM_ERR_INVALID, "global kron: null table");
  if (max_level < 0 || max_level > 20)
    return set_error(AKM_ERR_INVALID, "global kron: max_level %d not in [0, 20]", max_level);
  for (int64_t c = 0; c < num_cells; c++)
    for (int k = 0; k < d; k++)
    {
      const int L = table[c * 2 * d + k], p = table[c * 2 * d + d + k];
      if (L < 0 || L > max_level || p < 0 || (L <= 1 && p != 0) ||
          (L > 1 && p >= (1 << (L - 1))))
        return set_error(AKM_ERR_INVALID, "global kron: invalid cell (%d, %d)", L, p);
    }
  std::vector<int> p1, i1, dg;
  AKM_TRY(build_connect_1d(max_level, true, p1, i1, dg));
  const int64_t nconn = static_cast<int64_t>(i1.size());
  const int64_t expect_vA = nconn * d * T * int64_t(n) * n;
  if (T > 0 && (vA == nullptr || num_vA != expect_vA))
    return set_error(AKM_ERR_INVALID,
                     "global kron: vA has %lld values, connect_1d(%d) needs %lld",
                     static_cast<long long>(num_vA), max_level,
                     static_cast<long long>(expect_vA));
  std::vector<int> elem;
  cell_1d_indices(d, num_cells, table, elem);
  {
    std::map<std::vector<int>, int> seen;
    for (int64_t c = 0; c < num_cells; c++)
      if (!seen.emplace(std::vector<int>(elem.begin() + c * d, elem.begin() + (c + 1) * d), 0)
               .second)
        return set_error(AKM_ERR_INVALID, "global kron: cell %lld repeated",
                         static_cast<long long>(c));
  }
  std::unique_ptr<akm_matrix_s> m(new akm_matrix_s());
  m->prec = prec;
  m->esize = (prec == AKM_F64) ? 8 : 4;
  m->d = d;
  m->n = n;
  m->rows = m->cols = static_cast<int>(num_cells);
  m->T = T;
  int64_t N = 1;
  for (int k = 0; k < d; k++)
    N *= n;
  if (N > 4096)
    return set_error(AKM_ERR_UNSUPPORTED, "global kron: n^d = %lld exceeds 4096",
                     static_cast<long long>(N));
  m->N = N;
  m->mode = mode_global;
  m->max_level = max_level;
  m->M = 1 << max_level;
  m->num_vA = (T > 0) ? num_vA : 0;
  m->nconn = static_cast<int>(nconn);
  if (T > 0)
    AKM_TRY(m->vA.upload(static_cast<const unsigned char *>(vA), num_vA * m->esize));
  AKM_TRY(build_global(*m, elem, active, flux_dim));
  if (T > 0)
    AKM_TRY(global_update(*m, vA));
  *mat = m.release();
  return AKM_SUCCESS;
}

akm_status akm_global_info(akm_matrix mat, int64_t *padded_cells, int *chains, int *stages)
{
  if (mat == nullptr || mat->mode != mode_global || !mat->gk)
    return set_error(AKM_ERR_INVALID, "global info: not a global kron matrix");
  if (padded_cells)
    *padded_cells = mat->gk->cells - mat->gk->active_cells;
  if (chains)
    *chains = mat->gk->nchains;
  if (stages)
    *stages = mat->gk->max_stages;
  return AKM_SUCCESS;
}

akm_status akm_global_jacobi(akm_matrix mat, double *pc)
{
  if (mat == nullptr || mat->mode != mode_global || !mat->gk || pc == nullptr)
    return set_error(AKM_ERR_INVALID, "global jacobi: not a global kron matrix");
  const int64_t count = mat->gk->active_cells * mat->N;
  if (count > 0 && mat->gk->pc.data() != nullptr)
  {
    AKM_HIP(hipMemcpyAsync(pc, mat->gk->pc.data(), sizeof(double) * count, hipMemcpyDeviceToHost,
                           mat->stream));
    AKM_HIP(hipStreamSynchronize(mat->stream));
  }
  else if (count > 0)
    std::fill(pc, pc + count, 0.0);
  return AKM_SUCCESS;
}

} // extern "C"
