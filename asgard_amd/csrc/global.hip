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

// ---------------------------------------------------------------------------
// Slab layout.  The terms with the same one or two non-identity dimensions S
// form a group.  A slab of the group is a class of cells equal in every 1D
// index outside S; every stage of the group's chains maps a slab into
// itself (a stage along k in S reads only the cell's fibre along k), so one
// workgroup loads a slab's x once, runs all the group's chains on it in LDS
// and writes the group's sum once.  Small slabs are bundled into one
// workgroup; a slab too large for LDS is split over the other-index
// combinations (the tensor entries outside S, on which the group's factors
// act as the identity).  Per pass the 1D factor blocks of one (term,
// dimension), restricted to the 1D indices below the bundle's level bound,
// and the bundle's connection lists along that dimension are staged in LDS
// as well, so the connection loops touch only LDS.

// a bundle holds at most 256 blocks (one per thread), a large slab alone up
// to 1024
constexpr int gs_blocks_small = 256;
//! Blocks of a small bundle tried first: FP64 bundles of 512 blocks (512
//! threads) measured faster than 256 on the d=6 Vlasov operator, FP32 ones
//! slower.  A group whose 512-block bundles do not fit LDS is planned again
//! with 256.
__host__ __device__ constexpr int gs_blocks_small_first(int esize)
{
  return esize == 8 ? 512 : gs_blocks_small;
}
constexpr int gs_blocks_large = 1024;
constexpr int gs_lds_max = 160 * 1024;
constexpr int gs_pad = 2; // values of padding after each cell's blocks in LDS
static_assert(gs_pad % 2 == 0, "pair-major FP64 blocks need cells an even count of values apart");
constexpr int gs_split_level = 5; // slabs with level bound >= this run one chain per workgroup
constexpr int gs_lds_resident = 64 * 1024; // stage everything at once up to this size
constexpr int gs_unroll = 8;               // independent loads per thread while staging

//! Consecutive slabs of one group, their cells in bundle-local order.
struct gs_bundle
{
  int group;
  int cell0;    // first entry in gs_args::cells
  int ncell;
  int C;        // other-index combinations per chunk
  int Lp;       // level bound of the slab dimensions (compact factor table)
  int resident; // 1: every factor table and both lists staged once
  int list0;    // first int of the bundle's list images in gs_args::lists
  int llen[2];  // per slab dimension: image length (starts, diagonals, entries)
};

//! A group: its slab dimensions, factor tables and passes.
struct gs_group
{
  int nd;           // slab dimensions (1 or 2)
  int dim[2];       // ascending
  int str[2];       // their strides in the cell tensor
  int call;         // other-index combinations, n^(d - nd)
  int off0;         // first of its call entries in gs_args::offs
  int tab0, ntab;   // its (term, dimension) factor tables in gs_args::tabs
  int pass0, npass; // its passes in gs_args::passes
  int slot;         // work-buffer slot of the group's sum
};

//! One pass: a chain's stage.  mode 0: x -> sum (one-stage chain); 1: x ->
//! intermediate; 2: intermediate -> sum.
struct gs_pass
{
  int tab;  // factor table index within the group
  int j;    // slab dimension index (0 or 1)
  int fill, mode;
};

//! One workgroup's work: a chunk of a bundle with its group's data, so that
//! one record read starts it.
struct gs_item
{
  int cell0, ncell; // the bundle's cells in gs_args::cells
  int C;            // other-index combinations per chunk
  int cb, nq;       // the group's first compact factor table in gs_ctab, nq blocks each
  int resident;     // 1: every factor table and both lists staged once
  int list0;        // first int of the bundle's list images
  int llen[2];      // per slab dimension: image length
  int nd;           // slab dimensions (1 or 2)
  int str[2];       // their strides in the cell tensor
  int off;          // first of the chunk's C entries in gs_args::offs
  int tab0, ntab;   // the group's factor tables in gs_args::tabs
  int pass0, npass; // the group's passes in gs_args::passes
  int slot;         // work-buffer slot of the group's sum
  // a split item (one chain of a large slab): its chain of the bundle's nch
  // and the bundle's first partial in the partial buffer (-1: not split);
  // cell c's partial of chain k is partial part + c * nch + k (N values)
  int part, nch, chain;
};

struct gs_args
{
  int d, n, N, T;
  int64_t active_cells;
  int64_t stride; // work-buffer slot size
  const gs_pass *passes;
  const int *cells; // bundle cells (padded cell indices)
  // per bundle and slab dimension, the LDS image of its lists: ncell + 1
  // starts and ncell diagonal positions (relative to the entries), then
  // the entries (the connected cell's element offset in the bundle's x |
  // its factor block's element offset in the compact table << 16)
  const int *lists;
  const int *offs; // tensor offsets of the other-index combinations
};

//! LDS bytes of one work item: x (and the intermediate when nd = 2), ntab
//! factor tables of nq blocks, then the int lists (nlist_ints).
__host__ __device__ inline int64_t gs_lds_bytes(int esize, int nd, int n, int ncell, int C, int nq,
                                                int ntab, int64_t nlist_ints, int npass)
{
  const int64_t ds = (nd == 1) ? n : int64_t(n) * n;
  const int64_t bd = int64_t(ncell) * (ds * C + gs_pad);
  const int64_t vals = bd * nd + int64_t(ntab) * nq * n * n;
  return int64_t(16) * npass + (vals * esize + 15) / 16 * 16 + 4 * nlist_ints;
}

//! dst[i] = src[i], i < count, by a workgroup of TPB threads with
//! gs_unroll loads per thread in flight.
template<int TPB, typename T>
__device__ __forceinline__ void gs_copy(T *dst, const T *__restrict__ src, int count)
{
  // whole batches without bounds tests, then the remainder
  const int full = count / (gs_unroll * TPB) * (gs_unroll * TPB);
  for (int base = threadIdx.x; base < full; base += gs_unroll * TPB)
  {
    T v[gs_unroll];
#pragma unroll
    for (int u = 0; u < gs_unroll; u++)
      v[u] = src[base + u * TPB];
#pragma unroll
    for (int u = 0; u < gs_unroll; u++)
      dst[base + u * TPB] = v[u];
  }
  for (int base = full + threadIdx.x; base < count; base += gs_unroll * TPB)
  {
    T v[gs_unroll];
#pragma unroll
    for (int u = 0; u < gs_unroll; u++)
    {
      const int i = base + u * TPB;
      v[u] = (i < count) ? src[i] : T{};
    }
#pragma unroll
    for (int u = 0; u < gs_unroll; u++)
    {
      const int i = base + u * TPB;
      if (i < count)
        dst[i] = v[u];
    }
  }
}

//! One connection's block product along slab dimension J of ND: v[s] +=
//! sum_r A(s_J, r) * xb[s with s_J = r] over r in [rlo(s_J), rhi(s_J)): all
//! columns (FILL_FULL), from the row's own index on (upper diagonal block),
//! or before it (lower diagonal block).  Blocks are n^ND values, the last
//! slab digit fastest; A(s, r) = A[s + r * NT].
enum gs_cols : int
{
  gs_all = 0,
  gs_from_diag = 1,
  gs_before_diag = 2,
};

template<typename P, int NT, int ND, int J, int COLS>
__device__ __forceinline__ void gs_block_fma(const P (&av)[NT * NT],
                                             const P (&xv)[(ND == 1) ? NT : NT * NT], P *v)
{
  constexpr int ds = (ND == 1) ? NT : NT * NT;
#pragma unroll
  for (int si = 0; si < ds; si++)
  {
    const int s0 = (ND == 1) ? si : si / NT;
    const int s1 = (ND == 1) ? 0 : si % NT;
    const int sj = (J == 0) ? s0 : s1;
#pragma unroll
    for (int r = 0; r < NT; r++)
    {
      if ((COLS == gs_from_diag && r < sj) || (COLS == gs_before_diag && r >= sj))
        continue;
      const int xi = (ND == 1) ? r : ((J == 0) ? r * NT + s1 : s0 * NT + r);
      v[si] = fma(av[sj + r * NT], xv[xi], v[si]);
    }
  }
}

//! How a cell's blocks of DS values sit in LDS.  FP64 blocks with an even
//! value count are pair-major: values 2q and 2q + 1 of chunk entry o at
//! 2 (q C + o), one 16-byte read.  Other FP64 blocks are value-major (value
//! s of entry o at s C + o).  FP32 blocks are entry-major (DS o + s).  In
//! the FP64 orders a wavefront's lanes, which hold consecutive entries o,
//! read consecutive words instead of 32-byte strides, which spreads them
//! over the LDS banks; an FP32 block of four values is one 16-byte read.
template<typename P>
constexpr bool gs_value_major = sizeof(P) == 8;
template<typename P, int DS>
constexpr bool gs_pair_major = sizeof(P) == 8 && DS % 2 == 0;

//! The LDS offset of value si of a block whose entry starts at ob, with
//! value (pair, for pair-major blocks) stride vs.
template<typename P, int DS>
__device__ __forceinline__ int gs_at(int si, int ob, int vs)
{
  if constexpr (gs_pair_major<P, DS>)
    return ob + (si >> 1) * vs + (si & 1);
  else
    return ob + si * vs;
}

//! The factor block and the x block of one list entry: the block starts at
//! in + ob + the entry's cell offset, its values (pairs, when pair-major)
//! vs apart.  Factor blocks of an even value count are read in 16-byte
//! pairs (FP64 tables start 16-byte aligned, blocks n*n values apart).
template<typename P, int NT, int ND>
__device__ __forceinline__ void gs_block_load(const P *tab, const P *in, unsigned e, int ob,
                                              int vs, P (&av)[NT * NT],
                                              P (&xv)[(ND == 1) ? NT : NT * NT])
{
  constexpr int ds = (ND == 1) ? NT : NT * NT;
  const P *A = tab + (e >> 16);
  const P *xb = in + (e & 0xffffu) + ob;
  if constexpr (gs_pair_major<P, NT * NT>)
  {
#pragma unroll
    for (int q = 0; q < NT * NT / 2; q++)
    {
      const double2 t = reinterpret_cast<const double2 *>(A)[q];
      av[2 * q] = t.x;
      av[2 * q + 1] = t.y;
    }
  }
  else
  {
#pragma unroll
    for (int k = 0; k < NT * NT; k++)
      av[k] = A[k];
  }
  if constexpr (gs_pair_major<P, ds>)
  {
#pragma unroll
    for (int q = 0; q < ds / 2; q++)
    {
      const double2 t = *reinterpret_cast<const double2 *>(xb + q * vs);
      xv[2 * q] = t.x;
      xv[2 * q + 1] = t.y;
    }
  }
  else
  {
#pragma unroll
    for (int k = 0; k < ds; k++)
      xv[k] = xb[k * vs];
  }
}

//! One pass along slab dimension J for one block (cell c >= 0, block base
//! ob within the cell, values or pairs vs apart): the fill part of the
//! staged factor applied to `in`.
//! List entries hold the connected cell's element offset in `in` (low 16
//! bits) and its factor block's offset in `tab` (high 16 bits).  Each
//! value's sum runs over the connections ascending, then the factor's
//! columns, as in gk_cell_part; the diagonal block, restricted by the fill,
//! is the first connection of an upper fill and the last of a lower one.
//! The loop is software-pipelined: the blocks of the next connection and
//! the list entry after it are read while this connection's products run.
template<typename P, int NT, int ND, int J>
__device__ __forceinline__ void gs_pass_block(const P *in, const P *tab, const int *st,
                                              const int *dg, const unsigned *ent, int c, int ob,
                                              int vs, int fill, P *v)
{
  constexpr int ds = (ND == 1) ? NT : NT * NT;
  const int d0 = dg[c];
  const int lo = (fill == gk_upper) ? d0 + 1 : st[c];
  const int hi = (fill == gk_lower) ? d0 : st[c + 1];
  P av[NT * NT], xv[ds];
  if (fill == gk_upper)
  {
    gs_block_load<P, NT, ND>(tab, in, ent[d0], ob, vs, av, xv);
    gs_block_fma<P, NT, ND, J, gs_from_diag>(av, xv, v);
  }
  if (lo < hi)
  {
    gs_block_load<P, NT, ND>(tab, in, ent[lo], ob, vs, av, xv);
    unsigned en = ent[(lo + 1 < hi) ? lo + 1 : d0];
    for (int p = lo; p < hi; p++)
    {
      P an[NT * NT], xn[ds];
      gs_block_load<P, NT, ND>(tab, in, en, ob, vs, an, xn);
      en = ent[(p + 2 < hi) ? p + 2 : d0];
      gs_block_fma<P, NT, ND, J, gs_all>(av, xv, v);
#pragma unroll
      for (int k = 0; k < NT * NT; k++)
        av[k] = an[k];
#pragma unroll
      for (int k = 0; k < ds; k++)
        xv[k] = xn[k];
    }
  }
  if (fill == gk_lower)
  {
    gs_block_load<P, NT, ND>(tab, in, ent[d0], ob, vs, av, xv);
    gs_block_fma<P, NT, ND, J, gs_before_diag>(av, xv, v);
  }
}

//! One work item (bundle, chunk) per workgroup: every chain of the bundle's
//! group on the bundle's slabs, the group's sum written to its slot (active
//! cells only).  Chunk entries are blocks of n^ND values (the slab digits,
//! the last fastest), one per other-index combination of the chunk; a
//! thread owns one block.  A resident item stages x, every factor table of
//! the group and both dimensions' lists once; otherwise a pass stages its
//! own table and list when they change.
template<typename P, int NT, int ND, int TPB>
__global__ void __launch_bounds__(TPB)
    gk_slab(gs_args a, const P *__restrict__ ctab, const gs_item *__restrict__ items,
            const P *__restrict__ x, P *work, P *part)
{
  extern __shared__ __align__(16) unsigned char gs_smem[];
  constexpr int nn = NT * NT;
  constexpr int ds = (ND == 1) ? NT : nn;
  const gs_item b = items[blockIdx.x];
  const int C = b.C;
  // a cell's blocks in LDS, padded by gs_pad values so that the blocks of
  // neighbouring cells fall on different LDS banks
  const int D = ds * C + gs_pad;
  const int BD = b.ncell * D;
  const int nq = b.nq;
  const int ntab_lds = b.resident ? b.ntab : 1;
  gs_pass *pass = reinterpret_cast<gs_pass *>(gs_smem);
  P *X = reinterpret_cast<P *>(gs_smem + 16 * b.npass);
  P *Z = X + BD; // used when ND = 2
  P *tab = X + int64_t(BD) * ND;
  int *lists = reinterpret_cast<int *>(
      gs_smem + 16 * b.npass +
      ((int64_t(BD) * ND + int64_t(ntab_lds) * nq * nn) * int64_t(sizeof(P)) + 15) / 16 * 16);
  gs_copy<TPB>(reinterpret_cast<int *>(pass), reinterpret_cast<const int *>(a.passes + b.pass0),
               4 * b.npass);

  // this thread's block: cell c, other-index combination o of the chunk
  const int blk = threadIdx.x;
  const bool mine = blk < b.ncell * C;
  const int c = mine ? blk / C : 0;
  const int o = blk - (blk / C) * C;
  // where this entry's block starts within a cell, and its value (pair)
  // stride (pair-major, value-major or entry-major, as gs_at reads them);
  // cells are D values apart, D even for pair-major blocks (gs_pad is even),
  // so pairs stay 16-byte aligned
  const int vs = gs_pair_major<P, ds> ? 2 * C : (gs_value_major<P> ? C : 1);
  const int ob = gs_pair_major<P, ds> ? 2 * o : (gs_value_major<P> ? o : ds * o);
  const int gc = mine ? a.cells[b.cell0 + c] : -1;
  const bool active = gc >= 0 && gc < a.active_cells;
  const int64_t gbase = active ? int64_t(gc) * a.N + a.offs[b.off + o] : 0;
  auto gidx = [&](int si) -> int64_t {
    const int s0 = (ND == 1) ? si : si / NT;
    const int s1 = (ND == 1) ? 0 : si % NT;
    return gbase + int64_t(s0) * b.str[0] + ((ND == 1) ? 0 : int64_t(s1) * b.str[1]);
  };
  if (b.resident)
    gs_copy<TPB>(lists, a.lists + b.list0, b.llen[0] + (ND == 2 ? b.llen[1] : 0));
  P xv[ds];
#pragma unroll
  for (int si = 0; si < ds; si++)
    xv[si] = active ? x[gidx(si)] : P{0};
  if (b.resident)
    gs_copy<TPB>(tab, ctab + b.cb, b.ntab * nq * nn);
  if (mine)
  {
#pragma unroll
    for (int si = 0; si < ds; si++)
      X[c * D + gs_at<P, ds>(si, ob, vs)] = xv[si];
  }
  __syncthreads(); // the pass list is read before the first pass's barrier

  P acc[ds];
#pragma unroll
  for (int si = 0; si < ds; si++)
    acc[si] = 0;
  int cur_tab = -1, cur_j = -1;
  for (int pi = 0; pi < b.npass; pi++)
  {
    const gs_pass ps = pass[pi];
    if (!b.resident && (ps.tab != cur_tab || ps.j != cur_j))
    {
      // every earlier pass is done with the staged table and lists
      __syncthreads();
      if (ps.j != cur_j)
        gs_copy<TPB>(lists, a.lists + b.list0 + (ps.j == 0 ? 0 : b.llen[0]),
                     ps.j == 0 ? b.llen[0] : b.llen[1]);
      if (ps.tab != cur_tab)
        gs_copy<TPB>(tab, ctab + b.cb + int64_t(ps.tab) * nq * nn, nq * nn);
      cur_tab = ps.tab;
      cur_j = ps.j;
    }
    __syncthreads();
    if (mine)
    {
      const P *tb = b.resident ? tab + int64_t(ps.tab) * nq * nn : tab;
      const int *st = (b.resident && ps.j == 1) ? lists + b.llen[0] : lists;
      const int *dg = st + b.ncell + 1;
      const unsigned *ent = reinterpret_cast<const unsigned *>(dg + b.ncell);
      const P *in = (ps.mode == 2) ? Z : X;
      P v[ds];
#pragma unroll
      for (int si = 0; si < ds; si++)
        v[si] = 0;
      if (ND == 1 || ps.j == 0)
        gs_pass_block<P, NT, ND, 0>(in, tb, st, dg, ent, c, ob, vs, ps.fill, v);
      else
        gs_pass_block<P, NT, ND, 1>(in, tb, st, dg, ent, c, ob, vs, ps.fill, v);
#pragma unroll
      for (int si = 0; si < ds; si++)
      {
        if (ps.mode == 1)
          Z[c * D + gs_at<P, ds>(si, ob, vs)] = v[si];
        else
          acc[si] += v[si];
      }
    }
    // the next pass may overwrite what this one read, or read what it wrote
    __syncthreads();
  }
  if (active)
  {
    // the cell in the group's slot, or this chain's partial of the cell
    // (split items); entries at their offsets within the cell
    P *out = (b.part < 0) ? work + int64_t(b.slot) * a.stride + int64_t(gc) * a.N
                          : part + (int64_t(b.part) + int64_t(c) * b.nch + b.chain) * a.N;
#pragma unroll
    for (int si = 0; si < ds; si++)
      out[gidx(si) - int64_t(gc) * a.N] = acc[si];
  }
}

//! Where gk_sum finds the slab groups' sums of cells whose chains ran split:
//! psplit[g * active cells + cell] is the cell's first partial in part (-1:
//! the group's slot holds the sum), gnch[g] the group's chains; group g's
//! slot is gslot0 + g.
template<typename P>
struct gs_split_view
{
  const int *psplit = nullptr;
  const int *gnch = nullptr;
  const P *part = nullptr;
  int gslot0 = 0, ngroups = 0;
};

//! y = (alpha * sum of the sources + beta*y + gamma*xg) / (1 - jdt * pc) (no
//! division when pc is null).  The sources, in order: x (src < 0, a term of
//! identity factors) or work-buffer slot src (a chain's last stage in chain
//! order, or a slab group's sum at the place of the group's first chain; for
//! a cell whose group chains ran split, the sum of its chain partials in
//! chain order).
template<typename P>
__global__ void __launch_bounds__(gk_threads)
    gk_sum(gk_args a, const int *__restrict__ src, int nsrc, const P *__restrict__ x,
           const P *__restrict__ work, P alpha, P beta, P *y, P gamma, const P *xg,
           const double *pc, double jdt, gs_split_view<P> sv)
{
  const int64_t dofs = a.active_cells * a.N;
  for (int64_t i = blockIdx.x * int64_t(gk_threads) + threadIdx.x; i < dofs;
       i += int64_t(gk_threads) * gridDim.x)
  {
    P sum = 0;
    const int64_t cell = (sv.psplit != nullptr) ? i / a.N : 0;
    const int64_t e = i - cell * a.N;
    for (int c = 0; c < nsrc; c++)
    {
      const int sl = src[c];
      const int gi = sl - sv.gslot0;
      const int pb = (sv.psplit != nullptr && gi >= 0 && gi < sv.ngroups)
                         ? sv.psplit[int64_t(gi) * a.active_cells + cell]
                         : -1;
      P v;
      if (sl < 0)
        v = x[i];
      else if (pb < 0)
        v = work[int64_t(sl) * a.stride + i];
      else
      {
        v = sv.part[int64_t(pb) * a.N + e];
        for (int k = 1; k < sv.gnch[gi]; k++)
          v += sv.part[(int64_t(pb) + k) * a.N + e];
      }
      sum += v;
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
  // summation sources of gk_sum: chain layouts, and the slab layout
  akm::dev_buffer<int> src_chain, src_slab;
  int nsrc_chain = 0, nsrc_slab = 0;
  // slab layout (layout 6): groups of terms with the same one or two
  // non-identity dimensions, their bundles and work items; the other chains
  // run through the cell-layout stages
  bool slab = false; // chosen by global_autotune or akm_set_tuning
  int ngroups = 0;
  akm::dev_buffer<int> gs_passes, gs_tabs, gs_items, gs_cells, gs_lists, gs_offs, gs_qmap;
  // compact factor tables (per level bound, every group table), built from
  // vA by gs_tables at the first slab apply after vA changed (the matrix's
  // vA_stamp differs from gs_ctab_stamp); gs_lv: qmap starts then table
  // starts per level bound (gs_levels + 1 each)
  akm::dev_buffer<unsigned char> gs_ctab;
  akm::dev_buffer<int> gs_lv;
  int64_t gs_ctab_count = 0;
  uint64_t gs_ctab_stamp = 0;
  bool gs_ctab_valid = false;
  int gs_levels = 0;
  struct gs_launch
  {
    int tpb;     // 256 or 1024 threads
    int nd;      // slab dimensions of its groups
    bool side;   // on the side stream (split or 1024-thread items)
    int64_t lds; // dynamic LDS bytes (largest item of the launch)
    int first, count; // its items
  };
  std::vector<gs_launch> gs_launches;
  std::vector<int64_t> gs_item_lds; // dynamic LDS bytes per item
  // the last gs_moved items of the side launches run on the main stream after
  // its own launches (chosen by global_autotune: the stream that ends last
  // should be the main one, whose join then costs no wait)
  int gs_moved = 0;
  int gs_side_items = 0;
  // the 1024-thread launch (one large slab per workgroup, few workgroups)
  // runs beside the others on a stream of its own, forked and joined by events
  akm::dev_buffer<int> gs_psplit;     // per group and active cell: first partial or -1
  akm::dev_buffer<int> gs_gnch;       // per group: its chains
  akm::dev_buffer<unsigned char> gs_part; // split items' partials (N values each)
  hipStream_t gs_side = nullptr;
  hipEvent_t gs_fork = nullptr, gs_join = nullptr;
  ~akm_global_s()
  {
    if (gs_side)
      (void)hipStreamDestroy(gs_side);
    if (gs_fork)
      (void)hipEventDestroy(gs_fork);
    if (gs_join)
      (void)hipEventDestroy(gs_join);
  }
  std::vector<akm::dev_buffer<int>> stage_list_rest; // per stage: chains outside the groups
  std::vector<int> stage_count_rest;
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

//! Launches one slab work-item class: `count` items from `items`.
template<typename P, int TPB>
static int gs_launch_class(hipStream_t stream, int n, int nd, gs_args const &ga, const P *ctab,
                           const gs_item *items, int count, int64_t lds, const P *x, P *work,
                           P *part)
{
  auto go = [&](auto kern) -> int {
    if (lds > 64 * 1024 &&
        hipFuncSetAttribute(reinterpret_cast<const void *>(kern),
                            hipFuncAttributeMaxDynamicSharedMemorySize,
                            static_cast<int>(lds)) != hipSuccess)
    {
      (void)hipGetLastError();
      return set_error(AKM_ERR_HIP, "global kron slab: %lld bytes of LDS refused",
                       static_cast<long long>(lds));
    }
    kern<<<count, TPB, static_cast<size_t>(lds), stream>>>(ga, ctab, items, x, work, part);
    return check_launch("global kron slab");
  };
  switch (n * 4 + nd)
  {
  case 2 * 4 + 1:
    return go(gk_slab<P, 2, 1, TPB>);
  case 2 * 4 + 2:
    return go(gk_slab<P, 2, 2, TPB>);
  case 3 * 4 + 1:
    return go(gk_slab<P, 3, 1, TPB>);
  case 3 * 4 + 2:
    return go(gk_slab<P, 3, 2, TPB>);
  case 4 * 4 + 1:
    return go(gk_slab<P, 4, 1, TPB>);
  case 4 * 4 + 2:
    return go(gk_slab<P, 4, 2, TPB>);
  default:
    return set_error(AKM_ERR_UNSUPPORTED, "global kron slab: n = %d", n);
  }
}

//! Compact factor tables: entry w of block qc of group table t at level
//! bound Lp is entry w of block qmap[qm0[Lp] + qc] of the (term, dimension)
//! tabs[t] of vA.  lv: qm0 (levels + 1 entries), then the table starts
//! cbase (levels + 1 entries).
template<typename P>
__global__ void __launch_bounds__(gk_threads)
    gs_tables(const P *__restrict__ vA, const int *__restrict__ tabs, int ntab,
              const int *__restrict__ qmap, const int *__restrict__ lv, int levels, int nn,
              int64_t tstride, int64_t cstride, int64_t count, P *ctab)
{
  for (int64_t i = blockIdx.x * int64_t(gk_threads) + threadIdx.x; i < count;
       i += int64_t(gk_threads) * gridDim.x)
  {
    int Lp = 0;
    while (Lp + 1 < levels && i >= lv[levels + 1 + Lp + 1])
      Lp++;
    const int64_t rem = i - lv[levels + 1 + Lp];
    const int64_t per = int64_t(lv[Lp + 1] - lv[Lp]) * nn;
    const int t = static_cast<int>(rem / per);
    const int64_t r2 = rem - t * per;
    const int64_t qc = r2 / nn;
    const int w = static_cast<int>(r2 - qc * nn);
    const int q = qmap[lv[Lp] + qc];
    ctab[i] = vA[int64_t(q) * cstride + int64_t(tabs[2 * t]) * tstride +
                 int64_t(tabs[2 * t + 1]) * nn + w];
  }
}

//! Every slab group's sum into its work-buffer slot.
template<typename P>
static int gs_apply(akm_matrix_s &m, const P *vA, const P *x, P *work)
{
  akm_global_s &g = *m.gk;
  gs_args ga{};
  ga.d = m.d;
  ga.n = m.n;
  ga.N = static_cast<int>(m.N);
  ga.T = m.T;
  ga.active_cells = g.active_cells;
  ga.stride = g.cells * m.N;
  ga.passes = reinterpret_cast<const gs_pass *>(g.gs_passes.data());
  ga.cells = g.gs_cells.data();
  ga.lists = g.gs_lists.data();
  ga.offs = g.gs_offs.data();
  // the compact factor tables, rebuilt when vA changed since they were built
  const P *ctab = reinterpret_cast<const P *>(g.gs_ctab.data());
  if (!g.gs_ctab_valid || g.gs_ctab_stamp != m.vA_stamp)
  {
    const int nn = m.n * m.n;
    const int64_t tstride = int64_t(m.d) * nn;
    gs_tables<P><<<gk_blocks(g.gs_ctab_count), gk_threads, 0, m.stream>>>(
        vA, g.gs_tabs.data(), static_cast<int>(g.gs_tabs.size() / 2), g.gs_qmap.data(),
        g.gs_lv.data(), g.gs_levels, nn, tstride, tstride * m.T, g.gs_ctab_count,
        reinterpret_cast<P *>(g.gs_ctab.data()));
    AKM_TRY(check_launch("global kron slab tables"));
    g.gs_ctab_stamp = m.vA_stamp;
    g.gs_ctab_valid = true;
  }
  // the side launches come first and run on the side stream; gk_sum
  // follows the join
  // with every side item moved there is nothing to fork
  const bool fork = g.gs_side != nullptr && g.gs_moved < g.gs_side_items;
  if (fork)
  {
    AKM_HIP(hipEventRecord(g.gs_fork, m.stream));
    AKM_HIP(hipStreamWaitEvent(g.gs_side, g.gs_fork, 0));
  }
  const hipStream_t side = fork ? g.gs_side : m.stream;
  // items [0, bound) of the side launches run on the side stream, the rest
  // after the main stream's own launches
  const int bound = fork ? g.gs_side_items - std::min(g.gs_moved, g.gs_side_items)
                         : g.gs_side_items;
  P *part = reinterpret_cast<P *>(g.gs_part.data());
  auto run = [&](hipStream_t s, int tpb, int nd, int first, int count) -> int {
    if (count <= 0)
      return AKM_SUCCESS;
    const gs_item *items = reinterpret_cast<const gs_item *>(g.gs_items.data()) + first;
    const int64_t lds = *std::max_element(g.gs_item_lds.begin() + first,
                                          g.gs_item_lds.begin() + first + count);
    if (tpb == 1024)
      return gs_launch_class<P, 1024>(s, m.n, nd, ga, ctab, items, count, lds, x, work, part);
    if (tpb == 512)
      return gs_launch_class<P, 512>(s, m.n, nd, ga, ctab, items, count, lds, x, work, part);
    return gs_launch_class<P, 256>(s, m.n, nd, ga, ctab, items, count, lds, x, work, part);
  };
  for (auto const &l : g.gs_launches)
    if (l.side)
      AKM_TRY(run(side, l.tpb, l.nd, l.first, std::min(l.first + l.count, bound) - l.first));
  for (auto const &l : g.gs_launches)
    if (!l.side)
      AKM_TRY(run(m.stream, l.tpb, l.nd, l.first, l.count));
  for (auto const &l : g.gs_launches)
    if (l.side)
    {
      const int f = std::max(l.first, bound);
      AKM_TRY(run(m.stream, l.tpb, l.nd, f, l.first + l.count - f));
    }
  if (fork)
  {
    AKM_HIP(hipEventRecord(g.gs_join, g.gs_side));
    AKM_HIP(hipStreamWaitEvent(m.stream, g.gs_join, 0));
  }
  return AKM_SUCCESS;
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
  // the slab layout runs the chains outside its groups in the cell layout
  const bool slab = g.slab;
  const int lanes = slab ? (m.N <= wave_size ? 0 : 1) : g.lanes;
  for (int st = 0; st < g.max_stages; st++)
  {
    const int nl = slab ? g.stage_count_rest[st] : g.stage_count[st];
    if (nl == 0)
      continue;
    const int *list = slab ? g.stage_list_rest[st].data() : g.stage_list[st].data();
    const int blocks = (lanes == 0) ? gk_blocks_xcd(g.cells * nl * wave_size)
                                    : gk_blocks(g.cells * m.N * nl * lanes);
    if (!slab && lanes == 0 && st == 0 && g.first_grouped && g.first_largest > 1 && m.n >= 2 &&
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
      gk_stage<P, 8><<<blocks, gk_threads, 0, m.stream>>>(a, vA, list, nl, st, x, work);
    else
      gk_stage<P, 1><<<blocks, gk_threads, 0, m.stream>>>(a, vA, list, nl, st, x, work);
    AKM_TRY(check_launch("global kron stage"));
  }
  if (slab)
    AKM_TRY(gs_apply<P>(m, vA, x, work));
  const int *src = slab ? g.src_slab.data() : g.src_chain.data();
  const int nsrc = slab ? g.nsrc_slab : g.nsrc_chain;
  gs_split_view<P> sv{};
  if (slab && g.gs_psplit.size() > 0)
  {
    sv.psplit = g.gs_psplit.data();
    sv.gnch = g.gs_gnch.data();
    sv.part = reinterpret_cast<const P *>(g.gs_part.data());
    sv.gslot0 = 2 * g.nslots;
    sv.ngroups = g.ngroups;
  }
  gk_sum<P><<<gk_blocks(dofs), gk_threads, 0, m.stream>>>(a, src, nsrc, x, work, alpha, beta, y,
                                                          gamma, xg, pc, jdt, sv);
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
  if (m.gk->slab)
    return 6;
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
  if (layout != 1 && layout != 8 && (layout < 2 || layout > 6))
    return set_error(AKM_ERR_INVALID, "global kron layout %d is not 1, 2, 3, 4, 5, 6 or 8",
                     layout);
  const bool cell = (layout >= 2 && layout <= 5);
  if (cell && m.N > wave_size)
    return set_error(AKM_ERR_UNSUPPORTED, "global kron: the cell layout needs n^d <= 64");
  m.gk->slab = (layout == 6);
  if (m.gk->slab)
  {
    m.gk->lanes = 0;
    m.gk->first_grouped = false;
    m.gk->xcd_order = false;
    return AKM_SUCCESS;
  }
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
  bool best_xcd = false, best_slab = false;
  int best_moved = 0;
  // layouts 1 and 8, then the cell layout with and without the grouped
  // first stage (shared input reads against fewer, longer wavefronts), each
  // cell layout with and without the XCD-contiguous workgroup order, then
  // the slab layout (when some term has one or two non-identity dimensions)
  // with 0-100 % of its side-stream items moved to the main stream (100:
  // one stream, no fork and join)
  const int shapes[7][4] = {{1, 0, 0, 0}, {8, 0, 0, 0}, {0, 0, 0, 0}, {0, 0, 1, 0},
                            {0, 1, 0, 0}, {0, 1, 1, 0}, {0, 0, 0, 1}};
  for (auto const &shape : shapes)
  {
    const int lanes = shape[0];
    const bool grouped = shape[1] != 0;
    const bool xcd = shape[2] != 0;
    const bool slab = shape[3] != 0;
    if (lanes == 0 && !slab && m.N > wave_size)
      continue;
    if (grouped && g.first_largest < 2)
      continue;
    if (slab && g.ngroups == 0)
      continue;
    g.lanes = lanes;
    g.first_grouped = grouped;
    g.xcd_order = xcd;
    g.slab = slab;
    const int percents[9] = {0, 10, 20, 30, 40, 50, 65, 80, 100};
    const int nmoves = (slab && g.gs_side != nullptr) ? 9 : 1;
    for (int mi = 0; mi < nmoves && rc == AKM_SUCCESS; mi++)
    {
      g.gs_moved = g.gs_side_items * percents[mi] / 100;
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
        best_slab = slab;
        best_moved = g.gs_moved;
      }
    }
    if (rc != AKM_SUCCESS)
      break;
  }
  g.lanes = best_lanes;
  g.first_grouped = best_grouped;
  g.xcd_order = best_xcd;
  g.slab = best_slab;
  g.gs_moved = best_moved;
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

//! Level of a 1D cell index (0 for cell 0, l for 2^(l-1) <= i < 2^l).
static int level_1d(int i)
{
  int l = 0;
  while (i > 0)
  {
    i >>= 1;
    l++;
  }
  return l;
}

//! Per-dimension host copies of the fibre connection lists.
struct gk_lists
{
  std::vector<int> ptr, cell, q, diag;
};

//! One bundle while the slab plan is made.
struct gs_plan_bundle
{
  std::vector<int> cells;
  int C = 1, Lp = 0, tpb = 256;
};

//! Builds the slab layout: groups of the terms with one or two non-identity
//! dimensions (act[t] after the flux swap), their slabs, bundles, lists and
//! work items.  A group whose largest slab does not fit one workgroup at
//! one other-index combination per chunk is left to the cell-layout stages.
static int build_slabs(akm_matrix_s &m, akm_global_s &g, std::vector<gk_lists> const &lists,
                       std::vector<std::vector<int>> const &term_act,
                       std::vector<int> const &ch_term, std::vector<int> const &ch_nst,
                       std::vector<int> const &ch_dim, std::vector<int> const &ch_fill,
                       std::vector<int> const &ch_slot)
{
  const int d = m.d, n = m.n, T = m.T;
  const int L = m.max_level;
  const int nn = n * n;
  const int esize = static_cast<int>(m.esize);
  // compact factor index per level bound Lp: connect_1d entries with row and
  // column below 2^Lp, in connect_1d order
  std::vector<std::vector<int>> qc_of(L + 1);
  std::vector<int> qmap, qm0;
  for (int Lp = 0; Lp <= L; Lp++)
  {
    qm0.push_back(static_cast<int>(qmap.size()));
    qc_of[Lp].assign(g.c1_indx_h.size(), -1);
    const int lim = 1 << Lp;
    int cnt = 0;
    for (int r = 0; r < lim && r + 1 < static_cast<int>(g.c1_pntr_h.size()); r++)
      for (int q = g.c1_pntr_h[r]; q < g.c1_pntr_h[r + 1]; q++)
        if (g.c1_indx_h[q] < lim)
        {
          qc_of[Lp][q] = cnt++;
          qmap.push_back(q);
        }
  }
  qm0.push_back(static_cast<int>(qmap.size()));

  // groups by the set of non-identity dimensions
  std::vector<int> term_group(T, -1);
  std::vector<std::vector<int>> gdims;
  for (int t = 0; t < T; t++)
  {
    const int na = static_cast<int>(term_act[t].size());
    if (na < 1 || na > 2)
      continue;
    std::vector<int> key = term_act[t];
    std::sort(key.begin(), key.end());
    size_t gi = 0;
    while (gi < gdims.size() && gdims[gi] != key)
      gi++;
    if (gi == gdims.size())
      gdims.push_back(key);
    term_group[t] = static_cast<int>(gi);
  }

  std::vector<int> bundles, groups, passes, tabs, items, cells, lists_img, offs;
  struct item_rec
  {
    int bundle, h, tpb, nd;
    int64_t lds;
    int resident;
    int chain = -1;          // >= 0: this chain of the group only (a split item)
    int Lp = 0;              // level bound of its bundle
    int pass_off = 0, pass_cnt = 0; // its passes within the group's
  };
  std::vector<item_rec> irec;
  std::vector<int> loc(g.cells, -1);
  std::vector<int> gslot(gdims.size(), -1);
  int ngroups = 0;
  for (size_t gi = 0; gi < gdims.size(); gi++)
    for (int attempt = 0; attempt < 2; attempt++)
  {
    // small bundles of gs_blocks_small_first blocks, else (second attempt,
    // when those do not fit) of gs_blocks_small; a failed attempt commits
    // nothing
    const int blocks_small = (attempt == 0) ? gs_blocks_small_first(esize) : gs_blocks_small;
    if (attempt == 1 && blocks_small == gs_blocks_small_first(esize))
      break;
    const std::vector<int> &S = gdims[gi];
    const int nd = static_cast<int>(S.size());
    const int ds = (nd == 1) ? n : nn;
    const int64_t gs_cap_small = int64_t(blocks_small) * ds; // values
    const int64_t gs_cap_large = int64_t(gs_blocks_large) * ds;
    if (n < 2 || n > 4)
      continue; // the slab kernels hold n^nd-value blocks in registers
    int call = 1;
    for (int k = nd; k < d; k++)
      call *= n;
    // slabs: cells equal in every 1D index outside S, cell order kept
    std::map<std::vector<int>, std::vector<int>> slabs;
    std::vector<int> key;
    for (int64_t c = 0; c < g.cells; c++)
    {
      key.clear();
      for (int k = 0; k < d; k++)
        if (std::find(S.begin(), S.end(), k) == S.end())
          key.push_back(g.elem_host[c * d + k]);
      slabs[key].push_back(static_cast<int>(c));
    }
    int group_chains = 0;
    for (size_t c = 0; c < ch_term.size(); c++)
      if (term_group[ch_term[c]] == static_cast<int>(gi))
        group_chains++;
    // bundles: small slabs packed up to gs_cap_small values, a larger slab
    // alone with its other-index combinations split into chunks
    std::vector<gs_plan_bundle> plan;
    bool ok = true;
    gs_plan_bundle open;
    open.tpb = blocks_small;
    int64_t open_vals = 0;
    auto slab_level = [&](std::vector<int> const &sc) {
      int lp = 0;
      for (int c : sc)
        for (int k : S)
          lp = std::max(lp, level_1d(g.elem_host[int64_t(c) * d + k]));
      return lp;
    };
    for (auto const &kv : slabs)
    {
      const std::vector<int> &sc = kv.second;
      const int64_t msz = static_cast<int64_t>(sc.size());
      const int64_t full = msz * ds * call;
      const int lp = slab_level(sc);
      if (full <= gs_cap_small)
      {
        if (open_vals + full > gs_cap_small)
        {
          plan.push_back(open);
          open = gs_plan_bundle();
          open.tpb = blocks_small;
          open_vals = 0;
        }
        open.cells.insert(open.cells.end(), sc.begin(), sc.end());
        open.C = call;
        open.Lp = std::max(open.Lp, lp);
        open_vals += full;
        continue;
      }
      // a slab whose chains run split (long lists) takes the widest chunks
      const bool wide = lp >= gs_split_level && group_chains > 1;
      const int64_t cap = (msz * ds <= gs_cap_small && !wide) ? gs_cap_small : gs_cap_large;
      if (msz * ds > cap)
      {
        ok = false;
        break;
      }
      gs_plan_bundle one;
      one.cells = sc;
      one.Lp = lp;
      one.tpb = (cap == gs_cap_small) ? blocks_small : 1024;
      one.C = call;
      while (msz * ds * one.C > cap)
        one.C /= n;
      // x and the intermediate of a wide chunk within 64 KB, which leaves LDS
      // for the factor table and the lists
      while (wide && one.C > 1 && msz * ds * one.C * esize * nd > 64 * 1024)
        one.C /= n;
      plan.push_back(one);
    }
    if (ok && !open.cells.empty())
      plan.push_back(open);
    if (!ok)
      continue;

    // the group's terms in chain order: factor tables (term, S[j]) at
    // term index * nd + j
    std::vector<int> gterms;
    for (size_t c = 0; c < ch_term.size(); c++)
      if (term_group[ch_term[c]] == static_cast<int>(gi) &&
          std::find(gterms.begin(), gterms.end(), ch_term[c]) == gterms.end())
        gterms.push_back(ch_term[c]);
    const int ntab = static_cast<int>(gterms.size()) * nd;
    int npass = 0;
    std::vector<std::pair<int, int>> gchains; // (first pass, passes) per chain
    for (size_t c = 0; c < ch_term.size(); c++)
      if (term_group[ch_term[c]] == static_cast<int>(gi))
      {
        gchains.emplace_back(npass, ch_nst[c]);
        npass += ch_nst[c];
      }

    // each bundle's cells by their longest list, longest first: a wavefront
    // walks as many connections as its longest list, so long lists share
    // wavefronts (the order changes no operation order within a value)
    for (auto &pb : plan)
    {
      auto longest = [&](int gc) {
        int len = 0;
        for (int k : S)
          len = std::max(len, lists[k].ptr[gc + 1] - lists[k].ptr[gc]);
        return len;
      };
      std::stable_sort(pb.cells.begin(), pb.cells.end(),
                       [&](int u, int v) { return longest(u) > longest(v); });
    }

    // lists, LDS sizes and items of the plan (checked before anything is kept)
    const int group_index = ngroups;
    std::vector<int> b_rec, p_cells, p_lists;
    std::vector<item_rec> p_items;
    const int bundle_base = static_cast<int>(bundles.size() / (sizeof(gs_bundle) / sizeof(int)));
    for (size_t bi = 0; bi < plan.size() && ok; bi++)
    {
      gs_plan_bundle const &pb = plan[bi];
      const int ncell = static_cast<int>(pb.cells.size());
      gs_bundle rec{};
      rec.group = group_index;
      rec.cell0 = static_cast<int>(cells.size() + p_cells.size());
      rec.ncell = ncell;
      rec.C = pb.C;
      rec.Lp = pb.Lp;
      rec.list0 = static_cast<int>(lists_img.size() + p_lists.size());
      for (int c = 0; c < ncell; c++)
        loc[pb.cells[c]] = c;
      int nent_max = 0, nent_sum = 0;
      for (int j = 0; j < nd && ok; j++)
      {
        // image: ncell + 1 starts, ncell diagonal positions, the entries
        gk_lists const &ls = lists[S[j]];
        const size_t img = p_lists.size();
        p_lists.resize(img + 2 * size_t(ncell) + 1);
        int cnt = 0;
        for (int c = 0; c < ncell && ok; c++)
        {
          const int gc = pb.cells[c];
          p_lists[img + c] = cnt;
          p_lists[img + ncell + 1 + c] = cnt + ls.diag[gc] - ls.ptr[gc];
          for (int p = ls.ptr[gc]; p < ls.ptr[gc + 1]; p++)
          {
            // element offset of the connected cell in x / the intermediate,
            // element offset of its factor block in the table
            const int64_t xo = int64_t(loc[ls.cell[p]]) * (ds * pb.C + gs_pad);
            const int64_t ao = int64_t(qc_of[pb.Lp][ls.q[p]]) * nn;
            if (xo < 0 || ao < 0 || xo > 0xffff || ao > 0xffff)
            {
              ok = false; // outside the slab (cannot happen) or past 16-bit offsets
              break;
            }
            p_lists.push_back(static_cast<int>(static_cast<unsigned>(xo) |
                                               (static_cast<unsigned>(ao) << 16)));
            cnt++;
          }
        }
        p_lists[img + ncell] = cnt;
        rec.llen[j] = static_cast<int>(p_lists.size() - img);
        nent_max = std::max(nent_max, cnt);
        nent_sum += cnt;
      }
      for (int c = 0; c < ncell; c++)
        loc[pb.cells[c]] = -1;
      if (!ok)
        break;
      const int nq = qm0[pb.Lp + 1] - qm0[pb.Lp];
      // everything staged once when that keeps two workgroups per CU, else
      // one table and one list at a time
      int64_t lds = gs_lds_bytes(esize, nd, n, ncell, pb.C, nq, ntab,
                                 int64_t(nd) * (2 * ncell + 1) + nent_sum, npass);
      rec.resident = (lds <= gs_lds_resident) ? 1 : 0;
      if (!rec.resident)
        lds = gs_lds_bytes(esize, nd, n, ncell, pb.C, nq, 1, 2 * ncell + 1 + nent_max, npass);
      if (lds > gs_lds_max)
      {
        ok = false;
        break;
      }
      const int *rp = reinterpret_cast<const int *>(&rec);
      b_rec.insert(b_rec.end(), rp, rp + sizeof(gs_bundle) / sizeof(int));
      p_cells.insert(p_cells.end(), pb.cells.begin(), pb.cells.end());
      // a slab with long connection lists (level bound >= gs_split_level) runs
      // one workgroup per chain: its passes are the longest
      const bool split = pb.Lp >= gs_split_level && gchains.size() > 1;
      for (int h = 0; h < call / pb.C; h++)
      {
        if (!split)
        {
          p_items.push_back({bundle_base + static_cast<int>(bi), h, pb.tpb, nd, lds,
                             rec.resident});
          continue;
        }
        for (size_t k = 0; k < gchains.size(); k++)
        {
          item_rec r{bundle_base + static_cast<int>(bi), h, pb.tpb, nd,
                     gs_lds_bytes(esize, nd, n, ncell, pb.C, nq, 1, 2 * ncell + 1 + nent_max,
                                  gchains[k].second),
                     0};
          r.chain = static_cast<int>(k);
          r.Lp = pb.Lp;
          r.pass_off = gchains[k].first;
          r.pass_cnt = gchains[k].second;
          p_items.push_back(r);
        }
      }
    }
    if (!ok)
      continue;
    // the group is kept
    gs_group grec{};
    grec.nd = nd;
    for (int j = 0; j < nd; j++)
    {
      grec.dim[j] = S[j];
      int str = 1;
      for (int k = S[j] + 1; k < d; k++)
        str *= n;
      grec.str[j] = str;
    }
    grec.call = call;
    grec.off0 = static_cast<int>(offs.size());
    for (int o = 0; o < call; o++)
    {
      // digits of o over the other dimensions, the last one fastest
      int rem = o, off = 0, str = 1;
      for (int k = d - 1; k >= 0; k--)
      {
        if (std::find(S.begin(), S.end(), k) == S.end())
        {
          off += (rem % n) * str;
          rem /= n;
        }
        str *= n;
      }
      offs.push_back(off);
    }
    grec.tab0 = static_cast<int>(tabs.size() / 2);
    grec.ntab = ntab;
    for (int t : gterms)
      for (int j = 0; j < nd; j++)
        tabs.insert(tabs.end(), {t, S[j]});
    grec.pass0 = static_cast<int>(passes.size() / 4);
    for (size_t c = 0; c < ch_term.size(); c++)
    {
      const int t = ch_term[c];
      if (term_group[t] != static_cast<int>(gi))
        continue;
      const int ti = static_cast<int>(std::find(gterms.begin(), gterms.end(), t) -
                                      gterms.begin());
      const int ns = ch_nst[c];
      for (int st = 0; st < ns; st++)
      {
        const int k = ch_dim[c * gk_max_dims + st];
        const int j = (k == S[0]) ? 0 : 1;
        const gs_pass ps{ti * nd + j, j, ch_fill[c * gk_max_dims + st],
                         (ns == 1) ? 0 : (st == 0 ? 1 : 2)};
        passes.insert(passes.end(), {ps.tab, ps.j, ps.fill, ps.mode});
      }
    }
    grec.npass = static_cast<int>(passes.size() / 4) - grec.pass0;
    grec.slot = 2 * g.nslots + group_index;
    const int *gp = reinterpret_cast<const int *>(&grec);
    groups.insert(groups.end(), gp, gp + sizeof(gs_group) / sizeof(int));
    bundles.insert(bundles.end(), b_rec.begin(), b_rec.end());
    cells.insert(cells.end(), p_cells.begin(), p_cells.end());
    lists_img.insert(lists_img.end(), p_lists.begin(), p_lists.end());
    irec.insert(irec.end(), p_items.begin(), p_items.end());
    gslot[gi] = group_index;
    ngroups++;
    break;
  }
  for (int t = 0; t < T; t++)
    if (term_group[t] >= 0 && gslot[term_group[t]] < 0)
      term_group[t] = -1; // its group did not fit: cell-layout stages
  g.ngroups = ngroups;

  // launches: 1024-thread items, then the small bundles' (512 or 256
  // threads), per slab dimension count (one launch for all items of a size,
  // largest LDS first, measured faster than separate launches above and
  // below 32 KB of LDS)
  // the side stream takes the 1024-thread items (the split items of the
  // largest slabs among them), about half the work on the Vlasov operator;
  // they come first, and gs_apply may run the last gs_moved of them on the
  // main stream
  auto on_side = [](item_rec const &r) { return r.tpb == 1024; };
  auto cls = [&](item_rec const &r) {
    const int side = on_side(r) ? 0 : 1;
    const int size = (r.tpb == 1024) ? 0 : (r.tpb == 512 ? 1 : 2);
    return (side * 3 + size) * 2 + (r.nd - 1);
  };
  // within a launch the largest items first (the longest-running workgroups
  // start early)
  std::stable_sort(irec.begin(), irec.end(), [&](item_rec const &u, item_rec const &v) {
    return cls(u) != cls(v) ? cls(u) < cls(v) : u.lds > v.lds;
  });
  g.gs_launches.clear();
  g.gs_item_lds.clear();
  g.gs_side_items = 0;
  // partials of split bundles: per bundle ncell * chains of N values
  const size_t nbundles = bundles.size() / (sizeof(gs_bundle) / sizeof(int));
  std::vector<int64_t> part0(nbundles, -1);
  std::vector<int> nch_of(nbundles, 0);
  int64_t nparts = 0;
  for (auto const &r : irec)
    if (r.chain >= 0)
      nch_of[r.bundle] = std::max(nch_of[r.bundle], r.chain + 1);
  for (size_t bi = 0; bi < nbundles; bi++)
    if (nch_of[bi] > 0)
    {
      part0[bi] = nparts;
      nparts += int64_t(reinterpret_cast<const gs_bundle *>(bundles.data())[bi].ncell) *
                nch_of[bi];
    }
  if (nparts > (int64_t(1) << 31) - 1)
    return set_error(AKM_ERR_UNSUPPORTED, "global kron: too many slab partials");
  // compact factor tables, per level bound every group table: cbase[Lp]
  const int64_t ntab_all = static_cast<int64_t>(tabs.size() / 2);
  std::vector<int64_t> cbase(L + 2, 0);
  for (int Lp = 0; Lp <= L; Lp++)
    cbase[Lp + 1] = cbase[Lp] + ntab_all * (qm0[Lp + 1] - qm0[Lp]) * nn;
  if (cbase[L + 1] > (int64_t(1) << 31) - 1)
    return set_error(AKM_ERR_UNSUPPORTED, "global kron: slab factor tables too large");
  const gs_bundle *hb = reinterpret_cast<const gs_bundle *>(bundles.data());
  const gs_group *hg = reinterpret_cast<const gs_group *>(groups.data());
  for (size_t i = 0; i < irec.size(); i++)
  {
    const gs_bundle &bb = hb[irec[i].bundle];
    const gs_group &gg = hg[bb.group];
    gs_item it{};
    it.cell0 = bb.cell0;
    it.ncell = bb.ncell;
    it.C = bb.C;
    it.nq = qm0[bb.Lp + 1] - qm0[bb.Lp];
    it.cb = static_cast<int>(cbase[bb.Lp] + int64_t(gg.tab0) * it.nq * nn);
    it.resident = irec[i].resident;
    it.list0 = bb.list0;
    it.llen[0] = bb.llen[0];
    it.llen[1] = bb.llen[1];
    it.nd = gg.nd;
    it.str[0] = gg.str[0];
    it.str[1] = gg.str[1];
    it.off = gg.off0 + irec[i].h * bb.C;
    it.tab0 = gg.tab0;
    it.ntab = gg.ntab;
    const bool split = irec[i].chain >= 0;
    it.pass0 = gg.pass0 + (split ? irec[i].pass_off : 0);
    it.npass = split ? irec[i].pass_cnt : gg.npass;
    // a split item writes its chain's partials, which gk_sum adds
    it.slot = gg.slot;
    it.part = split ? part0[irec[i].bundle] : -1;
    it.nch = split ? nch_of[irec[i].bundle] : 0;
    it.chain = split ? irec[i].chain : 0;
    const int *ip = reinterpret_cast<const int *>(&it);
    items.insert(items.end(), ip, ip + sizeof(gs_item) / sizeof(int));
    if (g.gs_launches.empty() || cls(irec[i - 1]) != cls(irec[i]))
      g.gs_launches.push_back({irec[i].tpb, irec[i].nd, on_side(irec[i]), 0,
                               static_cast<int>(i), 0});
    auto &l = g.gs_launches.back();
    l.lds = std::max(l.lds, irec[i].lds);
    l.count++;
    g.gs_item_lds.push_back(irec[i].lds);
    if (on_side(irec[i]))
      g.gs_side_items = static_cast<int>(i) + 1;
  }
  // until global_autotune measures it, a quarter of the side items run on
  // the main stream (near the best split of the FP64 Vlasov operator)
  g.gs_moved = g.gs_side_items / 4;
  // gk_sum's view of split bundles (every chunk of a split bundle is split):
  // per group and active cell the first of its chain partials
  if (nparts > 0)
  {
    std::vector<int> psplit(size_t(ngroups) * g.active_cells, -1), gnch(ngroups, 0);
    for (size_t bi = 0; bi < nbundles; bi++)
    {
      if (nch_of[bi] == 0)
        continue;
      const gs_bundle &bb = hb[bi];
      gnch[bb.group] = nch_of[bi];
      for (int c = 0; c < bb.ncell; c++)
      {
        const int gc = cells[bb.cell0 + c];
        if (gc < g.active_cells)
          psplit[size_t(bb.group) * g.active_cells + gc] =
              static_cast<int>(part0[bi] + int64_t(c) * nch_of[bi]);
      }
    }
    AKM_TRY(g.gs_psplit.upload(psplit.data(), int64_t(psplit.size())));
    AKM_TRY(g.gs_gnch.upload(gnch.data(), int64_t(gnch.size())));
  }
  if (nparts > 0)
    AKM_TRY(g.gs_part.allocate(nparts * m.N * esize));
  if (!g.gs_launches.empty() && g.gs_launches[0].side && !g.gs_launches.back().side &&
      g.gs_side == nullptr)
  {
    AKM_HIP(hipStreamCreateWithFlags(&g.gs_side, hipStreamNonBlocking));
    AKM_HIP(hipEventCreateWithFlags(&g.gs_fork, hipEventDisableTiming));
    AKM_HIP(hipEventCreateWithFlags(&g.gs_join, hipEventDisableTiming));
  }
  if (ngroups > 0)
  {
    AKM_TRY(g.gs_passes.upload(passes.data(), int64_t(passes.size())));
    AKM_TRY(g.gs_tabs.upload(tabs.data(), int64_t(tabs.size())));
    AKM_TRY(g.gs_items.upload(items.data(), int64_t(items.size())));
    AKM_TRY(g.gs_cells.upload(cells.data(), int64_t(cells.size())));
    AKM_TRY(g.gs_lists.upload(lists_img.data(), int64_t(lists_img.size())));
    AKM_TRY(g.gs_offs.upload(offs.data(), int64_t(offs.size())));
    AKM_TRY(g.gs_qmap.upload(qmap.data(), int64_t(qmap.size())));
    std::vector<int> lv(qm0.begin(), qm0.end());
    for (int64_t v : cbase)
      lv.push_back(static_cast<int>(v));
    AKM_TRY(g.gs_lv.upload(lv.data(), int64_t(lv.size())));
    AKM_TRY(g.gs_ctab.allocate(cbase[L + 1] * esize));
    g.gs_ctab_count = cbase[L + 1];
    g.gs_ctab_valid = false;
    g.gs_levels = L + 1;
  }

  // the chains outside the groups, per stage, and the summation sources
  g.stage_list_rest.clear();
  g.stage_list_rest.resize(g.max_stages);
  g.stage_count_rest.assign(g.max_stages, 0);
  for (int st = 0; st < g.max_stages; st++)
  {
    std::vector<int> list;
    for (size_t c = 0; c < ch_term.size(); c++)
      if (term_group[ch_term[c]] < 0 && ch_nst[c] >= st + 1)
        list.push_back(static_cast<int>(c));
    g.stage_count_rest[st] = static_cast<int>(list.size());
    if (!list.empty())
      AKM_TRY(g.stage_list_rest[st].upload(list.data(), int64_t(list.size())));
  }
  std::vector<int> src_chain, src_slab;
  std::vector<bool> emitted(gdims.size(), false);
  for (size_t c = 0; c < ch_term.size(); c++)
  {
    const int ns = ch_nst[c];
    const int own = (ns == 0) ? -1 : ((ns - 1) & 1) * g.nslots + ch_slot[c];
    src_chain.push_back(own);
    const int gi = term_group[ch_term[c]];
    if (gi < 0)
      src_slab.push_back(own);
    else if (!emitted[gi])
    {
      emitted[gi] = true;
      src_slab.push_back(2 * g.nslots + gslot[gi]);
    }
  }
  g.nsrc_chain = static_cast<int>(src_chain.size());
  g.nsrc_slab = static_cast<int>(src_slab.size());
  if (!src_chain.empty())
    AKM_TRY(g.src_chain.upload(src_chain.data(), int64_t(src_chain.size())));
  if (!src_slab.empty())
    AKM_TRY(g.src_slab.upload(src_slab.data(), int64_t(src_slab.size())));
  return AKM_SUCCESS;
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
  std::vector<gk_lists> lists(d);
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
    lists[k].ptr = std::move(ptr);
    lists[k].cell = std::move(ncell);
    lists[k].q = std::move(nq);
    lists[k].diag = std::move(ndiag);
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
  std::vector<std::vector<int>> term_act(m.T);
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
    term_act[t] = act;
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
  AKM_TRY(build_slabs(m, *g, lists, term_act, ch_term, ch_nst, ch_dim, ch_fill, ch_slot));
  const int64_t work_slots = int64_t(2) * nslots + g->ngroups;
  if (work_slots > 0)
    AKM_TRY(g->work.allocate(work_slots * g->cells * m.N * int64_t(m.esize)));
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

akm_status akm_global_slab_info(akm_matrix mat, int *groups, int *rest_chains, int *sum_sources)
{
  if (mat == nullptr || mat->mode != mode_global || !mat->gk)
    return set_error(AKM_ERR_INVALID, "global slab info: not a global kron matrix");
  akm_global_s const &g = *mat->gk;
  if (groups)
    *groups = g.ngroups;
  if (rest_chains)
    *rest_chains = g.max_stages > 0 ? g.stage_count_rest[0] : 0;
  if (sum_sources)
    *sum_sources = g.nsrc_slab;
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
