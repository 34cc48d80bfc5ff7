// Sum-factorised kronmult on the FP64 matrix cores: d = 3, n = 3, sparse-grid
// mode (included by kron_f64.hip).
//
// For an output row r and a term t the apply is
//   y_r += sum over connected columns c of (A0[r0,c0] x A1[r1,c1] x A2[r2,c2]) x_c
// and the factor of dimension k depends only on (r_k, c_k).  With the row's
// columns in lexicographic order of their 1D indices (c0, c1, c2) the sum
// factorises: dimension 2 is contracted column by column and summed over the
// columns that share (c0, c1); that sum is contracted over dimension 1 and
// summed over the groups that share c0; that sum is contracted over
// dimension 0 once per c0 group.  Per (pair, term) the first contraction is
// the only per-column work; the other two run once per group.
//
// One wavefront holds one 3 x 3 x 3 tensor, padded to 4 x 4 x 4, in the lane
// layout of v_mfma_f64_4x4x4_4b_f64: with lane = 16 q + 4 g + m the
// instruction runs four independent 4 x 4 x 4 products (block g) with
// A[i][k] in lane (i = m, k = q), B[k][j] in lane (k = q, j = m) and the
// result D[i][j] in lane (i = q, j = m).  The result therefore sits where the
// next product's B operand must be, so the dimension-2 and dimension-1
// contractions chain on the matrix cores without moving data, and the
// matrix cores' accumulator input does the group sums:
//   stage 1 (dim 2): block g = i0, A = x[i0][i1 = m][i2 = q],
//                    B = A2(j2 = m, i2 = q)          -> D1[i1 = q][j2 = m]
//   stage 2 (dim 1): A = A1(j1 = m, i1 = q), B = D1  -> D2[j1 = q][j2 = m]
//   stage 3 (dim 0): lane (j1 = q, j0 = g, j2 = m) sums A0(g, i0) D2 over the
//                    blocks i0 = 0..2 (lane reads within the 16-lane row).
// Padding entries (index 3) hold zeros in every operand.
//
// A workgroup of W wavefronts owns a row; wavefront w takes the w-th of W
// contiguous pieces of the row's column list (each piece closes its own
// groups, which is exact because the apply is linear in the columns).  The
// pieces' sums meet in LDS in wavefront order: results repeat bit for bit.
#pragma once

#include "kron_impl.hpp"

namespace akm
{
constexpr int mfma_max_terms = 8;

__device__ __forceinline__ double mfma4(double a, double b, double c)
{
  return __builtin_amdgcn_mfma_f64_4x4x4f64(a, b, c, 0, 0, 0);
}

//! 1D index of dimension k from a packed cell word (10 bits per dimension).
__device__ __forceinline__ int packed_1d(unsigned long long w, int k)
{
  return static_cast<int>((w >> (10 * k)) & 1023u);
}

template<int W, int TT>
__global__ void __launch_bounds__(W *wave_size) kron_mfma_d3n3(kron_args<double> a)
{
  constexpr int N = 27;
  __shared__ double red[W * wave_size];
  const int lane = threadIdx.x & (wave_size - 1);
  // uniform per wavefront: loop bounds and group keys stay scalar
  const int wave = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x) / wave_size);
  const int q = lane >> 4, g = (lane >> 2) & 3, m = lane & 3;
  const int T = a.T;

  const int row = a.row_order[blockIdx.x];
  const int64_t jbeg = a.pntr[row], jend = a.pntr[row + 1];
  const int64_t len = jend - jbeg;
  const int64_t cb = jbeg + (len * wave) / W, ce = jbeg + (len * (wave + 1)) / W;

  const int r0M = a.elem[int64_t(row) * 3 + 0] * a.M;
  const int r1M = a.elem[int64_t(row) * 3 + 1] * a.M;
  const int r2M = a.elem[int64_t(row) * 3 + 2] * a.M;
  const int64_t blk = int64_t(T) * 27; // one connection: T terms x 3 dims x 9

  // lane roles: the x entry of the stage-1 A operand, the factor entry of
  // the stage-1 B and stage-2 A operands (both F[q*3 + m]), the stage-3 rows
  const bool xvalid = (g < 3 && m < 3 && q < 3);
  const int xoff = g * 9 + m * 3 + q;
  const bool fvalid = (q < 3 && m < 3);
  const int foff = q * 3 + m;

  double s1[TT], s2[TT];
  double f1[TT];    // stage-2 factor of the open (c0, c1) group
  double f0[TT][3]; // stage-3 factor A0(g, i0) of the open c0 group
#pragma unroll
  for (int t = 0; t < TT; t++)
  {
    s1[t] = 0;
    s2[t] = 0;
    f1[t] = 0;
    f0[t][0] = f0[t][1] = f0[t][2] = 0;
  }
  double acc = 0;

  auto load_f1 = [&](int c1) {
    const int64_t base = int64_t(a.off1d[r1M + c1]) * blk + 9;
#pragma unroll
    for (int t = 0; t < TT; t++)
      if (t < T)
        f1[t] = fvalid ? a.vA[base + t * 27 + foff] : 0.0;
  };
  auto load_f0 = [&](int c0) {
    const int64_t base = int64_t(a.off1d[r0M + c0]) * blk;
#pragma unroll
    for (int t = 0; t < TT; t++)
      if (t < T)
#pragma unroll
        for (int i0 = 0; i0 < 3; i0++)
          f0[t][i0] = (g < 3) ? a.vA[base + t * 27 + i0 * 3 + g] : 0.0;
  };
  // closes the open (c0, c1) group: its dimension-2 sum is contracted over
  // dimension 1 into the c0 group's sum
  auto close_01 = [&]() {
#pragma unroll
    for (int t = 0; t < TT; t++)
      if (t < T)
      {
        s2[t] = mfma4(f1[t], s1[t], s2[t]);
        s1[t] = 0;
      }
  };
  // closes the open c0 group: contraction over dimension 0 into acc
  auto close_0 = [&]() {
    const int src = 16 * q + m;
#pragma unroll
    for (int t = 0; t < TT; t++)
      if (t < T)
      {
        const double v0 = __shfl(s2[t], src, wave_size);
        const double v1 = __shfl(s2[t], src + 4, wave_size);
        const double v2 = __shfl(s2[t], src + 8, wave_size);
        acc = fma(f0[t][0], v0, acc);
        acc = fma(f0[t][1], v1, acc);
        acc = fma(f0[t][2], v2, acc);
        s2[t] = 0;
      }
  };

  unsigned int prev01 = 0xFFFFFFFFu, prev0 = 0xFFFFFFFFu;
  for (int64_t j0 = cb; j0 < ce; j0 += wave_size)
  {
    // one batch of up to 64 columns: ids, packed 1D indices and dimension-2
    // factor offsets, one column per lane
    const int64_t jl = j0 + lane;
    int colv = 0, off2v = 0;
    unsigned int cplo = 0;
    if (jl < ce)
    {
      colv = a.indx[jl];
      const unsigned long long w = a.epack[colv];
      cplo = static_cast<unsigned int>(w & 0x3FFFFFFFull);
      off2v = a.off1d[r2M + packed_1d(w, 2)];
    }
    const int nb = static_cast<int>((ce - j0 < wave_size) ? ce - j0 : wave_size);
    // software pipeline: the x entry and stage-1 factors of column jj + 1
    // are loaded before column jj is contracted
    double xn = 0, fn[TT];
    auto fetch = [&](int jj, double &xv, double (&fv)[TT]) {
      const int col = __builtin_amdgcn_readlane(colv, jj);
      const int off2 = __builtin_amdgcn_readlane(off2v, jj);
      const int64_t base = int64_t(off2) * blk + 18;
      xv = xvalid ? a.x[int64_t(col) * N + xoff] : 0.0;
#pragma unroll
      for (int t = 0; t < TT; t++)
        if (t < T)
          fv[t] = fvalid ? a.vA[base + t * 27 + foff] : 0.0;
    };
    fetch(0, xn, fn);
    for (int jj = 0; jj < nb; jj++)
    {
      const double xv = xn;
      double fv[TT];
#pragma unroll
      for (int t = 0; t < TT; t++)
        fv[t] = fn[t];
      if (jj + 1 < nb)
        fetch(jj + 1, xn, fn);
      const unsigned int cp = static_cast<unsigned int>(__builtin_amdgcn_readlane(cplo, jj));
      const unsigned int k01 = cp & 0xFFFFFu, k0 = cp & 0x3FFu;
      if (k01 != prev01)
      {
        if (prev01 != 0xFFFFFFFFu)
          close_01();
        if (k0 != prev0)
        {
          if (prev0 != 0xFFFFFFFFu)
            close_0();
          load_f0(static_cast<int>(k0));
          prev0 = k0;
        }
        load_f1(static_cast<int>((cp >> 10) & 1023u));
        prev01 = k01;
      }
#pragma unroll
      for (int t = 0; t < TT; t++)
        if (t < T)
          s1[t] = mfma4(xv, fv[t], s1[t]);
    }
  }
  if (prev01 != 0xFFFFFFFFu)
  {
    close_01();
    close_0();
  }

  // the wavefronts' sums in wavefront order; lane (q = j1, g = j0, m = j2)
  red[wave * wave_size + lane] = acc;
  __syncthreads();
  const int64_t ybase = int64_t(row) * N;
  for (int e = threadIdx.x; e < N; e += W * wave_size)
  {
    const int j0 = e / 9, j1 = (e / 3) % 3, j2 = e % 3;
    const int src = 16 * j1 + 4 * j0 + j2;
    double tot = red[src];
    for (int w = 1; w < W; w++)
      tot += red[w * wave_size + src];
    store_row(a, ybase + e, tot);
  }
}

template<int TT>
void launch_mfma_terms(kron_args<double> const &a, int num_rows, int waves, hipStream_t stream)
{
  switch (waves)
  {
  case 1: kron_mfma_d3n3<1, TT><<<num_rows, 1 * wave_size, 0, stream>>>(a); break;
  case 2: kron_mfma_d3n3<2, TT><<<num_rows, 2 * wave_size, 0, stream>>>(a); break;
  case 8:
  case 16: kron_mfma_d3n3<8, TT><<<num_rows, 8 * wave_size, 0, stream>>>(a); break;
  default: kron_mfma_d3n3<4, TT><<<num_rows, 4 * wave_size, 0, stream>>>(a); break;
  }
}

// ---------------------------------------------------------- d = 6, n = 2 ---
//
// With n = 2 two dimensions together form a dense 4 x 4 factor, the
// Kronecker product of their 2 x 2 blocks: K45_t = A_t4 x A_t5 and so on.
// The row's columns are in lexicographic order of (c0, ..., c5), so per term
//   y = sum over (c0,c1) groups of K01 (sum over (c2,c3) groups of
//       K23 (sum over the group's columns of K45 x_c)),
// and three matrix-core layouts chain without moving data:
//
// stage 1, per column, v_mfma_f64_16x16x4_f64 (A 16 x 4 in lane i + 16 k,
//   B 4 x 16 in lane j + 16 k, D 16 x 16 with D[i][j] in lane j + 16 (i % 4),
//   item i / 4): A = the column's x with rows i = (i0,i1,i2,i3) and
//   k = (i4,i5); B = K45 of four terms side by side, j = (t', j4, j5).  The
//   16 x 16 result is summed over the columns sharing (c0, ..., c3); lane
//   (hi = (i2,i3), t', (j4,j5)) holds the four (i0,i1) items.
// stage 2, once per (c0..c3) group, v_mfma_f64_4x4x4_4b_f64 with the block
//   index g = t' (A[i][k] in lane 16 k + 4 g + i, B[k][j] in lane 16 k + 4 g
//   + j, D[i][j] in lane 16 i + 4 g + j): per item, D2 = K23_t D1, whose B
//   operand is the stage-1 item exactly where it sits.  Summed over the
//   groups sharing (c0, c1) in the accumulator; lane (hi = (j2,j3), t',
//   (j4,j5)).
// stage 3, once per (c0,c1) group, in each lane: the four (i0,i1) items
//   times K01_t as two 2 x 2 passes (VALU), summed over the groups.
// At the end the four terms of a lane quad-row are summed by a fixed xor
// tree, then the blocks of four terms, then the wavefronts in LDS in
// wavefront order.
//
// The factors of dimensions 2-5 change every column or every few columns;
// the workgroup stages them in LDS (every connect_1d connection of the row's
// 1D cells in those dimensions, in the [dim][connection][term][4] layout of
// vAk) with a table from a 1D index to its LDS offset, so the only per-column
// global loads are the x slices, XD columns of them in flight while a column
// is contracted.  Dimensions 0 and 1 change once per (c0,c1) group (about 30
// columns at d=6, |l|_1 <= 7): their blocks are read from global memory when
// the group opens and used when it closes.

typedef double mfma_d4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ mfma_d4 mfma16(double a, double b, mfma_d4 c)
{
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}


//! LDS bytes of kron_mfma_d6n2: 1D-index tables of the six dimensions, the
//! row's factor tables (dimensions 0..5 with ALL6, else 2..5), and the
//! wavefronts' row sums.
inline size_t mfma_d6n2_lds_bytes(kron_args<double> const &a, int W, bool all6)
{
  const size_t locb = ((size_t(6) * a.M * sizeof(int) + 15) / 16) * 16;
  const int64_t elems = all6 ? a.max_table_elems : a.max_table_last4;
  const size_t tab = ((size_t(elems) * sizeof(double) + 15) / 16) * 16;
  return locb + tab + size_t(W) * wave_size * sizeof(double);
}

#ifndef D6_XD8
#define D6_XD8 16
#endif

template<int W, int TG, int XD, bool ALL6>
__global__ void __launch_bounds__(W *wave_size) kron_mfma_d6n2(kron_args<double> a)
{
  constexpr int N = 64, nnp = 4;
  constexpr int K0 = ALL6 ? 0 : 2; // first dimension whose factors are staged in LDS
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int M = a.M;
  const int T = a.T;
  const int64_t tstride = int64_t(T) * nnp; // one connection in vAk
  int *loc = reinterpret_cast<int *>(smem_raw);
  double *fac = reinterpret_cast<double *>(smem_raw + ((6 * M * sizeof(int) + 15) / 16) * 16);
  double *red = fac + (((ALL6 ? a.max_table_elems : a.max_table_last4) + 1) / 2) * 2;

  const int tid = threadIdx.x;
  const int lane = tid & (wave_size - 1);
  // uniform per wavefront: loop bounds and group keys stay scalar
  const int wave = __builtin_amdgcn_readfirstlane(tid / wave_size);
  const int lo = lane & 15, hi = lane >> 4, tq = lo >> 2;

  const int row = a.row_order[blockIdx.x];
  const int64_t jbeg = a.pntr[row], jend = a.pntr[row + 1];
  const int64_t len = jend - jbeg;
  const int64_t cb = jbeg + (len * wave) / W, ce = jbeg + (len * (wave + 1)) / W;

  int rr[6];
#pragma unroll
  for (int k = 0; k < 6; k++)
    rr[k] = a.elem[int64_t(row) * 6 + k];
  // staged dimensions: first connect_1d connection of the row's 1D cell and
  // LDS offset of its table
  int cbk[6], fbk[6];
  {
    int base = 0;
#pragma unroll
    for (int k = K0; k < 6; k++)
    {
      cbk[k] = a.conn_pntr[rr[k]];
      fbk[k] = base;
      base += (a.conn_pntr[rr[k] + 1] - cbk[k]) * T * nnp;
    }
  }

  // loc[k][c]: for a dimension staged in LDS the LDS offset of the term-0
  // block of (rr[k], c) (0 if unconnected), otherwise its connect_1d position
  for (int e = tid; e < 6 * M; e += W * wave_size)
  {
    const int k = e / M, c = e - k * M;
    int v = 0;
#pragma unroll
    for (int q = 0; q < 6; q++)
      if (q == k)
      {
        const int off = a.off1d[rr[q] * M + c];
        if (q < K0)
          v = off;
        else
          v = (off < 0) ? 0 : (off - cbk[q]) * T * nnp + fbk[q];
      }
    loc[e] = v;
  }
#pragma unroll
  for (int k = K0; k < 6; k++)
  {
    const int len = (a.conn_pntr[rr[k] + 1] - cbk[k]) * T * nnp;
    const double *src = a.vAk + (int64_t(k) * a.nconn + cbk[k]) * tstride;
    for (int e = tid; e < len; e += W * wave_size)
      fac[fbk[k] + e] = src[e];
  }
  __syncthreads();

  // lane roles
  const int xoff = (lo << 2) | hi; // x[(i0,i1,i2,i3) = lo, (i4,i5) = hi]
  const int ehi = ((hi >> 1) << 1) | ((lo >> 1) & 1);        // dims 4 / 2 block entry
  const int elo = ((hi & 1) << 1) | (lo & 1);                // dims 5 / 3 block entry
  int tcl[TG];
  bool tok[TG];
#pragma unroll
  for (int b = 0; b < TG; b++)
  {
    tok[b] = (4 * b + tq) < T;
    tcl[b] = tok[b] ? 4 * b + tq : 0;
  }

  mfma_d4 d1[TG], acc2[TG];
  double acc3[TG][4];
  double a2v[TG], a3v[TG]; // dims 2, 3 block entries of the open (c0..c3) group
  int q0 = 0, q1 = 0;      // connect_1d positions of the open (c0,c1) group
#pragma unroll
  for (int b = 0; b < TG; b++)
  {
    d1[b] = mfma_d4{0, 0, 0, 0};
    acc2[b] = mfma_d4{0, 0, 0, 0};
    a2v[b] = a3v[b] = 0;
#pragma unroll
    for (int q = 0; q < 4; q++)
      acc3[b][q] = 0;
  }

  const double *__restrict__ vAk = a.vAk;
  auto block_of = [&](int k, int conn, int t) -> const double * {
    return vAk + (int64_t(k) * a.nconn + conn) * tstride + int64_t(t) * nnp;
  };
  // LDS factor reads of lane term t = 4 b + t': lanes of terms past T read
  // past their connection's T blocks, into the next connection's blocks or,
  // at the end of the tables, the row-sum area (at most 12 values, always
  // inside the workgroup's LDS); whatever they read stays in their own
  // lanes through every stage and is dropped at the end
  constexpr int bstride = 4 * nnp; // LDS distance between the blocks b, b + 1
  // the group's dims 2, 3 entries are read from LDS when it opens and used
  // when it closes; l23: the group's LDS offsets of dims 2 (low 16 bits) and 3
  auto open_23 = [&](unsigned int l23) {
    const double *p2 = fac + static_cast<int>(l23 & 0xFFFFu) + tq * nnp + ehi;
    const double *p3 = fac + static_cast<int>(l23 >> 16) + tq * nnp + elo;
#pragma unroll
    for (int b = 0; b < TG; b++)
    {
      a2v[b] = p2[b * bstride];
      a3v[b] = p3[b * bstride];
    }
  };
  // l01: the (c0,c1) group's dims 0 (low 16 bits) and 1 entries of loc
  auto open_01 = [&](unsigned int l01) {
    q0 = static_cast<int>(l01 & 0xFFFFu);
    q1 = static_cast<int>(l01 >> 16);
  };
  // the open (c0..c3) group: its stage-1 sums through K23 into acc2
  auto close_23 = [&]() {
#pragma unroll
    for (int b = 0; b < TG; b++)
    {
      const double k23 = a2v[b] * a3v[b];
      acc2[b][0] = mfma4(k23, d1[b][0], acc2[b][0]);
      acc2[b][1] = mfma4(k23, d1[b][1], acc2[b][1]);
      acc2[b][2] = mfma4(k23, d1[b][2], acc2[b][2]);
      acc2[b][3] = mfma4(k23, d1[b][3], acc2[b][3]);
      d1[b] = mfma_d4{0, 0, 0, 0};
    }
  };
  // the open (c0,c1) group: items (i0,i1) through K01 into acc3[(j0,j1)];
  // A(j, i) is entry i * 2 + j of a column-major 2 x 2 block
  auto close_01 = [&]() {
#pragma unroll
    for (int b = 0; b < TG; b++)
    {
      // dims 0, 1 blocks, once per (c0,c1) group: from LDS when staged
      // there, otherwise from global memory
      double k0r[4], k1r[4];
      // (global reads keep lanes past T on term 0, inside the array)
      const double *p0 = ALL6 ? fac + q0 + tq * nnp + b * bstride : block_of(0, q0, tcl[b]);
      const double *p1 = ALL6 ? fac + q1 + tq * nnp + b * bstride : block_of(1, q1, tcl[b]);
#pragma unroll
      for (int q = 0; q < 4; q++)
      {
        k0r[q] = p0[q];
        k1r[q] = p1[q];
      }
      double u[2][2]; // [j0][i1]
#pragma unroll
      for (int j0 = 0; j0 < 2; j0++)
#pragma unroll
        for (int i1 = 0; i1 < 2; i1++)
          u[j0][i1] = fma(k0r[2 + j0], acc2[b][2 + i1], k0r[j0] * acc2[b][i1]);
#pragma unroll
      for (int j0 = 0; j0 < 2; j0++)
#pragma unroll
        for (int j1 = 0; j1 < 2; j1++)
          acc3[b][2 * j0 + j1] += fma(k1r[2 + j1], u[j0][1], k1r[j1] * u[j0][0]);
      acc2[b] = mfma_d4{0, 0, 0, 0};
    }
  };

  constexpr unsigned long long mask01 = (1ull << 20) - 1, mask03 = (1ull << 40) - 1, none = ~0ull;
  // key of the column before the batch (none before the wavefront's first)
  unsigned int last_lo = 0xFFFFFFFFu, last_hi = 0xFFFFFFFFu;
  int opened = 0; // a (c0..c3) group is open
  for (int64_t j0 = cb; j0 < ce; j0 += wave_size)
  {
    // one batch of up to 64 columns, one per lane: id, the loc entries of
    // its six 1D indices (two per word), and in bits 30 and 31 of the
    // dims-4/5 word whether it opens a new (c0..c3) / (c0,c1) group
    const int64_t jl = j0 + lane;
    int colv = 0;
    unsigned int l01v = 0, l23v = 0, l45v = 0;
    unsigned long long wv = none;
    if (jl < ce)
    {
      colv = a.indx[jl];
      wv = a.epack[colv];
      l01v = static_cast<unsigned int>(loc[packed_1d(wv, 0)]) |
             (static_cast<unsigned int>(loc[M + packed_1d(wv, 1)]) << 16);
      l23v = static_cast<unsigned int>(loc[2 * M + packed_1d(wv, 2)]) |
             (static_cast<unsigned int>(loc[3 * M + packed_1d(wv, 3)]) << 16);
      l45v = static_cast<unsigned int>(loc[4 * M + packed_1d(wv, 4)]) |
             (static_cast<unsigned int>(loc[5 * M + packed_1d(wv, 5)]) << 15);
    }
    {
      const unsigned int wlo = static_cast<unsigned int>(wv), whi = static_cast<unsigned int>(wv >> 32);
      unsigned int plo = __shfl_up(wlo, 1, wave_size), phi = __shfl_up(whi, 1, wave_size);
      if (lane == 0)
      {
        plo = last_lo;
        phi = last_hi;
      }
      const unsigned long long pv = (static_cast<unsigned long long>(phi) << 32) | plo;
      const bool first = (pv == none);
      l45v |= ((first || ((wv ^ pv) & mask03) != 0) ? (1u << 30) : 0u) |
              ((first || ((wv ^ pv) & mask01) != 0) ? (1u << 31) : 0u);
      const int nb0 = static_cast<int>((ce - j0 < wave_size) ? ce - j0 : wave_size);
      last_lo = static_cast<unsigned int>(__builtin_amdgcn_readlane(wlo, nb0 - 1));
      last_hi = static_cast<unsigned int>(__builtin_amdgcn_readlane(whi, nb0 - 1));
    }
    const int nb = static_cast<int>((ce - j0 < wave_size) ? ce - j0 : wave_size);
    double xq[XD];
#pragma unroll
    for (int u = 0; u < XD; u++)
    {
      const int col = __builtin_amdgcn_readlane(colv, u < nb ? u : 0);
      xq[u] = a.x[int64_t(col) * N + xoff];
    }
    // dimension-4/5 entries of the next column, read one column ahead
    double f4n[TG], f5n[TG];
    // the dims-4/5 word of the next column (its factors are read one column
    // ahead; its group bits are used when it is processed)
    unsigned int l45n = 0;
    auto fetch_factors = [&](int jj) {
      l45n = static_cast<unsigned int>(__builtin_amdgcn_readlane(l45v, jj));
      const double *p4 = fac + static_cast<int>(l45n & 0x7FFFu) + tq * nnp + ehi;
      const double *p5 = fac + static_cast<int>((l45n >> 15) & 0x7FFFu) + tq * nnp + elo;
#pragma unroll
      for (int b = 0; b < TG; b++)
      {
        f4n[b] = p4[b * bstride];
        f5n[b] = p5[b * bstride];
      }
    };
    fetch_factors(0);
    // one column: group changes, then its stage-1 product; the next
    // column's dimension-4/5 entries go out before the group work
    auto column = [&](int jj, double xv) {
      double bv[TG];
#pragma unroll
      for (int b = 0; b < TG; b++)
        bv[b] = f4n[b] * f5n[b];
      const unsigned int flags = l45n >> 30; // this column's group bits
      fetch_factors((jj + 1 < nb) ? jj + 1 : jj);
      if (flags & 1u)
      {
        if (opened)
          close_23();
        if (flags & 2u)
        {
          if (opened)
            close_01();
          open_01(static_cast<unsigned int>(__builtin_amdgcn_readlane(l01v, jj)));
        }
        open_23(static_cast<unsigned int>(__builtin_amdgcn_readlane(l23v, jj)));
        opened = 1;
      }
#pragma unroll
      for (int b = 0; b < TG; b++)
        d1[b] = mfma16(xv, bv[b], d1[b]);
    };
    // full chunks of XD columns: every slot of the x ring is consumed and
    // refilled once per chunk, unconditionally (the batch's last chunk
    // re-reads its last column), so no branch joins an x load in flight
    int jb = 0;
    for (; jb + XD <= nb; jb += XD)
    {
#pragma unroll
      for (int u = 0; u < XD; u++)
      {
        const double xv = xq[u];
        const int nxt = (jb + u + XD < nb) ? jb + u + XD : nb - 1;
        const int col = __builtin_amdgcn_readlane(colv, nxt);
        xq[u] = a.x[int64_t(col) * N + xoff];
        column(jb + u, xv);
      }
    }
    // the last nb - jb (< XD) columns, already in the ring
#pragma unroll
    for (int u = 0; u < XD; u++)
      if (jb + u < nb)
        column(jb + u, xq[u]);
  }
  if (opened)
  {
    close_23();
    close_01();
  }

  // terms: blocks in order, then the four t' of a lane row by a fixed tree;
  // lanes with t' = 0 hold entry (j0,j1, j2,j3 = hi, j4,j5 = lane & 3).
  // Lanes of terms past T carried term 0's factors through every stage
  // (stage results never cross lanes of different t'); they are dropped here.
  double v[4];
#pragma unroll
  for (int q = 0; q < 4; q++)
  {
    double s = tok[0] ? acc3[0][q] : 0.0;
#pragma unroll
    for (int b = 1; b < TG; b++)
      s += tok[b] ? acc3[b][q] : 0.0;
    s += __shfl_xor(s, 4, wave_size);
    s += __shfl_xor(s, 8, wave_size);
    v[q] = s;
  }
  if (tq == 0)
  {
#pragma unroll
    for (int q = 0; q < 4; q++)
      red[wave * wave_size + q * 16 + hi * 4 + (lane & 3)] = v[q];
  }
  __syncthreads();
  const int64_t ybase = int64_t(row) * N;
  for (int e = tid; e < N; e += W * wave_size)
  {
    double tot = red[e];
    for (int w = 1; w < W; w++)
      tot += red[w * wave_size + e];
    store_row(a, ybase + e, tot);
  }
}

template<int TG, bool ALL6>
bool launch_mfma6_tables(kron_args<double> const &a, int num_rows, int waves,
                         hipStream_t stream)
{
  auto go = [&](auto kern, int W) -> bool {
    const size_t lds = mfma_d6n2_lds_bytes(a, W, ALL6);
    if (lds > 160 * 1024)
      return false;
    if (lds > 64 * 1024 &&
        hipFuncSetAttribute(reinterpret_cast<const void *>(kern),
                            hipFuncAttributeMaxDynamicSharedMemorySize,
                            static_cast<int>(lds)) != hipSuccess)
    {
      (void)hipGetLastError();
      return false;
    }
    kern<<<num_rows, W * wave_size, lds, stream>>>(a);
    return true;
  };
  switch (waves)
  {
  // XD x loads in flight per wavefront; 16 wavefronts keep 4 to fit 128 VGPRs
  case 1: return go(kron_mfma_d6n2<1, TG, 8, ALL6>, 1);
  case 2: return go(kron_mfma_d6n2<2, TG, 16, ALL6>, 2);
  case 8: return go(kron_mfma_d6n2<8, TG, D6_XD8, ALL6>, 8);
  case 16: return go(kron_mfma_d6n2<16, TG, 4, ALL6>, 16);
  default: return go(kron_mfma_d6n2<4, TG, 16, ALL6>, 4);
  }
}

//! All six dimensions' tables in LDS when they fit (no global loads but the
//! x slices), else dimensions 2..5.
template<int TG>
bool launch_mfma6_terms(kron_args<double> const &a, int num_rows, int waves, hipStream_t stream)
{
  const int W = (waves == 1 || waves == 2 || waves == 8 || waves == 16) ? waves : 4;
  if (mfma_d6n2_lds_bytes(a, W, true) <= 160 * 1024)
    return launch_mfma6_tables<TG, true>(a, num_rows, waves, stream);
  // dimensions 0, 1 then go by connect_1d position, packed in 16 bits
  if (a.nconn > 0xFFFF)
    return false;
  return launch_mfma6_tables<TG, false>(a, num_rows, waves, stream);
}

inline bool launch_mfma_d6n2(kron_args<double> const &a, int num_rows, int waves,
                             hipStream_t stream)
{
  if (a.T > mfma_max_terms || a.epack == nullptr || a.vAk == nullptr || a.M > 1024)
    return false;
  if (a.T <= 4)
    return launch_mfma6_terms<1>(a, num_rows, waves, stream);
  return launch_mfma6_terms<2>(a, num_rows, waves, stream);
}
// ---------------------------------------------------------- d = 4, n = 2 ---

//! Element of a 2 x 2 column-major block pair product for lane (q, m):
//! A(jhi, ihi) B(jlo, ilo) with (ihi, ilo) = bits of q, (jhi, jlo) = bits of m.
__device__ __forceinline__ int pair_hi(int q, int m) { return (q >> 1) * 2 + (m >> 1); }
__device__ __forceinline__ int pair_lo(int q, int m) { return (q & 1) * 2 + (m & 1); }
//
// A 16-entry tensor fills one block of the 4 x 4 x 4 product, so the four
// blocks carry four terms of the same column at once (term 4 p + g in block
// g for term group p).  Entry e = 4 m + q of the column sits in lane
// 16 q + 4 g + m of every block:
//   stage 1 (dims 2, 3):  A = x (i = (i0, i1) = m, k = (i2, i3) = q),
//                         B[k][j] = A2(j2, i2) A3(j3, i3) of the block's term
//                         -> D1[(i0, i1) = q][(j2, j3) = m]
//   stage 2 (dims 0, 1):  A[i][k] = A0(j0, i0) A1(j1, i1) (i = m, k = q),
//                         B = D1 -> D2[(j0, j1) = q][(j2, j3) = m]
// Stage 1 runs per column and is summed over the columns sharing (c0, c1);
// stage 2 runs once per such group.  The terms (blocks) are summed at the end.
template<int W, int TG>
__global__ void __launch_bounds__(W *wave_size) kron_mfma_d4n2(kron_args<double> a)
{
  constexpr int N = 16;
  __shared__ double red[W * wave_size];
  const int lane = threadIdx.x & (wave_size - 1);
  // uniform per wavefront: loop bounds and group keys stay scalar
  const int wave = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x) / wave_size);
  const int q = lane >> 4, g = (lane >> 2) & 3, m = lane & 3;
  const int T = a.T;

  const int row = a.row_order[blockIdx.x];
  const int64_t jbeg = a.pntr[row], jend = a.pntr[row + 1];
  const int64_t len = jend - jbeg;
  const int64_t cb = jbeg + (len * wave) / W, ce = jbeg + (len * (wave + 1)) / W;

  int rM[4];
#pragma unroll
  for (int k = 0; k < 4; k++)
    rM[k] = a.elem[int64_t(row) * 4 + k] * a.M;
  const int64_t blk = int64_t(T) * 16; // one connection: T terms x 4 dims x 4
  const int xoff = 4 * m + q;
  const int ehi = pair_hi(q, m), elo = pair_lo(q, m);
  // term of this lane's block in each term group (-1: no such term)
  int term[TG];
#pragma unroll
  for (int p = 0; p < TG; p++)
    term[p] = (4 * p + g < T) ? 4 * p + g : -1;

  double s1[TG], s2[TG], a2[TG], k01[TG];
#pragma unroll
  for (int p = 0; p < TG; p++)
    s1[p] = s2[p] = a2[p] = k01[p] = 0;

  auto block_of = [&](int k, int c) -> int64_t {
    return int64_t(a.off1d[rM[k] + c]) * blk + k * 4;
  };
  auto load_a2 = [&](int c2) {
    const int64_t b2 = block_of(2, c2);
#pragma unroll
    for (int p = 0; p < TG; p++)
      a2[p] = (term[p] >= 0) ? a.vA[b2 + term[p] * 16 + ehi] : 0.0;
  };
  auto load_k01 = [&](int c0, int c1) {
    const int64_t b0 = block_of(0, c0), b1 = block_of(1, c1);
#pragma unroll
    for (int p = 0; p < TG; p++)
      k01[p] = (term[p] >= 0)
                   ? a.vA[b0 + term[p] * 16 + ehi] * a.vA[b1 + term[p] * 16 + elo]
                   : 0.0;
  };
  auto close_2 = [&]() {
#pragma unroll
    for (int p = 0; p < TG; p++)
    {
      s2[p] = mfma4(k01[p], s1[p], s2[p]);
      s1[p] = 0;
    }
  };

  constexpr unsigned long long mask2 = (1ull << 20) - 1, mask3 = (1ull << 30) - 1, none = ~0ull;
  unsigned long long prev2 = none, prev3 = none;
  for (int64_t j0 = cb; j0 < ce; j0 += wave_size)
  {
    const int64_t jl = j0 + lane;
    int colv = 0, off3v = 0;
    unsigned long long wv = 0;
    if (jl < ce)
    {
      colv = a.indx[jl];
      wv = a.epack[colv];
      off3v = a.off1d[rM[3] + packed_1d(wv, 3)];
    }
    const unsigned int wlo = static_cast<unsigned int>(wv), whi = static_cast<unsigned int>(wv >> 32);
    const int nb = static_cast<int>((ce - j0 < wave_size) ? ce - j0 : wave_size);
    double xn = 0, fn[TG];
    auto fetch = [&](int jj, double &xv, double (&fv)[TG]) {
      const int col = __builtin_amdgcn_readlane(colv, jj);
      const int64_t b3 = int64_t(__builtin_amdgcn_readlane(off3v, jj)) * blk + 12;
      xv = a.x[int64_t(col) * N + xoff];
#pragma unroll
      for (int p = 0; p < TG; p++)
        fv[p] = (term[p] >= 0) ? a.vA[b3 + term[p] * 16 + elo] : 0.0;
    };
    fetch(0, xn, fn);
    for (int jj = 0; jj < nb; jj++)
    {
      const double xv = xn;
      double f3[TG];
#pragma unroll
      for (int p = 0; p < TG; p++)
        f3[p] = fn[p];
      if (jj + 1 < nb)
        fetch(jj + 1, xn, fn);
      const unsigned long long w =
          (static_cast<unsigned long long>(
               static_cast<unsigned int>(__builtin_amdgcn_readlane(whi, jj)))
           << 32) |
          static_cast<unsigned int>(__builtin_amdgcn_readlane(wlo, jj));
      const unsigned long long p3 = w & mask3;
      if (p3 != prev3)
      {
        const unsigned long long p2 = w & mask2;
        if (p2 != prev2)
        {
          if (prev2 != none)
            close_2();
          load_k01(packed_1d(w, 0), packed_1d(w, 1));
          prev2 = p2;
        }
        load_a2(packed_1d(w, 2));
        prev3 = p3;
      }
#pragma unroll
      for (int p = 0; p < TG; p++)
        s1[p] = mfma4(xv, a2[p] * f3[p], s1[p]);
    }
  }
  if (prev2 != none)
    close_2();

  // lane (q = (j0, j1), g = term, m = (j2, j3)) holds entry 4 q + m of its
  // blocks' terms; the terms are summed in block order
  double acc = 0;
#pragma unroll
  for (int p = 0; p < TG; p++)
    acc += s2[p];
  red[wave * wave_size + lane] = acc;
  __syncthreads();
  const int64_t ybase = int64_t(row) * N;
  for (int e = threadIdx.x; e < N; e += W * wave_size)
  {
    const int src = 16 * (e >> 2) + (e & 3);
    double tot = 0;
    for (int w = 0; w < W; w++)
      for (int gg = 0; gg < 4; gg++)
        tot += red[w * wave_size + src + 4 * gg];
    store_row(a, ybase + e, tot);
  }
}

template<int TG>
void launch_mfma4_terms(kron_args<double> const &a, int num_rows, int waves, hipStream_t stream)
{
  switch (waves)
  {
  case 1: kron_mfma_d4n2<1, TG><<<num_rows, 1 * wave_size, 0, stream>>>(a); break;
  case 2: kron_mfma_d4n2<2, TG><<<num_rows, 2 * wave_size, 0, stream>>>(a); break;
  case 8:
  case 16: kron_mfma_d4n2<8, TG><<<num_rows, 8 * wave_size, 0, stream>>>(a); break;
  default: kron_mfma_d4n2<4, TG><<<num_rows, 4 * wave_size, 0, stream>>>(a); break;
  }
}

inline bool launch_mfma_d4n2(kron_args<double> const &a, int num_rows, int waves,
                             hipStream_t stream)
{
  if (a.T > mfma_max_terms || a.epack == nullptr)
    return false;
  if (a.T <= 4)
    launch_mfma4_terms<1>(a, num_rows, waves, stream);
  else
    launch_mfma4_terms<2>(a, num_rows, waves, stream);
  return true;
}

//! Launches the sum-factorised matrix-core kernel; false when it does not
//! apply (other (d, n), more than mfma_max_terms terms, not sparse-grid).
inline bool launch_mfma_d3n3(kron_args<double> const &a, int num_rows, int waves,
                             hipStream_t stream)
{
  if (a.T > mfma_max_terms || a.epack == nullptr)
    return false;
  if (a.T <= 4)
    launch_mfma_terms<4>(a, num_rows, waves, stream);
  else
    launch_mfma_terms<mfma_max_terms>(a, num_rows, waves, stream);
  return true;
}

} // namespace akm
