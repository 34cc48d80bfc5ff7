// Kronmult kernel templates (included by kron_f64.hip and kron_f32.hip, one
// translation unit per precision so the two compile in parallel).
#pragma once

#include "akm_internal.hpp"
#include "kron_kernels.hpp"

namespace akm
{
// ------------------------------------------------------ register stages ---

//! Contraction of a leading dimension spread over the lane group: the lane
//! with digit `digit` (stride `stride` lanes) gathers the n values of each
//! fibre from its group partners.
template<typename P, int n, int R, int stride>
__device__ __forceinline__ void exchange_stage(P (&w)[R], const P *__restrict__ A, int lane,
                                               int digit)
{
  P arow[n];
#pragma unroll
  for (int j = 0; j < n; j++)
    arow[j] = A[j * n + digit];
  if constexpr (n == 2 && stride == 1)
  {
    const P a_self = arow[digit], a_other = arow[1 - digit];
#pragma unroll
    for (int r = 0; r < R; r++)
    {
      P v = swap_pair(w[r]);
      w[r] = fma(a_other, v, a_self * w[r]);
    }
  }
  else
  {
    const int base = lane - digit * stride;
#pragma unroll
    for (int r = 0; r < R; r++)
    {
      P v = 0;
#pragma unroll
      for (int j = 0; j < n; j++)
        v = fma(arow[j], lane_read(w[r], base + j * stride), v);
      w[r] = v;
    }
  }
}

template<typename P, int D, int n, int g, int R, int K>
__device__ __forceinline__ void exchange_stages(P (&w)[R], const P *const (&A)[D], int lane,
                                                int sub)
{
  if constexpr (K >= 0)
  {
    constexpr int stride = ipow<n, g - 1 - K>::value;
    const int digit = (sub / stride) % n;
    exchange_stage<P, n, R, stride>(w, A[K], lane, digit);
    exchange_stages<P, D, n, g, R, K - 1>(w, A, lane, sub);
  }
}

//! One contraction stage from `src` into `dst` (which may be the same
//! array): dst[o, s, i] = sum_j A(s, j) src[o, j, i], or dst += ... when
//! ACCUM.  Every fibre is read before it is written, so in-place is safe.
template<typename P, int n, int R, int inner, bool ACCUM>
__device__ __forceinline__ void stage_into(const P (&src)[R], P (&dst)[R],
                                           const P *__restrict__ A)
{
  P a[n * n];
#pragma unroll
  for (int q = 0; q < n * n; q++)
    a[q] = A[q];
  constexpr int outer = R / (n * inner);
#pragma unroll
  for (int o = 0; o < outer; o++)
#pragma unroll
    for (int i = 0; i < inner; i++)
    {
      P f[n];
#pragma unroll
      for (int j = 0; j < n; j++)
        f[j] = src[(o * n + j) * inner + i];
#pragma unroll
      for (int s = 0; s < n; s++)
      {
        const int pos = (o * n + s) * inner + i;
        P v = ACCUM ? fma(a[s], f[0], dst[pos]) : a[s] * f[0];
#pragma unroll
        for (int j = 1; j < n; j++)
          v = fma(a[j * n + s], f[j], v);
        dst[pos] = v;
      }
    }
}

//! Local stages K = D-1 .. g of one (column, term): the first reads the
//! column tensor X (kept for the next term), the rest work in place on w,
//! and when no exchange stage follows (g == 0) the last one accumulates
//! straight into acc.
template<typename P, int D, int n, int g, int R, int K, bool FROM_W>
__device__ __forceinline__ void contract_local(const P (&X)[R], P (&w)[R], P (&acc)[R],
                                               const P *const (&A)[D])
{
  if constexpr (K >= g)
  {
    constexpr int inner = ipow<n, D - 1 - K>::value;
    constexpr bool first = (K == D - 1), last = (K == g), accum = last && (g == 0);
    if constexpr (first && accum)
    {
      if constexpr (FROM_W)
        stage_into<P, n, R, inner, true>(w, acc, A[K]);
      else
        stage_into<P, n, R, inner, true>(X, acc, A[K]);
    }
    else if constexpr (first && !FROM_W)
      stage_into<P, n, R, inner, false>(X, w, A[K]);
    else if constexpr (accum)
      stage_into<P, n, R, inner, true>(w, acc, A[K]);
    else
      stage_into<P, n, R, inner, false>(w, w, A[K]);
    contract_local<P, D, n, g, R, K - 1, FROM_W>(X, w, acc, A);
  }
}

//! Full contraction of one (column, term) into acc; the column tensor is
//! read from X (FROM_W = false) or from a copy already in w (FROM_W = true,
//! which frees X for the next column's prefetch).
template<typename P, int D, int n, int g, int R, bool FROM_W>
__device__ __forceinline__ void contract_term(const P (&X)[R], P (&w)[R], P (&acc)[R],
                                              const P *const (&A)[D], int lane, int sub)
{
  contract_local<P, D, n, g, R, D - 1, FROM_W>(X, w, acc, A);
  if constexpr (g > 0)
  {
    exchange_stages<P, D, n, g, R, g - 1>(w, A, lane, sub);
#pragma unroll
    for (int r = 0; r < R; r++)
      acc[r] += w[r];
  }
}

//! Lane-spread dimensions are exchanged through LDS (one transpose per
//! (column, term)) rather than by lane reads when n >= 3: a lane read moves
//! one value per instruction and n of them are needed per entry and stage.
template<int n, int g>
constexpr bool transposed_exchange()
{
  return g > 0 && n != 2;
}

//! Orders the LDS traffic of one wavefront around a cross-lane exchange.
__device__ __forceinline__ void wave_lds_sync()
{
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

//! Stages K .. 0 in the transposed layout, where the registers hold the
//! leading D - g dimensions (so dims 0 .. g-1 are local); the last one
//! (K == 0) accumulates into acc.
template<typename P, int D, int n, int g, int R, int K>
__device__ __forceinline__ void transposed_stages(P (&w)[R], P (&acc)[R],
                                                  const P *const (&A)[D])
{
  if constexpr (K >= 0)
  {
    constexpr int inner = ipow<n, D - g - 1 - K>::value;
    if constexpr (K == 0)
      stage_into<P, n, R, inner, true>(w, acc, A[0]);
    else
      stage_into<P, n, R, inner, false>(w, w, A[K]);
    transposed_stages<P, D, n, g, R, K - 1>(w, acc, A);
  }
}

//! Full contraction of one (column, term) into acc with the exchange done
//! as a transpose through the lane group's LDS slice tb (N entries): before
//! it lane `sub` holds entries sub*R + r (lane-spread dims leading), after
//! it entries r*G + sub (lane-spread dims trailing), which is also the
//! layout of acc.
template<typename P, int D, int n, int g, int R, bool FROM_W>
__device__ __forceinline__ void contract_term_t(const P (&X)[R], P (&w)[R], P (&acc)[R],
                                                const P *const (&A)[D], int sub, P *tb)
{
  constexpr int G = ipow<n, g>::value;
  static_assert(D - g >= g, "the transposed layout needs D - g >= g");
  contract_local<P, D, n, g, R, D - 1, FROM_W>(X, w, acc, A);
#pragma unroll
  for (int r = 0; r < R; r++)
    tb[sub * R + r] = w[r];
  wave_lds_sync();
#pragma unroll
  for (int r = 0; r < R; r++)
    w[r] = tb[r * G + sub];
  wave_lds_sync();
  transposed_stages<P, D, n, g, R, g - 1>(w, acc, A);
}

//! Position in the row tensor of accumulator entry r of lane `sub`.
template<int N, int G, int R, bool TR>
__device__ __forceinline__ int acc_index(int sub, int r)
{
  return TR ? r * G + sub : sub * R + r;
}

//! Column range [beg, end) of a row.
template<typename P, int MODE>
__device__ __forceinline__ void row_range(kron_args<P> const &a, int row, int64_t &beg,
                                          int64_t &end)
{
  if constexpr (MODE == mode_dense)
  {
    beg = 0;
    end = a.cols;
  }
  else
  {
    beg = a.pntr[row];
    end = a.pntr[row + 1];
  }
}

// ------------------------------------------------- row-end reduction ---

//! True when the column slots of a wavefront are the lane bits above the
//! lane-group bits (G a power of two, registers in the untransposed
//! layout), so that the slots can be summed by a tree of lane exchanges.
//! G <= 8 keeps CPW >= 8 >= W, so the W * N partial sums fit in the
//! CPW * (N + 1) reduction buffer.
template<int n, int g>
constexpr bool slot_tree_reduce()
{
  constexpr int G = ipow<n, g>::value;
  return !transposed_exchange<n, g>() && (G & (G - 1)) == 0 && G <= 8;
}

//! One level of the slot tree.  Lanes `lane` and `lane ^ MASK` own the same
//! accumulator entries [e, end) (stored from v[0]; slots at or past `end`
//! are padding and hold zero).  The lower lane keeps the first ceil(LEN/2)
//! slots, the upper lane the rest, each summed with the partner's copy;
//! after the last level (MASK < STOP) every lane hands its real entries to
//! `emit(entry, value)`.  The tree is fixed, so the sums are the same run to
//! run.
template<typename P, int LEN, int MASK, int STOP, typename F>
__device__ __forceinline__ void slot_tree(const P (&v)[LEN], int lane, int e, int end, F &&emit)
{
  if constexpr (MASK < STOP)
  {
#pragma unroll
    for (int i = 0; i < LEN; i++)
      if (e + i < end)
        emit(e + i, v[i]);
  }
  else
  {
    constexpr int H = (LEN + 1) / 2;
    const bool upper = (lane & MASK) != 0;
    P nv[H];
#pragma unroll
    for (int i = 0; i < H; i++)
    {
      const P lo = v[i];
      const P hi = (H + i < LEN) ? v[(H + i < LEN) ? H + i : 0] : P(0);
      const P recv = __shfl_xor(upper ? lo : hi, MASK, wave_size);
      nv[i] = (upper ? hi : lo) + recv;
    }
    const int ne = upper ? e + H : e;
    const int nend = upper ? end : (end < e + H ? end : e + H);
    slot_tree<P, H, MASK / 2, STOP>(nv, lane, ne, nend, emit);
  }
}

//! Sums the per-lane accumulators of a row over the column slots of every
//! wavefront and stores the row.  With the slot tree each wavefront sums its
//! slots in registers and writes N partial sums; otherwise the wavefronts
//! take turns writing their slots to LDS.  The wavefronts' partial sums are
//! then added in wavefront order.
template<typename P, int D, int n, int g, int W>
__device__ __forceinline__ void row_reduce_store(kron_args<P> const &a, int row,
                                                 P (&acc)[ipow<n, D - g>::value], P *red)
{
  constexpr int N = ipow<n, D>::value;
  constexpr int G = ipow<n, g>::value;
  constexpr int R = N / G;
  constexpr int CPW = wave_size / G;
  constexpr int LDN = N + 1;
  constexpr int nthreads = W * wave_size;
  constexpr int EPT = (N + nthreads - 1) / nthreads;
  constexpr bool TR = transposed_exchange<n, g>();
  const int tid = threadIdx.x;
  const int wave = tid / wave_size;
  const int lane = tid - wave * wave_size;
  const int slot = lane / G;
  const int sub = lane - slot * G;
  const int64_t ybase = int64_t(row) * N;
  if constexpr (slot_tree_reduce<n, g>())
  {
    P *mine = red + wave * N + sub * R;
    slot_tree<P, R, wave_size / 2, G>(acc, lane, 0, R, [&](int r, P v) { mine[r] = v; });
    __syncthreads();
    for (int e = tid; e < N; e += nthreads)
    {
      P tot = red[e];
      for (int wv = 1; wv < W; wv++)
        tot += red[wv * N + e];
      store_row(a, ybase + e, tot);
    }
  }
  else
  {
    P tot[EPT];
#pragma unroll
    for (int i = 0; i < EPT; i++)
      tot[i] = 0;
    for (int wv = 0; wv < W; wv++)
    {
      if (wave == wv && slot < CPW)
      {
#pragma unroll
        for (int r = 0; r < R; r++)
          red[slot * LDN + acc_index<N, G, R, TR>(sub, r)] = acc[r];
      }
      __syncthreads();
#pragma unroll
      for (int i = 0; i < EPT; i++)
      {
        const int e = tid + i * nthreads;
        if (e < N)
          for (int sl = 0; sl < CPW; sl++)
            tot[i] += red[sl * LDN + e];
      }
      __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < EPT; i++)
    {
      const int e = tid + i * nthreads;
      if (e < N)
        store_row(a, ybase + e, tot[i]);
    }
  }
}

// ---------------------------------------------------- register kernel ---

#ifndef AKM_WAVES_PER_EU
#define AKM_WAVES_PER_EU 1
#endif

//! One output row of the register kernel (the whole workgroup works on it);
//! `red` is the LDS buffer of the final reduction (CPW * (N + 1) entries).
template<typename P, int D, int n, int g, int MODE, int W>
__device__ __forceinline__ void kron_row_body(kron_args<P> const &a, int row, P *red, P *tbuf)
{
  constexpr int N = ipow<n, D>::value;
  constexpr int G = ipow<n, g>::value;
  constexpr int R = N / G;
  constexpr int CPW = wave_size / G; // columns per wavefront step
  constexpr int nn = n * n;
  static_assert(G <= wave_size, "lane group larger than a wavefront");

  constexpr bool TR = transposed_exchange<n, g>();
  const int tid = threadIdx.x;
  const int wave = tid / wave_size;
  const int lane = tid - wave * wave_size;
  const int slot = lane / G;
  const int sub = lane - slot * G;
  const bool lane_used = slot < CPW;
  P *tb = TR ? tbuf + (wave * CPW + slot) * N : nullptr;

  int64_t jbeg, jend;
  row_range<P, MODE>(a, row, jbeg, jend);

  P acc[R];
#pragma unroll
  for (int r = 0; r < R; r++)
    acc[r] = 0;

  int r1d[D];
  if constexpr (MODE == mode_grid)
  {
#pragma unroll
    for (int k = 0; k < D; k++)
      r1d[k] = a.elem[int64_t(row) * D + k] * a.M;
  }
  else if constexpr (MODE == mode_dense)
  {
#pragma unroll
    for (int k = 0; k < D; k++)
      r1d[k] = a.elem[int64_t(row + a.row_offset) * D + k];
  }

  const int64_t tstride = int64_t(D) * nn; // term stride inside a grid block
  if constexpr (MODE == mode_grid)
  {
    // Software pipeline over the wavefront's column steps: the column id two
    // steps ahead and the next column's 1D indices are fetched at the start
    // of a step, its connect_1d positions and x slice under the last term.
    const int64_t blk = tstride * a.T; // one connection's factor blocks
    const int64_t jstep = int64_t(W) * CPW;
    int64_t j = jbeg + int64_t(wave) * CPW + slot;
    bool valid = lane_used && j < jend && a.T > 0;
    P X[R];
    int64_t cbase[D];
    if (valid)
    {
      const int col = a.indx[j];
#pragma unroll
      for (int k = 0; k < D; k++)
        cbase[k] = int64_t(a.off1d[r1d[k] + a.elem[int64_t(col) * D + k]]) * blk + k * nn;
      const P *xc = a.x + int64_t(col) * N + sub * R;
#pragma unroll
      for (int r = 0; r < R; r++)
        X[r] = xc[r];
    }
    int col_n = 0;
    if (valid && j + jstep < jend)
      col_n = a.indx[j + jstep];
    for (int64_t j0 = jbeg + int64_t(wave) * CPW; j0 < jend; j0 += jstep)
    {
      const int64_t jn = j + jstep;
      const bool valid_next = valid && jn < jend;
      int en[D];
      if (valid_next)
      {
#pragma unroll
        for (int k = 0; k < D; k++)
          en[k] = a.elem[int64_t(col_n) * D + k];
      }
      int col_nn = 0;
      if (valid_next && jn + jstep < jend)
        col_nn = a.indx[jn + jstep];
      if (valid)
      {
        const P *Ak[D];
#pragma unroll
        for (int k = 0; k < D; k++)
          Ak[k] = a.vA + cbase[k];
        P w[R];
        for (int t = 0; t < a.T - 1; t++)
        {
          if constexpr (TR)
            contract_term_t<P, D, n, g, R, false>(X, w, acc, Ak, sub, tb);
          else
            contract_term<P, D, n, g, R, false>(X, w, acc, Ak, lane, sub);
#pragma unroll
          for (int k = 0; k < D; k++)
            Ak[k] += tstride;
        }
        // last term: contract from a copy in w while the next column's
        // factor offsets and x slice arrive
#pragma unroll
        for (int r = 0; r < R; r++)
          w[r] = X[r];
        if (valid_next)
        {
#pragma unroll
          for (int k = 0; k < D; k++)
            cbase[k] = int64_t(a.off1d[r1d[k] + en[k]]) * blk + k * nn;
          const P *xn = a.x + int64_t(col_n) * N + sub * R;
#pragma unroll
          for (int r = 0; r < R; r++)
            X[r] = xn[r];
        }
        if constexpr (TR)
          contract_term_t<P, D, n, g, R, true>(X, w, acc, Ak, sub, tb);
        else
          contract_term<P, D, n, g, R, true>(X, w, acc, Ak, lane, sub);
      }
      j = jn;
      valid = valid_next;
      col_n = col_nn;
    }
  }
  else
  {
    for (int64_t j0 = jbeg + int64_t(wave) * CPW; j0 < jend; j0 += int64_t(W) * CPW)
    {
      const int64_t j = j0 + slot;
      if (!(lane_used && j < jend))
        continue;
      const int col = (MODE == mode_dense) ? int(j) : a.indx[j];
      const P *__restrict__ xc = a.x + int64_t(col) * N + sub * R;
      int64_t cbase[D];
      if constexpr (MODE == mode_dense)
      {
        const int64_t vstride = int64_t(a.nb) * a.nb * nn;
#pragma unroll
        for (int k = 0; k < D; k++)
        {
          const int ix = a.elem[int64_t(col + a.col_offset) * D + k];
          cbase[k] = k * vstride + int64_t(nn) * (int64_t(ix) * a.nb + r1d[k]);
        }
      }
      P X[R];
#pragma unroll
      for (int r = 0; r < R; r++)
        X[r] = xc[r];
      for (int t = 0; t < a.T; t++)
      {
        const P *Ak[D];
        if constexpr (MODE == mode_dense)
        {
          const P *tp = a.terms[t];
#pragma unroll
          for (int k = 0; k < D; k++)
            Ak[k] = tp + cbase[k];
        }
        else
        {
          const int *ia = a.iA + (j * a.T + t) * D;
#pragma unroll
          for (int k = 0; k < D; k++)
            Ak[k] = a.vA + ia[k];
        }
        P w[R];
        if constexpr (TR)
          contract_term_t<P, D, n, g, R, false>(X, w, acc, Ak, sub, tb);
        else
          contract_term<P, D, n, g, R, false>(X, w, acc, Ak, lane, sub);
      }
    }
  }
  if constexpr (TR)
    __syncthreads(); // red may share LDS with the transpose slices
  row_reduce_store<P, D, n, g, W>(a, row, acc, red);
}

template<typename P, int D, int n, int g, int MODE, int W>
__global__ void __launch_bounds__(W *wave_size)
    __attribute__((amdgpu_waves_per_eu(AKM_WAVES_PER_EU))) kron_rows(kron_args<P> a)
{
  constexpr int N = ipow<n, D>::value;
  constexpr int CPW = wave_size / ipow<n, g>::value;
  if constexpr (transposed_exchange<n, g>())
  {
    // dynamic LDS: the wavefronts' transpose slices, then the reduction
    extern __shared__ __align__(16) unsigned char rows_lds[];
    P *buf = reinterpret_cast<P *>(rows_lds);
    kron_row_body<P, D, n, g, MODE, W>(a, a.row_order[blockIdx.x], buf, buf);
  }
  else
  {
    __shared__ P red[CPW * (N + 1)];
    kron_row_body<P, D, n, g, MODE, W>(a, a.row_order[blockIdx.x], red, nullptr);
  }
}

//! Dynamic LDS bytes of kron_rows / kron_rows_batch.
template<typename P, int D, int n, int g, int W>
constexpr size_t rows_lds_bytes()
{
  constexpr int N = ipow<n, D>::value;
  constexpr int CPW = wave_size / ipow<n, g>::value;
  constexpr size_t tr = size_t(W) * CPW * N, red = size_t(CPW) * (N + 1);
  return transposed_exchange<n, g>() ? sizeof(P) * (tr > red ? tr : red) : 0;
}

//! Launches with `lds` dynamic bytes, raising the kernel's limit above
//! 64 KB when needed; false if the device refuses.
template<typename K>
bool launch_with_lds(K kern, size_t lds)
{
  if (lds > 64 * 1024)
  {
    if (hipFuncSetAttribute(reinterpret_cast<const void *>(kern),
                            hipFuncAttributeMaxDynamicSharedMemorySize,
                            static_cast<int>(lds)) != hipSuccess)
    {
      (void)hipGetLastError();
      return false;
    }
  }
  return true;
}

// -------------------------------------------------- batched row kernel ---

//! Persistent kernel over the (system, row) work list of a batch of
//! sparse-grid operators with the same (d, n): workgroup b takes items b,
//! b + gridDim.x, ...  Phase 0 forms the residual of every active system,
//! V_s[0] = b_s - (I - dt_s A_s) x_s (V_s[0] holds b_s on entry); phase 1
//! applies the implicit operator to the current Krylov column j = *jcol,
//! V_s[j+1] = (I - dt_s A_s) V_s[j].
// Two wavefronts per SIMD: the persistent loop would otherwise take ~300
// VGPRs at d=5 n=3 (one wavefront per SIMD); the few spills this costs
// measured cheaper (batch of 64: 1.17 against 1.22 s).
template<int D, int n, int g, int W>
__global__ void __launch_bounds__(W *wave_size) __attribute__((amdgpu_waves_per_eu(2)))
kron_rows_batch(const kron_args<double> *args, const batch_system *sys, const int2 *work,
                int nwork, const int *jcol, const int *active, int phase)
{
  constexpr int N = ipow<n, D>::value;
  constexpr int CPW = wave_size / ipow<n, g>::value;
  constexpr bool TR = transposed_exchange<n, g>();
  extern __shared__ __align__(16) unsigned char batch_lds[];
  __shared__ double red_static[TR ? 1 : CPW * (N + 1)];
  double *red = TR ? reinterpret_cast<double *>(batch_lds) : red_static;
  double *tbuf = TR ? reinterpret_cast<double *>(batch_lds) : nullptr;
  for (int item = blockIdx.x; item < nwork; item += gridDim.x)
  {
    const int2 wk = work[item];
    if (active[wk.x] == 0)
      continue;
    kron_args<double> a = args[wk.x];
    const batch_system v = sys[wk.x];
    if (phase == 0)
    {
      a.x = v.X;
      a.y = v.V;
      a.xg = v.X;
      a.alpha = v.dt;
      a.beta = 1.0;
      a.gamma = -1.0;
    }
    else
    {
      const double *xj = v.V + int64_t(*jcol) * v.n;
      a.x = xj;
      a.y = v.V + int64_t(*jcol + 1) * v.n;
      a.xg = xj;
      a.alpha = -v.dt;
      a.beta = 0.0;
      a.gamma = 1.0;
    }
    kron_row_body<double, D, n, g, mode_grid, W>(a, wk.y, red, tbuf);
    __syncthreads();
  }
}

// ----------------------------------------- sparse-grid kernel, LDS tables ---

constexpr int epack_bits = 10; // bits per dimension in a packed cell word

//! The sparse-grid kernel with the row's factor tables in LDS.  At the start
//! the workgroup copies, for every dimension k, the factor blocks of all 1D
//! connections of the row's 1D cell (one contiguous range of the
//! [dim][connection][term][nnp] array) into LDS, and builds a per-dimension
//! table from a column's 1D index to the LDS offset of its term-0 block.
//! The main loop then reads factors from LDS only; the next column's x slice
//! is fetched while the current column's last term is contracted.
template<typename P, int D, int n, int g, int W>
__global__ void __launch_bounds__(W *wave_size) kron_grid_lds(kron_args<P> a)
{
  constexpr int N = ipow<n, D>::value;
  constexpr int G = ipow<n, g>::value;
  constexpr int R = N / G;
  constexpr int CPW = wave_size / G;
  constexpr int nnp = padded_block<P>(n);
  constexpr int nthreads = W * wave_size;
  static_assert(G <= wave_size, "lane group larger than a wavefront");

  constexpr bool TR = transposed_exchange<n, g>();
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int M = a.M;
  int *loc = reinterpret_cast<int *>(smem_raw);
  P *fac = reinterpret_cast<P *>(smem_raw + ((D * M * sizeof(int) + 15) / 16) * 16);
  P *red = fac; // reused once the main loop is done
  // transpose slices after the larger of the factor tables and the reduction
  const int64_t fac_elems =
      a.max_table_elems > int64_t(CPW) * (N + 1) ? a.max_table_elems : int64_t(CPW) * (N + 1);

  const int tid = threadIdx.x;
  const int wave = tid / wave_size;
  const int lane = tid - wave * wave_size;
  const int slot = lane / G;
  const int sub = lane - slot * G;
  const bool lane_used = slot < CPW;
  const int T = a.T;
  P *tb = TR ? fac + fac_elems + (wave * CPW + slot) * N : nullptr;

  const int row = a.row_order[blockIdx.x];
  const int64_t jbeg = a.pntr[row], jend = a.pntr[row + 1];

  // stage the row's factor tables and the 1D-index -> LDS-offset tables
  int rr[D], cb[D], fb[D];
  {
    int base = 0;
#pragma unroll
    for (int k = 0; k < D; k++)
    {
      rr[k] = a.elem[int64_t(row) * D + k];
      cb[k] = a.conn_pntr[rr[k]];
      fb[k] = base;
      base += (a.conn_pntr[rr[k] + 1] - cb[k]) * T * nnp;
    }
  }
  // the first column's loads go out before the tables are staged, so their
  // latency overlaps the staging
  const int64_t jstep = int64_t(W) * CPW;
  int64_t j = jbeg + int64_t(wave) * CPW + slot;
  bool valid = lane_used && j < jend;
  P X[R];
  unsigned long long cp = 0;
  if (valid)
  {
    const int col = a.indx[j];
    cp = a.epack[col];
    const P *xc = a.x + int64_t(col) * N + sub * R;
#pragma unroll
    for (int r = 0; r < R; r++)
      X[r] = xc[r];
  }

  for (int e = tid; e < D * M; e += nthreads)
  {
    const int k = e / M, c = e - k * M;
    int v = 0;
#pragma unroll
    for (int q = 0; q < D; q++)
      if (q == k)
      {
        const int off = a.off1d[rr[q] * M + c];
        v = (off < 0) ? 0 : (off - cb[q]) * T * nnp + fb[q];
      }
    loc[e] = v;
  }
#pragma unroll
  for (int k = 0; k < D; k++)
  {
    const int len = (a.conn_pntr[rr[k] + 1] - cb[k]) * T * nnp;
    const P *src = a.vAk + (int64_t(k) * a.nconn + cb[k]) * T * nnp;
    for (int e = tid; e < len; e += nthreads)
      fac[fb[k] + e] = src[e];
  }
  __syncthreads();

  P acc[R];
#pragma unroll
  for (int r = 0; r < R; r++)
    acc[r] = 0;
  for (int64_t j0 = jbeg + int64_t(wave) * CPW; j0 < jend; j0 += jstep)
  {
    const int64_t jn = j + jstep;
    const bool valid_next = lane_used && jn < jend;
    int col_next = 0;
    if (valid_next)
      col_next = a.indx[jn];
    if (valid)
    {
      const P *Ak[D];
#pragma unroll
      for (int k = 0; k < D; k++)
      {
        const int c1d = int((cp >> (epack_bits * k)) & ((1u << epack_bits) - 1u));
        Ak[k] = fac + loc[k * M + c1d];
      }
      unsigned long long cp_next = 0;
      P w[R];
      for (int t = 0; t < T - 1; t++)
      {
        if constexpr (TR)
          contract_term_t<P, D, n, g, R, false>(X, w, acc, Ak, sub, tb);
        else
          contract_term<P, D, n, g, R, false>(X, w, acc, Ak, lane, sub);
#pragma unroll
        for (int k = 0; k < D; k++)
          Ak[k] += nnp;
      }
      // last term: copy the column into w, fetch the next column into X
      // under the contraction
#pragma unroll
      for (int r = 0; r < R; r++)
        w[r] = X[r];
      if (valid_next)
      {
        cp_next = a.epack[col_next];
        const P *xn = a.x + int64_t(col_next) * N + sub * R;
#pragma unroll
        for (int r = 0; r < R; r++)
          X[r] = xn[r];
      }
      if constexpr (TR)
        contract_term_t<P, D, n, g, R, true>(X, w, acc, Ak, sub, tb);
      else
        contract_term<P, D, n, g, R, true>(X, w, acc, Ak, lane, sub);
      cp = cp_next;
    }
    j = jn;
    valid = valid_next;
  }
  __syncthreads(); // factor tables are dead from here on
  row_reduce_store<P, D, n, g, W>(a, row, acc, red);
}

// ------------------------------------------------------- generic kernel ---

constexpr int generic_threads = 256;
constexpr int generic_max_ept = 16; // N <= 4096

//! Block-wide kernel for any (d, n) with n^d <= 4096: the tensor of one
//! (column, term) lives in LDS (one buffer, updated in place stage by
//! stage), each thread owns the entries it writes.  One block per row.
template<typename P, int MODE>
__global__ void __launch_bounds__(generic_threads) kron_generic(kron_args<P> a, int D, int n)
{
  extern __shared__ unsigned char smem_raw[];
  int N = 1;
  for (int k = 0; k < D; k++)
    N *= n;
  P *buf = reinterpret_cast<P *>(smem_raw);
  const int tid = threadIdx.x;
  const int nn = n * n;

  const int row = a.row_order[blockIdx.x];
  int64_t jbeg, jend;
  row_range<P, MODE>(a, row, jbeg, jend);

  int st[6];
  {
    int s = 1;
    for (int k = D - 1; k >= 0; k--)
    {
      st[k] = s;
      s *= n;
    }
  }
  int r1d[6] = {0, 0, 0, 0, 0, 0};
  if (MODE == mode_grid)
    for (int k = 0; k < D; k++)
      r1d[k] = a.elem[int64_t(row) * D + k] * a.M;
  else if (MODE == mode_dense)
    for (int k = 0; k < D; k++)
      r1d[k] = a.elem[int64_t(row + a.row_offset) * D + k];

  P acc[generic_max_ept];
  for (int i = 0; i < generic_max_ept; i++)
    acc[i] = 0;

  const int64_t tstride = int64_t(D) * nn;
  for (int64_t j = jbeg; j < jend; j++)
  {
    const int col = (MODE == mode_dense) ? int(j) : a.indx[j];
    const P *xc = a.x + int64_t(col) * N;
    int64_t cbase[6];
    if (MODE == mode_grid)
      for (int k = 0; k < D; k++)
        cbase[k] = int64_t(a.off1d[r1d[k] + a.elem[int64_t(col) * D + k]]) * (tstride * a.T) +
                   k * nn;
    else if (MODE == mode_dense)
      for (int k = 0; k < D; k++)
        cbase[k] = k * (int64_t(a.nb) * a.nb * nn) +
                   int64_t(nn) * (int64_t(a.elem[int64_t(col + a.col_offset) * D + k]) * a.nb +
                                  r1d[k]);
    for (int t = 0; t < a.T; t++)
    {
      __syncthreads();
      for (int e = tid; e < N; e += generic_threads)
        buf[e] = xc[e];
      __syncthreads();
      for (int k = 0; k < D; k++)
      {
        const P *A;
        if (MODE == mode_grid)
          A = a.vA + cbase[k] + t * tstride;
        else if (MODE == mode_dense)
          A = a.terms[t] + cbase[k];
        else
          A = a.vA + a.iA[(j * a.T + t) * D + k];
        // every thread computes its entries into registers, then the
        // stage result replaces the tensor in place
        P v[generic_max_ept];
#pragma unroll
        for (int i = 0; i < generic_max_ept; i++)
        {
          const int e = tid + i * generic_threads;
          v[i] = 0;
          if (e < N)
          {
            const int s = (e / st[k]) % n;
            const int base = e - s * st[k];
            for (int q = 0; q < n; q++)
              v[i] = fma(A[q * n + s], buf[base + q * st[k]], v[i]);
          }
        }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < generic_max_ept; i++)
        {
          const int e = tid + i * generic_threads;
          if (e < N)
            buf[e] = v[i];
        }
        __syncthreads();
      }
#pragma unroll
      for (int i = 0; i < generic_max_ept; i++)
      {
        const int e = tid + i * generic_threads;
        if (e < N)
          acc[i] += buf[e];
      }
    }
  }
  const int64_t ybase = int64_t(row) * N;
#pragma unroll
  for (int i = 0; i < generic_max_ept; i++)
  {
    const int e = tid + i * generic_threads;
    if (e < N)
      store_row(a, ybase + e, acc[i]);
  }
}

// --------------------------------------------------------- scale kernel ---

//! y = beta*y + gamma*xg (operator with no terms, or alpha == 0).
template<typename P>
__global__ void __launch_bounds__(256) kron_scale(int64_t size, P beta, P *y, P gamma,
                                                  const P *xg)
{
  for (int64_t i = blockIdx.x * int64_t(256) + threadIdx.x; i < size;
       i += int64_t(gridDim.x) * 256)
  {
    P v = (beta == P{0}) ? P{0} : beta * y[i];
    if (gamma != P{0})
      v += gamma * xg[i];
    y[i] = v;
  }
}

// ------------------------------------------------------------- dispatch ---

//! Number of leading dimensions spread over the lane group: the smallest g
//! with n^(d-g) <= the register budget (32 FP64 / 64 FP32 entries per lane).
#ifndef AKM_F64_REG_ENTRIES
#define AKM_F64_REG_ENTRIES 32
#endif

template<typename P>
constexpr int lead_dims(int D, int n)
{
  const int budget = (sizeof(P) == 8) ? AKM_F64_REG_ENTRIES : 64;
  int g = 0, R = 1;
  for (int k = 0; k < D; k++)
    R *= n;
  // n >= 3 exchanges by transpose, which needs at least as many register
  // dimensions as lane-spread ones (D - g >= g)
  while (R > budget && g < D && (n == 2 || 2 * (g + 1) <= D))
  {
    R /= n;
    g++;
  }
  return g;
}

template<typename P, int D, int n, int MODE, int W>
bool launch_rows_w(kron_args<P> const &a, int num_rows, hipStream_t stream)
{
  constexpr int g = lead_dims<P>(D, n);
  if constexpr (ipow<n, g>::value > wave_size)
  {
    return false;
  }
  else
  {
    constexpr size_t lds = rows_lds_bytes<P, D, n, g, W>();
    auto kern = kron_rows<P, D, n, g, MODE, W>;
    if (!launch_with_lds(kern, lds))
      return false;
    kern<<<num_rows, W * wave_size, lds, stream>>>(a);
    return true;
  }
}

//! LDS bytes of kron_grid_lds for the matrix, or 0 if it does not fit.
template<typename P, int D, int n, int g, int W>
size_t grid_lds_bytes(kron_args<P> const &a)
{
  constexpr int N = ipow<n, D>::value;
  constexpr int CPW = wave_size / ipow<n, g>::value;
  const size_t locb = ((size_t(D) * a.M * sizeof(int) + 15) / 16) * 16;
  const size_t tab = size_t(a.max_table_elems) * sizeof(P);
  const size_t redb = size_t(CPW) * (N + 1) * sizeof(P);
  const size_t trb = transposed_exchange<n, g>() ? size_t(W) * CPW * N * sizeof(P) : 0;
  const size_t total = locb + (tab > redb ? tab : redb) + trb;
  return (a.vAk != nullptr && total <= size_t(a.lds_cap)) ? total : 0;
}

template<typename P, int D, int n, int W>
bool launch_grid_lds_w(kron_args<P> const &a, int num_rows, hipStream_t stream)
{
  constexpr int g = lead_dims<P>(D, n);
  if constexpr (ipow<n, g>::value > wave_size)
  {
    return false;
  }
  else
  {
    const size_t lds = grid_lds_bytes<P, D, n, g, W>(a);
    if (lds == 0)
      return false;
    auto kern = kron_grid_lds<P, D, n, g, W>;
    if (lds > 64 * 1024)
    {
      if (hipFuncSetAttribute(reinterpret_cast<const void *>(kern),
                              hipFuncAttributeMaxDynamicSharedMemorySize,
                              static_cast<int>(lds)) != hipSuccess)
      {
        (void)hipGetLastError();
        return false;
      }
    }
    kern<<<num_rows, W * wave_size, lds, stream>>>(a);
    return true;
  }
}

template<typename P, int D, int n, int MODE>
bool launch_rows(kron_args<P> const &a, int num_rows, int waves, hipStream_t stream)
{
  if constexpr (MODE == mode_grid)
  {
    if (a.lds_cap > 0)
    {
      bool done = false;
      switch (waves)
      {
      case 1: done = launch_grid_lds_w<P, D, n, 1>(a, num_rows, stream); break;
      case 2: done = launch_grid_lds_w<P, D, n, 2>(a, num_rows, stream); break;
      case 8: done = launch_grid_lds_w<P, D, n, 8>(a, num_rows, stream); break;
      default: done = launch_grid_lds_w<P, D, n, 4>(a, num_rows, stream); break;
      }
      if (done)
        return true;
    }
    // the performance path carries every workgroup size the tuner may pick
    switch (waves)
    {
    case 1: return launch_rows_w<P, D, n, MODE, 1>(a, num_rows, stream);
    case 2: return launch_rows_w<P, D, n, MODE, 2>(a, num_rows, stream);
    case 8: return launch_rows_w<P, D, n, MODE, 8>(a, num_rows, stream);
    default: return launch_rows_w<P, D, n, MODE, 4>(a, num_rows, stream);
    }
  }
  else
  {
    return launch_rows_w<P, D, n, MODE, 4>(a, num_rows, stream);
  }
}

template<typename P, int MODE>
bool dispatch_rows(int D, int n, kron_args<P> const &a, int num_rows, int waves,
                   hipStream_t stream)
{
#define AKM_CASE(DD, NN)                                                                           \
  if (D == DD && n == NN)                                                                          \
    return launch_rows<P, DD, NN, MODE>(a, num_rows, waves, stream);
  AKM_CASE(1, 1) AKM_CASE(1, 2) AKM_CASE(1, 3) AKM_CASE(1, 4) AKM_CASE(1, 5)
  AKM_CASE(2, 1) AKM_CASE(2, 2) AKM_CASE(2, 3) AKM_CASE(2, 4) AKM_CASE(2, 5)
  AKM_CASE(3, 1) AKM_CASE(3, 2) AKM_CASE(3, 3) AKM_CASE(3, 4) AKM_CASE(3, 5)
  AKM_CASE(4, 1) AKM_CASE(4, 2) AKM_CASE(4, 3) AKM_CASE(4, 4)
  AKM_CASE(5, 1) AKM_CASE(5, 2) AKM_CASE(5, 3)
  AKM_CASE(6, 1) AKM_CASE(6, 2) AKM_CASE(6, 3)
#undef AKM_CASE
  return false;
}

template<typename P>
int launch_kronmult_impl(int D, int n, int mode, kron_args<P> const &a, int num_rows,
                         int waves, int force_generic, hipStream_t stream)
{
  if (num_rows == 0)
    return AKM_SUCCESS;
  int N = 1;
  for (int k = 0; k < D; k++)
    N *= n;
  bool done = false;
  if (!force_generic)
  {
    if (mode == mode_ia)
      done = dispatch_rows<P, mode_ia>(D, n, a, num_rows, waves, stream);
    else if (mode == mode_grid)
      done = dispatch_rows<P, mode_grid>(D, n, a, num_rows, waves, stream);
    else
      done = dispatch_rows<P, mode_dense>(D, n, a, num_rows, waves, stream);
  }
  if (!done)
  {
    if (N > generic_threads * generic_max_ept)
      return set_error(AKM_ERR_UNSUPPORTED, "kronmult: tensor size n^d = %d exceeds 4096", N);
    const size_t lds = sizeof(P) * size_t(N);
    if (mode == mode_ia)
      kron_generic<P, mode_ia><<<num_rows, generic_threads, lds, stream>>>(a, D, n);
    else if (mode == mode_grid)
      kron_generic<P, mode_grid><<<num_rows, generic_threads, lds, stream>>>(a, D, n);
    else
      kron_generic<P, mode_dense><<<num_rows, generic_threads, lds, stream>>>(a, D, n);
  }
  return check_launch("kronmult");
}

template<int D, int n, int W>
bool launch_rows_batch_w(const kron_args<double> *args, const batch_system *sys, const int2 *work,
                         int nwork, int grid, const int *jcol, const int *active, int phase,
                         hipStream_t stream)
{
  constexpr int g = lead_dims<double>(D, n);
  if constexpr (ipow<n, g>::value > wave_size)
  {
    return false;
  }
  else
  {
    constexpr size_t lds = rows_lds_bytes<double, D, n, g, W>();
    auto kern = kron_rows_batch<D, n, g, W>;
    if (!launch_with_lds(kern, lds))
      return false;
    kern<<<grid, W * wave_size, lds, stream>>>(args, sys, work, nwork, jcol, active, phase);
    return true;
  }
}

template<int D, int n>
bool launch_rows_batch_dn(int waves, const kron_args<double> *args, const batch_system *sys,
                          const int2 *work, int nwork, int grid, const int *jcol,
                          const int *active, int phase, hipStream_t stream)
{
  switch (waves)
  {
  case 1: return launch_rows_batch_w<D, n, 1>(args, sys, work, nwork, grid, jcol, active, phase, stream);
  case 2: return launch_rows_batch_w<D, n, 2>(args, sys, work, nwork, grid, jcol, active, phase, stream);
  case 8: return launch_rows_batch_w<D, n, 8>(args, sys, work, nwork, grid, jcol, active, phase, stream);
  default: return launch_rows_batch_w<D, n, 4>(args, sys, work, nwork, grid, jcol, active, phase, stream);
  }
}

inline int launch_rows_batch_impl(int D, int n, int waves, const kron_args<double> *args,
                                  const batch_system *sys, const int2 *work, int nwork, int grid,
                                  const int *jcol, const int *active, int phase,
                                  hipStream_t stream)
{
  if (nwork == 0)
    return AKM_SUCCESS;
  bool done = false;
#define AKM_BCASE(DD, NN)                                                                          \
  if (!done && D == DD && n == NN)                                                                 \
    done = launch_rows_batch_dn<DD, NN>(waves, args, sys, work, nwork, grid, jcol, active, phase,  \
                                        stream);
  AKM_BCASE(1, 2) AKM_BCASE(1, 3) AKM_BCASE(1, 4)
  AKM_BCASE(2, 2) AKM_BCASE(2, 3) AKM_BCASE(2, 4)
  AKM_BCASE(3, 2) AKM_BCASE(3, 3) AKM_BCASE(3, 4)
  AKM_BCASE(4, 2) AKM_BCASE(4, 3)
  AKM_BCASE(5, 2) AKM_BCASE(5, 3)
  AKM_BCASE(6, 2)
#undef AKM_BCASE
  if (!done)
    return set_error(AKM_ERR_UNSUPPORTED, "batched kronmult: (d, n) = (%d, %d) not supported", D, n);
  return check_launch("batched kronmult");
}

template<typename P>
int launch_scale_impl(int64_t size, P beta, P *y, P gamma, const P *xg, hipStream_t stream)
{
  if (size == 0)
    return AKM_SUCCESS;
  int64_t blocks = (size + 255) / 256;
  if (blocks > 4096)
    blocks = 4096;
  kron_scale<P><<<int(blocks), 256, 0, stream>>>(size, beta, y, gamma, xg);
  return check_launch("kronmult scale");
}

} // namespace akm
