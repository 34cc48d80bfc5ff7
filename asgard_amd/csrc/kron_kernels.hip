// Kronmult kernels for gfx950: the register-resident wavefront kernel for
// tensors that fit a lane group, and a block-wide LDS kernel for larger ones.
#include "kron_kernels.hpp"

#include "akm_internal.hpp"

namespace akm
{
// ------------------------------------------------------------- helpers ---

//! Relaxed device-scope load: bypasses the non-coherent per-CU cache when a
//! wavefront reads partial sums another CU has just published.
__device__ __forceinline__ double load_coherent(const double *p)
{
  unsigned long long v = __hip_atomic_load(reinterpret_cast<const unsigned long long *>(p),
                                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  return __longlong_as_double(static_cast<long long>(v));
}
__device__ __forceinline__ float load_coherent(const float *p)
{
  unsigned v = __hip_atomic_load(reinterpret_cast<const unsigned *>(p), __ATOMIC_RELAXED,
                                 __HIP_MEMORY_SCOPE_AGENT);
  return __uint_as_float(v);
}

//! Finishes a chunk: single-chunk rows store directly; otherwise the chunk
//! publishes its partial sums and the last chunk of the row to arrive adds
//! all partials in chunk order.  `vals[i]` holds entry tid + i*nthreads.
template<typename P, int EPT>
__device__ __forceinline__ void finish_chunk(kron_args<P> const &a, int chunk, int row,
                                             int N, int tid, int nthreads,
                                             P const (&vals)[EPT], int *s_flag)
{
  const int first = a.row_chunk_begin[row];
  const int nch = a.row_chunk_begin[row + 1] - first;
  const int64_t ybase = int64_t(row) * N;
  if (nch == 1)
  {
#pragma unroll
    for (int i = 0; i < EPT; i++)
    {
      int e = tid + i * nthreads;
      if (e < N)
        store_row(a, ybase + e, vals[i]);
    }
    return;
  }
  P *part = a.partial + int64_t(chunk) * N;
#pragma unroll
  for (int i = 0; i < EPT; i++)
  {
    int e = tid + i * nthreads;
    if (e < N)
      part[e] = vals[i];
  }
  __threadfence();
  __syncthreads();
  if (tid == 0)
    *s_flag = (atomicAdd(&a.counters[row], 1) == nch - 1);
  __syncthreads();
  if (*s_flag == 0)
    return;
  __threadfence();
#pragma unroll
  for (int i = 0; i < EPT; i++)
  {
    int e = tid + i * nthreads;
    if (e < N)
    {
      P tot = 0;
      for (int c = 0; c < nch; c++)
        tot += load_coherent(a.partial + int64_t(first + c) * N + e);
      store_row(a, ybase + e, tot);
    }
  }
  if (tid == 0)
    a.counters[row] = 0; // ready for the next launch
}

// ------------------------------------------------------ register stages ---

//! In-place contraction of one trailing dimension held in registers:
//! w viewed as (outer, n, inner), w[o, s, i] <- sum_j A(s, j) w[o, j, i].
template<typename P, int n, int R, int inner>
__device__ __forceinline__ void local_stage(P (&w)[R], const P *__restrict__ A)
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
        f[j] = w[(o * n + j) * inner + i];
#pragma unroll
      for (int s = 0; s < n; s++)
      {
        P v = a[s] * f[0];
#pragma unroll
        for (int j = 1; j < n; j++)
          v = fma(a[j * n + s], f[j], v);
        w[(o * n + s) * inner + i] = v;
      }
    }
}

template<typename P, int D, int n, int g, int R, int K>
__device__ __forceinline__ void local_stages(P (&w)[R], const P *const (&A)[D])
{
  if constexpr (K >= g)
  {
    local_stage<P, n, R, ipow<n, D - 1 - K>::value>(w, A[K]);
    local_stages<P, D, n, g, R, K - 1>(w, A);
  }
}

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

// ---------------------------------------------------- register kernel ---

template<typename P, int D, int n, int g, int MODE>
__global__ void __launch_bounds__(wave_size) kron_reg(kron_args<P> a)
{
  constexpr int N = ipow<n, D>::value;
  constexpr int G = ipow<n, g>::value;
  constexpr int R = N / G;
  constexpr int CPW = wave_size / G; // columns per wavefront step
  constexpr int LDN = N + 1;         // padded LDS row
  constexpr int nn = n * n;
  static_assert(G <= wave_size, "lane group larger than a wavefront");

  __shared__ P red[CPW * LDN];
  __shared__ int s_flag;

  const int lane = threadIdx.x;
  const int slot = lane / G;
  const int sub = lane - slot * G;
  const bool lane_used = slot < CPW;

  const int chunk = xcd_chunk(blockIdx.x, a.num_chunks);
  const int row = a.chunk_row[chunk];
  const int64_t jbeg = a.chunk_beg[chunk], jend = a.chunk_end[chunk];

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
  for (int64_t j0 = jbeg; j0 < jend; j0 += CPW)
  {
    const int64_t j = j0 + slot;
    if (lane_used && j < jend)
    {
      const int col = (MODE == mode_dense) ? int(j) : a.indx[j];
      const P *__restrict__ xc = a.x + int64_t(col) * N + sub * R;
      int64_t cbase[D];
      if constexpr (MODE == mode_grid)
      {
#pragma unroll
        for (int k = 0; k < D; k++)
        {
          const int off = a.off1d[r1d[k] + a.elem[int64_t(col) * D + k]];
          cbase[k] = int64_t(off) * (tstride * a.T) + k * nn;
        }
      }
      else if constexpr (MODE == mode_dense)
      {
        const int64_t vstride = int64_t(a.nb) * a.nb * nn;
#pragma unroll
        for (int k = 0; k < D; k++)
        {
          const int ix = a.elem[int64_t(col + a.col_offset) * D + k];
          cbase[k] = k * vstride + int64_t(nn) * (int64_t(ix) * a.nb + r1d[k]);
        }
      }
      for (int t = 0; t < a.T; t++)
      {
        const P *Ak[D];
        if constexpr (MODE == mode_grid)
        {
#pragma unroll
          for (int k = 0; k < D; k++)
            Ak[k] = a.vA + cbase[k] + t * tstride;
        }
        else if constexpr (MODE == mode_dense)
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
#pragma unroll
        for (int r = 0; r < R; r++)
          w[r] = xc[r];
        local_stages<P, D, n, g, R, D - 1>(w, Ak);
        if constexpr (g > 0)
          exchange_stages<P, D, n, g, R, g - 1>(w, Ak, lane, sub);
#pragma unroll
        for (int r = 0; r < R; r++)
          acc[r] += w[r];
      }
    }
  }

  // sum the column slots in a fixed order
  if (lane_used)
  {
#pragma unroll
    for (int r = 0; r < R; r++)
      red[slot * LDN + sub * R + r] = acc[r];
  }
  __syncthreads();
  constexpr int EPT = (N + wave_size - 1) / wave_size;
  P vals[EPT];
#pragma unroll
  for (int i = 0; i < EPT; i++)
  {
    const int e = lane + i * wave_size;
    P s = 0;
    if (e < N)
      for (int sl = 0; sl < CPW; sl++)
        s += red[sl * LDN + e];
    vals[i] = s;
  }
  finish_chunk<P, EPT>(a, chunk, row, N, lane, wave_size, vals, &s_flag);
}

// ------------------------------------------------------- generic kernel ---

constexpr int generic_threads = 256;
constexpr int generic_max_ept = 16; // N <= 4096

//! Block-wide kernel for any (d, n) with n^d <= 4096: the tensor of one
//! (column, term) lives in LDS (one buffer, updated in place stage by
//! stage), each thread owns the entries it writes.
template<typename P, int MODE>
__global__ void __launch_bounds__(generic_threads) kron_generic(kron_args<P> a, int D, int n)
{
  extern __shared__ unsigned char smem_raw[];
  __shared__ int s_flag;
  int N = 1;
  for (int k = 0; k < D; k++)
    N *= n;
  P *buf = reinterpret_cast<P *>(smem_raw);
  const int tid = threadIdx.x;
  const int nn = n * n;

  const int chunk = xcd_chunk(blockIdx.x, a.num_chunks);
  const int row = a.chunk_row[chunk];
  const int64_t jbeg = a.chunk_beg[chunk], jend = a.chunk_end[chunk];

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
  finish_chunk<P, generic_max_ept>(a, chunk, row, N, tid, generic_threads, acc, &s_flag);
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
template<typename P>
constexpr int lead_dims(int D, int n)
{
  const int budget = (sizeof(P) == 8) ? 32 : 64;
  int g = 0, R = 1;
  for (int k = 0; k < D; k++)
    R *= n;
  while (R > budget && g < D)
  {
    R /= n;
    g++;
  }
  return g;
}

template<typename P, int D, int n, int MODE>
bool launch_reg(kron_args<P> const &a, hipStream_t stream)
{
  constexpr int g = lead_dims<P>(D, n);
  if constexpr (ipow<n, g>::value > wave_size)
  {
    return false;
  }
  else
  {
    kron_reg<P, D, n, g, MODE><<<a.num_chunks, wave_size, 0, stream>>>(a);
    return true;
  }
}

template<typename P, int MODE>
bool dispatch_reg(int D, int n, kron_args<P> const &a, hipStream_t stream)
{
#define AKM_CASE(DD, NN)                                                                           \
  if (D == DD && n == NN)                                                                          \
    return launch_reg<P, DD, NN, MODE>(a, stream);
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
int launch_kronmult(int D, int n, int mode, kron_args<P> const &a, int force_generic,
                    hipStream_t stream)
{
  int N = 1;
  for (int k = 0; k < D; k++)
    N *= n;
  bool done = false;
  if (!force_generic)
  {
    if (mode == mode_ia)
      done = dispatch_reg<P, mode_ia>(D, n, a, stream);
    else if (mode == mode_grid)
      done = dispatch_reg<P, mode_grid>(D, n, a, stream);
    else
      done = dispatch_reg<P, mode_dense>(D, n, a, stream);
  }
  if (!done)
  {
    if (N > generic_threads * generic_max_ept)
      return set_error(AKM_ERR_UNSUPPORTED, "kronmult: tensor size n^d = %d exceeds 4096", N);
    const size_t lds = sizeof(P) * size_t(N);
    if (mode == mode_ia)
      kron_generic<P, mode_ia><<<a.num_chunks, generic_threads, lds, stream>>>(a, D, n);
    else if (mode == mode_grid)
      kron_generic<P, mode_grid><<<a.num_chunks, generic_threads, lds, stream>>>(a, D, n);
    else
      kron_generic<P, mode_dense><<<a.num_chunks, generic_threads, lds, stream>>>(a, D, n);
  }
  return check_launch("kronmult");
}

template<typename P>
int launch_scale(int64_t size, P beta, P *y, P gamma, const P *xg, hipStream_t stream)
{
  if (size == 0)
    return AKM_SUCCESS;
  int64_t blocks = (size + 255) / 256;
  if (blocks > 4096)
    blocks = 4096;
  kron_scale<P><<<int(blocks), 256, 0, stream>>>(size, beta, y, gamma, xg);
  return check_launch("kronmult scale");
}

template int launch_kronmult<double>(int, int, int, kron_args<double> const &, int, hipStream_t);
template int launch_kronmult<float>(int, int, int, kron_args<float> const &, int, hipStream_t);
template int launch_scale<double>(int64_t, double, double *, double, const double *,
                                  hipStream_t);
template int launch_scale<float>(int64_t, float, float *, float, const float *, hipStream_t);

} // namespace akm
