// Learned right preconditioner of the flexible GMRES: a per-cell MLP.
//
// For every grid cell c with residual block r_c (the cell's n^d
// coefficients) and norm s = |r_c|, the network reads
//   x = [ r_c / s , f_c , 0 pad ]            (K1 = roundup16(n^d + d) inputs)
// where f_c holds one feature per dimension: level + (position + 1/2) /
// cells-of-the-level, divided by max_level + 1.  Two hidden layers of H ReLU
// units and a linear output o (n^d values, padded to N3 = roundup16(n^d))
// give the correction
//   z_c = r_c + s * o(x)
// so an untrained network (zero output layer) is the identity.  Training
// fits o to (e_c - r_c) / s on (residual, correction) pairs of the plain
// solver, e = (I - dt A)^-1 r, with the mean squared error.
//
// Parameters live as one flat FP32 array
//   [ W1 (H x K1) | b1 (H) | W2 (H x H) | b2 (H) | W3 (N3 x H) | b3 (N3) ]
// with every W row-major (output x input).  Training (forward, backward,
// Adam) runs in FP32 on an LDS-tiled GEMM kernel with fused epilogues; the
// preconditioner itself runs on BF16 copies of W1..W3 with the MFMA
// v_mfma_f32_16x16x16_bf16 (FP32 accumulation), one wavefront per 16 cells,
// activations in LDS.
#include "precond.hpp"

#include <algorithm>
#include <cmath>
#include <memory>
#include <numeric>
#include <random>
#include <vector>

namespace akm
{
constexpr int mm_tile = 64;    // GEMM output tile (rows and columns)
constexpr int mm_k = 16;       // GEMM k step
constexpr int mm_threads = 256;

enum mm_epilogue
{
  ep_store = 0,     // C = A B
  ep_bias = 1,      // C = A B + bias[col]
  ep_bias_relu = 2, // C = max(A B + bias[col], 0)
  ep_mask = 3       // C = (A B) * (mask[row, col] > 0)
};

//! C[m, c] = sum_k A(m, k) B(k, c) with A(m, k) = A[m*sam + k*sak] and
//! B(k, c) = B[k*sbk + c*sbn], then the epilogue.  64 x 64 tile per block,
//! 4 x 4 outputs per thread, k in steps of 16 through LDS.
template<int EP>
__global__ void __launch_bounds__(mm_threads)
mlp_gemm(int M, int N, int K, const float *A, int64_t sam, int64_t sak, const float *B,
         int64_t sbk, int64_t sbn, float *C, int64_t ldc, const float *bias, const float *mask,
         int64_t ldm, int kchunk, int64_t zstride)
{
  __shared__ float As[mm_k][mm_tile + 4];
  __shared__ float Bs[mm_k][mm_tile + 4];
  const int t = threadIdx.x;
  const int m0 = blockIdx.y * mm_tile, c0 = blockIdx.x * mm_tile;
  const int tm = (t / 16) * 4, tc = (t % 16) * 4;
  // split K: block z sums k in [z*kchunk, min(K, (z+1)*kchunk)) into its own
  // output slice C + z*zstride
  const int kb = blockIdx.z * kchunk;
  const int ke = min(K, kb + kchunk);
  C += blockIdx.z * zstride;
  float acc[4][4] = {};
  for (int k0 = kb; k0 < ke; k0 += mm_k)
  {
#pragma unroll
    for (int q = 0; q < 4; q++)
    {
      const int e = t + mm_threads * q; // 0 .. 1023
      const int mm = e % mm_tile, kk = e / mm_tile;
      const int gm = m0 + mm, gk = k0 + kk;
      As[kk][mm] = (gm < M && gk < ke) ? A[gm * sam + gk * sak] : 0.0f;
      const int gc = c0 + mm;
      Bs[kk][mm] = (gc < N && gk < ke) ? B[gk * sbk + gc * sbn] : 0.0f;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < mm_k; kk++)
    {
      float a[4], b[4];
#pragma unroll
      for (int i = 0; i < 4; i++)
      {
        a[i] = As[kk][tm + i];
        b[i] = Bs[kk][tc + i];
      }
#pragma unroll
      for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++)
          acc[i][j] = fmaf(a[i], b[j], acc[i][j]);
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 4; i++)
  {
    const int gm = m0 + tm + i;
    if (gm >= M)
      continue;
#pragma unroll
    for (int j = 0; j < 4; j++)
    {
      const int gc = c0 + tc + j;
      if (gc >= N)
        continue;
      float v = acc[i][j];
      if constexpr (EP == ep_bias || EP == ep_bias_relu)
        v += bias[gc];
      if constexpr (EP == ep_bias_relu)
        v = fmaxf(v, 0.0f);
      if constexpr (EP == ep_mask)
        v = (mask[gm * ldm + gc] > 0.0f) ? v : 0.0f;
      C[gm * ldc + gc] = v;
    }
  }
}

//! dO = scale * (O - T) on the first N of N3 columns (0 on the padding);
//! partial[block] = sum of (O - T)^2 over the block's entries.
__global__ void __launch_bounds__(256)
mlp_out_grad(int64_t rows, int N, int N3, const float *O, const float *T, float scale, float *dO,
             double *partial)
{
  __shared__ double sh[256];
  double sq = 0;
  const int64_t total = rows * N3;
  for (int64_t e = blockIdx.x * 256ll + threadIdx.x; e < total; e += 256ll * gridDim.x)
  {
    const int col = static_cast<int>(e % N3);
    float g = 0.0f;
    if (col < N)
    {
      const float diff = O[e] - T[e];
      sq += double(diff) * double(diff);
      g = scale * diff;
    }
    dO[e] = g;
  }
  sh[threadIdx.x] = sq;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1)
  {
    if (threadIdx.x < s)
      sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0)
    partial[blockIdx.x] = sh[0];
}

//! out[0] (+)= sum of partial[0 .. count) (one block, fixed order).
__global__ void __launch_bounds__(256) mlp_sum(int count, const double *partial, double *out,
                                               int accumulate)
{
  __shared__ double sh[256];
  double s = 0;
  for (int i = threadIdx.x; i < count; i += 256)
    s += partial[i];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1)
  {
    if (threadIdx.x < k)
      sh[threadIdx.x] += sh[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0)
    out[0] = accumulate ? out[0] + sh[0] : sh[0];
}

//! partial[z*cols + c] = sum over rows r in chunk z of X[r*ld + c]
//! (block (x, z): 256 columns of chunk z).
__global__ void __launch_bounds__(256)
mlp_colsum_part(int64_t rows, int cols, const float *X, int64_t ld, int64_t chunk,
                float *partial)
{
  const int c = blockIdx.x * 256 + threadIdx.x;
  if (c >= cols)
    return;
  const int64_t r0 = blockIdx.z * chunk;
  const int64_t r1 = min(rows, r0 + chunk);
  float s = 0.0f;
  for (int64_t r = r0; r < r1; r++)
    s += X[r * ld + c];
  partial[int64_t(blockIdx.z) * cols + c] = s;
}

//! out[e] = sum over z < parts of partial[z*count + e] (fixed order).
__global__ void __launch_bounds__(256)
mlp_sum_parts(int64_t count, int parts, const float *partial, float *out)
{
  for (int64_t e = blockIdx.x * 256ll + threadIdx.x; e < count; e += 256ll * gridDim.x)
  {
    float s = 0.0f;
    for (int z = 0; z < parts; z++)
      s += partial[z * count + e];
    out[e] = s;
  }
}

//! Adam (Kingma & Ba, 2015): m = b1 m + (1-b1) g, v = b2 v + (1-b2) g^2,
//! p -= lr * (m c1) / (sqrt(v c2) + eps) with the bias corrections
//! c1 = 1/(1-b1^t), c2 = 1/(1-b2^t).
__global__ void __launch_bounds__(256)
mlp_adam(int64_t count, float *p, const float *g, float *m, float *v, float lr, float b1,
         float b2, float eps, float c1, float c2)
{
  for (int64_t i = blockIdx.x * 256ll + threadIdx.x; i < count; i += 256ll * gridDim.x)
  {
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi * c1) / (sqrtf(vi * c2) + eps);
  }
}

//! Samples of the pairs: for vector q and cell c (sample q*cells + c),
//! X row = [r/s, features, 0], T row = [(e - r)/s, 0], s = |r_c|.
__global__ void __launch_bounds__(256)
mlp_make_samples(int64_t nvec, int64_t cells, int N, int d, int K1, int N3, const double *R,
                 const double *E, const float *feat, float *X, float *T)
{
  const int64_t total = nvec * cells;
  for (int64_t q = blockIdx.x * 256ll + threadIdx.x; q < total; q += 256ll * gridDim.x)
  {
    const int64_t c = q % cells;
    const double *r = R + q * N;
    const double *e = E + q * N;
    double s = 0;
    for (int i = 0; i < N; i++)
      s = fma(r[i], r[i], s);
    s = sqrt(s);
    const double inv = (s > 0.0) ? 1.0 / s : 0.0;
    float *x = X + q * K1;
    float *y = T + q * N3;
    for (int i = 0; i < N; i++)
    {
      x[i] = static_cast<float>(r[i] * inv);
      y[i] = static_cast<float>((e[i] - r[i]) * inv);
    }
    for (int k = 0; k < d; k++)
      x[N + k] = feat[c * d + k];
    for (int i = N + d; i < K1; i++)
      x[i] = 0.0f;
    for (int i = N; i < N3; i++)
      y[i] = 0.0f;
  }
}

//! Rows perm[base .. base+rows) of X and T into the batch buffers.
__global__ void __launch_bounds__(256)
mlp_gather(int rows, const int64_t *perm, int64_t base, int K1, int N3, const float *X,
           const float *T, float *Xb, float *Tb)
{
  const int r = blockIdx.x;
  if (r >= rows)
    return;
  const int64_t src = perm[base + r];
  for (int i = threadIdx.x; i < K1; i += 256)
    Xb[int64_t(r) * K1 + i] = X[src * K1 + i];
  for (int i = threadIdx.x; i < N3; i += 256)
    Tb[int64_t(r) * N3 + i] = T[src * N3 + i];
}

//! FP32 -> BF16, round to nearest even (bias-and-truncate on the bit
//! pattern); NaN stays a quiet NaN.
__device__ __forceinline__ unsigned short to_bf16(float f)
{
  unsigned int u = __float_as_uint(f);
  if ((u & 0x7fffffffu) > 0x7f800000u)
    return 0x7fc0;
  u += 0x7fffu + ((u >> 16) & 1u);
  return static_cast<unsigned short>(u >> 16);
}

__global__ void __launch_bounds__(256)
mlp_to_bf16(int64_t count, const float *src, unsigned short *dst)
{
  for (int64_t i = blockIdx.x * 256ll + threadIdx.x; i < count; i += 256ll * gridDim.x)
    dst[i] = to_bf16(src[i]);
}

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int mlp_cells_per_wave = 16;

//! One dense layer of the 16 cells of a wavefront on the matrix cores:
//! out[m][j] = act(sum_k in[m][k] W[j][k] + b[j]) for j < J (J, K multiples
//! of 16), in/out BF16 rows in LDS.  v_mfma_f32_16x16x16_bf16 operands:
//! lane l holds A[l%16][4(l/16) .. +3] and B[4(l/16) .. +3][l%16], and gets
//! D[4(l/16) + i][l%16], i = 0..3.
__device__ __forceinline__ void mfma_layer(const unsigned short *in, int ldi, int K,
                                           const unsigned short *W, const float *bias, int J,
                                           unsigned short *out, int ldo)
{
  const int lane = threadIdx.x;
  const int row = lane % 16, quad = lane / 16;
  for (int j0 = 0; j0 < J; j0 += 16)
  {
    f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
    const unsigned short *wrow = W + int64_t(j0 + row) * K + 4 * quad;
    const unsigned short *arow = in + row * ldi + 4 * quad;
    for (int k0 = 0; k0 < K; k0 += 16)
    {
      const bf16x4 a = *reinterpret_cast<const bf16x4 *>(arow + k0);
      const bf16x4 b = *reinterpret_cast<const bf16x4 *>(wrow + k0);
      acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
    }
    const float bj = bias[j0 + row];
#pragma unroll
    for (int i = 0; i < 4; i++)
      out[(4 * quad + i) * ldo + j0 + row] = to_bf16(fmaxf(acc[i] + bj, 0.0f));
  }
}

//! z = v + s * o(x) per cell (see the file comment); one wavefront per 16
//! cells.  LDS: X [16][K1+4], H1 [16][H+4], H2 [16][H+4] (BF16), s [16].
__global__ void __launch_bounds__(64)
mlp_apply_bf16(int64_t cells, int N, int d, int K1, int H, int N3, const double *v, double *z,
               const float *feat, const unsigned short *W1, const float *b1,
               const unsigned short *W2, const float *b2, const unsigned short *W3,
               const float *b3)
{
  extern __shared__ unsigned char lds_raw[];
  const int ldx = K1 + 4, ldh = H + 4;
  float *snorm = reinterpret_cast<float *>(lds_raw);
  unsigned short *Xs = reinterpret_cast<unsigned short *>(lds_raw + 64);
  unsigned short *H1s = Xs + mlp_cells_per_wave * ldx;
  unsigned short *H2s = H1s + mlp_cells_per_wave * ldh;
  const int lane = threadIdx.x;
  const int64_t c0 = int64_t(blockIdx.x) * mlp_cells_per_wave;

  // cell norms: 4 lanes per cell
  {
    const int m = lane / 4, sub = lane % 4;
    const int64_t c = c0 + m;
    double sq = 0;
    if (c < cells)
      for (int i = sub; i < N; i += 4)
      {
        const double vi = v[c * N + i];
        sq = fma(vi, vi, sq);
      }
    sq += __shfl_xor(sq, 1);
    sq += __shfl_xor(sq, 2);
    if (sub == 0)
      snorm[m] = static_cast<float>(sqrt(sq));
  }
  __syncthreads();
  // inputs
  for (int e = lane; e < mlp_cells_per_wave * K1; e += 64)
  {
    const int m = e / K1, i = e % K1;
    const int64_t c = c0 + m;
    float x = 0.0f;
    if (c < cells)
    {
      if (i < N)
      {
        const float s = snorm[m];
        x = (s > 0.0f) ? static_cast<float>(v[c * N + i] / double(s)) : 0.0f;
      }
      else if (i < N + d)
        x = feat[c * d + (i - N)];
    }
    Xs[m * ldx + i] = to_bf16(x);
  }
  __syncthreads();
  mfma_layer(Xs, ldx, K1, W1, b1, H, H1s, ldh);
  __syncthreads();
  mfma_layer(H1s, ldh, H, W2, b2, H, H2s, ldh);
  __syncthreads();
  // output layer, linear, straight to z
  const int row = lane % 16, quad = lane / 16;
  for (int j0 = 0; j0 < N3; j0 += 16)
  {
    f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
    const unsigned short *wrow = W3 + int64_t(j0 + row) * H + 4 * quad;
    const unsigned short *arow = H2s + row * ldh + 4 * quad;
    for (int k0 = 0; k0 < H; k0 += 16)
    {
      const bf16x4 a = *reinterpret_cast<const bf16x4 *>(arow + k0);
      const bf16x4 b = *reinterpret_cast<const bf16x4 *>(wrow + k0);
      acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
    }
    const int j = j0 + row;
    if (j >= N)
      continue;
    const float bj = b3[j];
#pragma unroll
    for (int i = 0; i < 4; i++)
    {
      const int m = 4 * quad + i;
      const int64_t c = c0 + m;
      if (c < cells)
        z[c * N + j] = v[c * N + j] + double(snorm[m]) * double(acc[i] + bj);
    }
  }
}

static int64_t round16(int64_t v) { return (v + 15) / 16 * 16; }

static int grid_for(int64_t count)
{
  return static_cast<int>(std::max<int64_t>(1, std::min<int64_t>((count + 255) / 256, 4096)));
}

} // namespace akm

using namespace akm;

struct akm_mlp_s final : akm_precond_s
{
  int d = 0, n = 0, N = 0, N3 = 0, K1 = 0, H = 0;
  int64_t cells = 0;
  int max_level = 0;
  int64_t oW1 = 0, ob1 = 0, oW2 = 0, ob2 = 0, oW3 = 0, ob3 = 0, num_params = 0;
  dev_buffer<float> params, grad, adam_m, adam_v, feat;
  dev_buffer<unsigned short> wbf; // W1 | W2 | W3 in BF16
  bool bf16_current = false;
  int64_t adam_t = 0;
  // batch buffers (sized for `batch` rows)
  int batch = 0;
  dev_buffer<float> Xb, Tb, H1, H2, O, dO, dH2, dH1;
  dev_buffer<float> scratch; // split-K partial products
  static constexpr int split_parts = 16;
  dev_buffer<double> partial, loss;
  hipStream_t stream = nullptr;

  int64_t size() const override { return cells * N; }

  int ensure_batch(int B)
  {
    if (B <= batch)
      return AKM_SUCCESS;
    batch = B;
    AKM_TRY(Xb.allocate(int64_t(B) * K1));
    AKM_TRY(Tb.allocate(int64_t(B) * N3));
    AKM_TRY(H1.allocate(int64_t(B) * H));
    AKM_TRY(H2.allocate(int64_t(B) * H));
    AKM_TRY(O.allocate(int64_t(B) * N3));
    AKM_TRY(dO.allocate(int64_t(B) * N3));
    AKM_TRY(dH2.allocate(int64_t(B) * H));
    AKM_TRY(dH1.allocate(int64_t(B) * H));
    AKM_TRY(partial.allocate(4096));
    AKM_TRY(loss.allocate(2));
    const int64_t widest = std::max<int64_t>({int64_t(H) * K1, int64_t(H) * H, int64_t(N3) * H});
    AKM_TRY(scratch.allocate(split_parts * widest));
    return AKM_SUCCESS;
  }

  template<int EP>
  int gemm(hipStream_t s, int M, int Nc, int K, const float *A, int64_t sam, int64_t sak,
           const float *B, int64_t sbk, int64_t sbn, float *C, int64_t ldc,
           const float *bias = nullptr, const float *mask = nullptr, int64_t ldm = 0)
  {
    dim3 grid((Nc + mm_tile - 1) / mm_tile, (M + mm_tile - 1) / mm_tile, 1);
    mlp_gemm<EP><<<grid, mm_threads, 0, s>>>(M, Nc, K, A, sam, sak, B, sbk, sbn, C, ldc, bias,
                                              mask, ldm, K, 0);
    return check_launch("mlp gemm");
  }

  //! C (M x Nc, ldc = Nc) = A B for a long K (the batch): K is split over
  //! blocks into partial products that are then summed in a fixed order.
  int gemm_long_k(hipStream_t s, int M, int Nc, int K, const float *A, int64_t sam, int64_t sak,
                  const float *B, int64_t sbk, int64_t sbn, float *C)
  {
    const int parts = std::max(1, std::min(split_parts, (K + 127) / 128));
    const int kchunk = (K + parts - 1) / parts;
    const int64_t zstride = int64_t(M) * Nc;
    dim3 grid((Nc + mm_tile - 1) / mm_tile, (M + mm_tile - 1) / mm_tile, parts);
    mlp_gemm<ep_store><<<grid, mm_threads, 0, s>>>(M, Nc, K, A, sam, sak, B, sbk, sbn,
                                                    scratch.data(), Nc, nullptr, nullptr, 0,
                                                    kchunk, zstride);
    mlp_sum_parts<<<grid_for(zstride), 256, 0, s>>>(zstride, parts, scratch.data(), C);
    return check_launch("mlp gemm (split k)");
  }

  //! out[c] = sum over the B rows of X[r*ld + c].
  int colsum(hipStream_t s, int64_t rows, int cols, const float *X, int64_t ld, float *out)
  {
    const int parts = static_cast<int>(std::max<int64_t>(1, std::min<int64_t>(split_parts, (rows + 63) / 64)));
    const int64_t chunk = (rows + parts - 1) / parts;
    mlp_colsum_part<<<dim3((cols + 255) / 256, 1, parts), 256, 0, s>>>(rows, cols, X, ld, chunk,
                                                                      scratch.data());
    mlp_sum_parts<<<grid_for(cols), 256, 0, s>>>(cols, parts, scratch.data(), out);
    return check_launch("mlp column sums");
  }

  //! FP32 forward of B rows of Xb into H1, H2, O.
  int forward(hipStream_t s, int B, const float *X)
  {
    const float *p = params.data();
    AKM_TRY(gemm<ep_bias_relu>(s, B, H, K1, X, K1, 1, p + oW1, 1, K1, H1.data(), H, p + ob1));
    AKM_TRY(gemm<ep_bias_relu>(s, B, H, H, H1.data(), H, 1, p + oW2, 1, H, H2.data(), H,
                               p + ob2));
    AKM_TRY(gemm<ep_bias>(s, B, N3, H, H2.data(), H, 1, p + oW3, 1, H, O.data(), N3, p + ob3));
    return AKM_SUCCESS;
  }

  //! Loss (mean squared error over B rows and N outputs) into loss[slot]
  //! (accumulated if `accumulate`) and the gradient of every parameter into
  //! grad, for the rows of X / T (forward must have run on X).
  int backward(hipStream_t s, int B, const float *X, const float *T, int accumulate)
  {
    const float *p = params.data();
    float *g = grad.data();
    const float scale = 2.0f / (float(B) * float(N));
    const int blocks = std::min(4096, grid_for(int64_t(B) * N3));
    mlp_out_grad<<<blocks, 256, 0, s>>>(B, N, N3, O.data(), T, scale, dO.data(), partial.data());
    mlp_sum<<<1, 256, 0, s>>>(blocks, partial.data(), loss.data(), accumulate);
    AKM_TRY(check_launch("mlp loss"));
    // layer 3
    AKM_TRY(gemm_long_k(s, N3, H, B, dO.data(), 1, N3, H2.data(), H, 1, g + oW3));
    AKM_TRY(colsum(s, B, N3, dO.data(), N3, g + ob3));
    AKM_TRY(gemm<ep_mask>(s, B, H, N3, dO.data(), N3, 1, p + oW3, H, 1, dH2.data(), H, nullptr,
                          H2.data(), H));
    // layer 2
    AKM_TRY(gemm_long_k(s, H, H, B, dH2.data(), 1, H, H1.data(), H, 1, g + oW2));
    AKM_TRY(colsum(s, B, H, dH2.data(), H, g + ob2));
    AKM_TRY(gemm<ep_mask>(s, B, H, H, dH2.data(), H, 1, p + oW2, H, 1, dH1.data(), H, nullptr,
                          H1.data(), H));
    // layer 1
    AKM_TRY(gemm_long_k(s, H, K1, B, dH1.data(), 1, H, X, K1, 1, g + oW1));
    AKM_TRY(colsum(s, B, H, dH1.data(), H, g + ob1));
    return check_launch("mlp backward");
  }

  int adam(hipStream_t s, const float *g, float lr)
  {
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    ++adam_t;
    const float c1 = static_cast<float>(1.0 / (1.0 - std::pow(double(b1), double(adam_t))));
    const float c2 = static_cast<float>(1.0 / (1.0 - std::pow(double(b2), double(adam_t))));
    mlp_adam<<<grid_for(num_params), 256, 0, s>>>(num_params, params.data(), g, adam_m.data(),
                                                  adam_v.data(), lr, b1, b2, eps, c1, c2);
    bf16_current = false;
    return check_launch("mlp adam");
  }

  int refresh_bf16(hipStream_t s)
  {
    if (bf16_current)
      return AKM_SUCCESS;
    const float *p = params.data();
    unsigned short *w = wbf.data();
    const int64_t n1 = int64_t(H) * K1, n2 = int64_t(H) * H, n3 = int64_t(N3) * H;
    mlp_to_bf16<<<grid_for(n1), 256, 0, s>>>(n1, p + oW1, w);
    mlp_to_bf16<<<grid_for(n2), 256, 0, s>>>(n2, p + oW2, w + n1);
    mlp_to_bf16<<<grid_for(n3), 256, 0, s>>>(n3, p + oW3, w + n1 + n2);
    AKM_TRY(check_launch("mlp bf16 weights"));
    bf16_current = true;
    return AKM_SUCCESS;
  }

  size_t apply_lds_bytes() const
  {
    return 64 + sizeof(unsigned short) * mlp_cells_per_wave * size_t((K1 + 4) + 2 * (H + 4));
  }

  int apply(const double *v, double *z, hipStream_t s) override
  {
    AKM_TRY(refresh_bf16(s));
    const int64_t blocks = (cells + mlp_cells_per_wave - 1) / mlp_cells_per_wave;
    const unsigned short *w = wbf.data();
    const int64_t n1 = int64_t(H) * K1, n2 = int64_t(H) * H;
    const float *p = params.data();
    mlp_apply_bf16<<<static_cast<unsigned>(blocks), 64, apply_lds_bytes(), s>>>(
        cells, N, d, K1, H, N3, v, z, feat.data(), w, p + ob1, w + n1, p + ob2, w + n1 + n2,
        p + ob3);
    return check_launch("mlp preconditioner");
  }
};

static akm_mlp_s *as_mlp(akm_precond p) { return dynamic_cast<akm_mlp_s *>(p); }

extern "C" {

akm_status akm_mlp_create(int d, int n, int64_t num_cells, const int *table, int max_level,
                          int hidden, uint64_t seed, akm_precond *out)
{
  if (out == nullptr || table == nullptr)
    return set_error(AKM_ERR_INVALID, "mlp: null argument");
  *out = nullptr;
  if (d < 1 || d > 6 || n < 1 || num_cells < 1 || max_level < 0 || max_level > 30)
    return set_error(AKM_ERR_INVALID, "mlp: invalid grid (d %d, n %d, cells %lld, level %d)", d,
                     n, static_cast<long long>(num_cells), max_level);
  if (hidden <= 0)
    hidden = 256;
  if (hidden % 16 != 0 || hidden > 512)
    return set_error(AKM_ERR_INVALID, "mlp: hidden width %d must be a multiple of 16, <= 512",
                     hidden);
  int64_t N = 1;
  for (int k = 0; k < d; k++)
    N *= n;
  if (N > 256)
    return set_error(AKM_ERR_UNSUPPORTED, "mlp: n^d = %lld above 256", static_cast<long long>(N));
  auto m = std::make_unique<akm_mlp_s>();
  m->d = d;
  m->n = n;
  m->N = static_cast<int>(N);
  m->N3 = static_cast<int>(round16(N));
  m->K1 = static_cast<int>(round16(N + d));
  m->H = hidden;
  m->cells = num_cells;
  m->max_level = max_level;
  const int64_t K1 = m->K1, H = m->H, N3 = m->N3;
  m->oW1 = 0;
  m->ob1 = H * K1;
  m->oW2 = m->ob1 + H;
  m->ob2 = m->oW2 + H * H;
  m->oW3 = m->ob2 + H;
  m->ob3 = m->oW3 + N3 * H;
  m->num_params = m->ob3 + N3;
  // He-normal hidden layers, zero biases, zero output layer (identity map)
  std::vector<float> p(m->num_params, 0.0f);
  std::mt19937_64 rng(seed);
  std::normal_distribution<float> nd(0.0f, 1.0f);
  const float s1 = std::sqrt(2.0f / float(m->N + d)), s2 = std::sqrt(2.0f / float(H));
  for (int64_t j = 0; j < H; j++)
    for (int64_t i = 0; i < m->N + d; i++)
      p[m->oW1 + j * K1 + i] = s1 * nd(rng);
  for (int64_t j = 0; j < H * H; j++)
    p[m->oW2 + j] = s2 * nd(rng);
  AKM_TRY(m->params.upload(p.data(), m->num_params));
  std::vector<float> zeros(m->num_params, 0.0f);
  AKM_TRY(m->adam_m.upload(zeros.data(), m->num_params));
  AKM_TRY(m->adam_v.upload(zeros.data(), m->num_params));
  AKM_TRY(m->grad.upload(zeros.data(), m->num_params));
  AKM_TRY(m->wbf.allocate(H * K1 + H * H + N3 * H));
  // features: (level + (position + 1/2) / cells of the level) / (max_level + 1)
  std::vector<float> f(num_cells * d);
  for (int64_t c = 0; c < num_cells; c++)
    for (int k = 0; k < d; k++)
    {
      const int l = table[c * 2 * d + k], pos = table[c * 2 * d + d + k];
      const double count = (l <= 1) ? 1.0 : double(int64_t(1) << (l - 1));
      f[c * d + k] = static_cast<float>((l + (pos + 0.5) / count) / double(max_level + 1));
    }
  AKM_TRY(m->feat.upload(f.data(), num_cells * d));
  *out = m.release();
  return AKM_SUCCESS;
}

akm_status akm_mlp_info(akm_precond p, int64_t *num_params, int *inputs, int *padded_inputs,
                        int *hidden, int *outputs, int *padded_outputs)
{
  auto *m = as_mlp(p);
  if (m == nullptr)
    return set_error(AKM_ERR_INVALID, "mlp: not an MLP preconditioner");
  if (num_params)
    *num_params = m->num_params;
  if (inputs)
    *inputs = m->N + m->d;
  if (padded_inputs)
    *padded_inputs = m->K1;
  if (hidden)
    *hidden = m->H;
  if (outputs)
    *outputs = m->N;
  if (padded_outputs)
    *padded_outputs = m->N3;
  return AKM_SUCCESS;
}

akm_status akm_mlp_set_stream(akm_precond p, void *stream)
{
  auto *m = as_mlp(p);
  if (m == nullptr)
    return set_error(AKM_ERR_INVALID, "mlp: not an MLP preconditioner");
  m->stream = static_cast<hipStream_t>(stream);
  return AKM_SUCCESS;
}

akm_status akm_mlp_get_params(akm_precond p, float *host)
{
  auto *m = as_mlp(p);
  if (m == nullptr || host == nullptr)
    return set_error(AKM_ERR_INVALID, "mlp: null argument");
  AKM_HIP(hipMemcpyAsync(host, m->params.data(), sizeof(float) * m->num_params,
                         hipMemcpyDeviceToHost, m->stream));
  AKM_HIP(hipStreamSynchronize(m->stream));
  return AKM_SUCCESS;
}

akm_status akm_mlp_set_params(akm_precond p, const float *host)
{
  auto *m = as_mlp(p);
  if (m == nullptr || host == nullptr)
    return set_error(AKM_ERR_INVALID, "mlp: null argument");
  AKM_HIP(hipMemcpyAsync(m->params.data(), host, sizeof(float) * m->num_params,
                         hipMemcpyHostToDevice, m->stream));
  AKM_HIP(hipStreamSynchronize(m->stream));
  m->bf16_current = false;
  return AKM_SUCCESS;
}

akm_status akm_mlp_forward(akm_precond p, const float *X, int64_t rows, float *O)
{
  auto *m = as_mlp(p);
  if (m == nullptr || X == nullptr || O == nullptr)
    return set_error(AKM_ERR_INVALID, "mlp: null argument");
  if (rows < 1 || rows > (1 << 24))
    return set_error(AKM_ERR_INVALID, "mlp: rows %lld out of range", static_cast<long long>(rows));
  AKM_TRY(m->ensure_batch(static_cast<int>(rows)));
  AKM_TRY(m->forward(m->stream, static_cast<int>(rows), X));
  AKM_HIP(hipMemcpyAsync(O, m->O.data(), sizeof(float) * rows * m->N3, hipMemcpyDeviceToDevice,
                         m->stream));
  AKM_HIP(hipStreamSynchronize(m->stream));
  return AKM_SUCCESS;
}

akm_status akm_mlp_gradients(akm_precond p, const float *X, const float *T, int64_t rows,
                             double *loss, float *grad)
{
  auto *m = as_mlp(p);
  if (m == nullptr || X == nullptr || T == nullptr)
    return set_error(AKM_ERR_INVALID, "mlp: null argument");
  if (rows < 1 || rows > (1 << 24))
    return set_error(AKM_ERR_INVALID, "mlp: rows %lld out of range", static_cast<long long>(rows));
  const int B = static_cast<int>(rows);
  AKM_TRY(m->ensure_batch(B));
  AKM_TRY(m->forward(m->stream, B, X));
  AKM_TRY(m->backward(m->stream, B, X, T, 0));
  if (grad != nullptr)
    AKM_HIP(hipMemcpyAsync(grad, m->grad.data(), sizeof(float) * m->num_params,
                           hipMemcpyDeviceToDevice, m->stream));
  double l = 0;
  AKM_HIP(hipMemcpyAsync(&l, m->loss.data(), sizeof(double), hipMemcpyDeviceToHost, m->stream));
  AKM_HIP(hipStreamSynchronize(m->stream));
  if (loss)
    *loss = l / (double(B) * m->N);
  return AKM_SUCCESS;
}

akm_status akm_mlp_adam_step(akm_precond p, const float *grad, double lr)
{
  auto *m = as_mlp(p);
  if (m == nullptr || grad == nullptr)
    return set_error(AKM_ERR_INVALID, "mlp: null argument");
  AKM_TRY(m->adam(m->stream, grad, static_cast<float>(lr)));
  AKM_HIP(hipStreamSynchronize(m->stream));
  return AKM_SUCCESS;
}

akm_status akm_mlp_train(akm_precond p, const double *residuals, const double *corrections,
                         int64_t num_vectors, int epochs, int batch, double lr, uint64_t seed,
                         double *final_loss)
{
  auto *m = as_mlp(p);
  if (m == nullptr || residuals == nullptr || corrections == nullptr)
    return set_error(AKM_ERR_INVALID, "mlp: null argument");
  if (num_vectors < 1 || epochs < 1 || batch < 1 || !(lr > 0))
    return set_error(AKM_ERR_INVALID, "mlp train: vectors %lld, epochs %d, batch %d, lr %g",
                     static_cast<long long>(num_vectors), epochs, batch, lr);
  hipStream_t s = m->stream;
  const int64_t S = num_vectors * m->cells;
  const int B = static_cast<int>(std::min<int64_t>(batch, S));
  AKM_TRY(m->ensure_batch(B));
  dev_buffer<float> X, T;
  dev_buffer<int64_t> perm_d;
  AKM_TRY(X.allocate(S * m->K1));
  AKM_TRY(T.allocate(S * m->N3));
  AKM_TRY(perm_d.allocate(S));
  mlp_make_samples<<<grid_for(S), 256, 0, s>>>(num_vectors, m->cells, m->N, m->d, m->K1, m->N3,
                                               residuals, corrections, m->feat.data(), X.data(),
                                               T.data());
  AKM_TRY(check_launch("mlp samples"));
  std::vector<int64_t> perm(S);
  std::iota(perm.begin(), perm.end(), int64_t{0});
  std::mt19937_64 rng(seed);
  const int64_t nb = S / B; // full batches only
  double epoch_loss = 0;
  for (int ep = 0; ep < epochs; ep++)
  {
    std::shuffle(perm.begin(), perm.end(), rng);
    AKM_HIP(hipMemcpyAsync(perm_d.data(), perm.data(), sizeof(int64_t) * S,
                           hipMemcpyHostToDevice, s));
    for (int64_t bi = 0; bi < nb; bi++)
    {
      mlp_gather<<<B, 256, 0, s>>>(B, perm_d.data(), bi * B, m->K1, m->N3, X.data(), T.data(),
                                    m->Xb.data(), m->Tb.data());
      AKM_TRY(check_launch("mlp gather"));
      AKM_TRY(m->forward(s, B, m->Xb.data()));
      AKM_TRY(m->backward(s, B, m->Xb.data(), m->Tb.data(), bi > 0));
      AKM_TRY(m->adam(s, m->grad.data(), static_cast<float>(lr)));
    }
    // the host copy of perm is reshuffled next epoch: wait for the upload
    AKM_HIP(hipStreamSynchronize(s));
    if (ep == epochs - 1)
    {
      double l = 0;
      AKM_HIP(hipMemcpy(&l, m->loss.data(), sizeof(double), hipMemcpyDeviceToHost));
      epoch_loss = l / (double(nb) * B * m->N);
    }
  }
  AKM_TRY(m->refresh_bf16(s));
  AKM_HIP(hipStreamSynchronize(s));
  if (final_loss)
    *final_loss = epoch_loss;
  return AKM_SUCCESS;
}

akm_status akm_precond_apply(akm_precond p, const double *v, double *z, void *stream)
{
  if (p == nullptr || v == nullptr || z == nullptr)
    return set_error(AKM_ERR_INVALID, "precond: null argument");
  return p->apply(v, z, static_cast<hipStream_t>(stream));
}

akm_status akm_precond_destroy(akm_precond p)
{
  delete p;
  return AKM_SUCCESS;
}

} // extern "C"
