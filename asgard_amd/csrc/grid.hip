// Sparse-grid index construction: the 1D cell hierarchy connectivity and the
// cell-pair sparsity pattern of the kronmult operator, built on the GPU.
//
// Connection rule (the reference's check_connected,
// src/asgard_kronmult_matrix.cpp:230-300): two 1D cells (L1, p1), (L2, p2) are
// connected when their closed supports touch on the periodic unit interval:
// levels 0 and 1 span the domain; on one level the cell and its two
// neighbours (with wrap-around); across levels the ancestor, plus the
// ancestor's neighbour when the finer cell sits on that edge of the
// ancestor's support.  A d-dimensional pair is connected when every
// dimension is.
#include "akm_internal.hpp"
#include "grid.hpp"

#include <algorithm>
#include <cstring>
#include <vector>

namespace akm
{
__host__ __device__ inline bool connected_1d(int L1, int p1, int L2, int p2, bool full)
{
  if (L1 > L2)
  {
    int t = L1;
    L1 = L2;
    L2 = t;
    t = p1;
    p1 = p2;
    p2 = t;
  }
  if (L1 <= 1)
    return true;
  if (!full)
  {
    // volume connectivity: (L1, p1) is an ancestor of (L2, p2)
    return (p2 >> (L2 - L1)) == p1;
  }
  const int last1 = (1 << (L1 - 1)) - 1, last2 = (1 << (L2 - 1)) - 1;
  if ((p1 == 0 && p2 == last2) || (p2 == 0 && p1 == last1))
    return true; // periodic wrap
  const int shift = L2 - L1;
  const int anc = p2 >> shift;
  if (shift == 0)
    return abs(p1 - p2) <= 1;
  if (anc == p1)
    return true;
  // the finer cell touches a neighbour of its ancestor only when it sits at
  // the matching edge of the ancestor's support
  const int mask = (1 << shift) - 1;
  const int within = p2 & mask;
  if (within == 0)
    return anc - 1 == p1;
  if (within == mask)
    return anc + 1 == p1;
  return false;
}

__host__ __device__ inline int cell_index_1d(int level, int pos)
{
  return (level == 0) ? 0 : ((1 << (level - 1)) + pos);
}

int build_connect_1d(int max_level, bool full, std::vector<int> &pntr, std::vector<int> &indx,
                     std::vector<int> &diag)
{
  if (max_level < 1 || max_level > 20)
    return set_error(AKM_ERR_INVALID, "connect_1d: max_level %d out of range [1, 20]",
                     max_level);
  const int rows = 1 << max_level;
  std::vector<int> lev(rows), pos(rows);
  for (int i = 0; i < rows; i++)
  {
    int L = 0;
    while (i >= (1 << L))
      L++;
    // cell 0 is level 0; cell i >= 1 is level floor(log2 i) + 1
    lev[i] = (i == 0) ? 0 : L;
    pos[i] = (i == 0) ? 0 : i - (1 << (L - 1));
  }
  pntr.assign(rows + 1, 0);
  diag.assign(rows, 0);
  indx.clear();
  for (int r = 0; r < rows; r++)
  {
    for (int c = 0; c < rows; c++)
    {
      // levels 0 and 1 are connected to everything in both hierarchies
      bool conn = (r <= 1 || c <= 1) || connected_1d(lev[r], pos[r], lev[c], pos[c], full);
      if (conn)
      {
        if (c == r)
          diag[r] = static_cast<int>(indx.size());
        indx.push_back(c);
      }
    }
    pntr[r + 1] = static_cast<int>(indx.size());
  }
  return AKM_SUCCESS;
}

int build_offset_table(int max_level, std::vector<int> &table)
{
  if (max_level > 12)
    return set_error(AKM_ERR_UNSUPPORTED,
                     "sparse-grid kronmult: max_level %d > 12 (1D offset table too large)",
                     max_level);
  std::vector<int> pntr, indx, diag;
  AKM_TRY(build_connect_1d(max_level, true, pntr, indx, diag));
  const int M = 1 << max_level;
  table.assign(size_t(M) * M, -1);
  for (int r = 0; r < M; r++)
    for (int j = pntr[r]; j < pntr[r + 1]; j++)
      table[size_t(r) * M + indx[j]] = j;
  return AKM_SUCCESS;
}

void cell_1d_indices(int d, int64_t num_cells, const int *table, std::vector<int> &elem)
{
  elem.resize(size_t(num_cells) * d);
  for (int64_t c = 0; c < num_cells; c++)
    for (int k = 0; k < d; k++)
      elem[size_t(c) * d + k] = cell_index_1d(table[c * 2 * d + k], table[c * 2 * d + d + k]);
}

// ----------------------------------------------------------- GPU scan ---

template<int D>
__device__ __forceinline__ bool connected_cells(const int *__restrict__ rc,
                                                const int *__restrict__ cc)
{
#pragma unroll
  for (int k = 0; k < D; k++)
    if (!connected_1d(rc[k], rc[D + k], cc[k], cc[D + k], true))
      return false;
  return true;
}

//! One wavefront per row; lanes test 64 columns at a time.  In the count
//! pass `indx` is null and the row length goes to counts[row]; in the fill
//! pass the connected columns are written at pntr[row] in the order of the
//! candidate list: ascending cell ids, or order[0..num_cells) when given.
template<int D>
__global__ void __launch_bounds__(256) pattern_scan(int64_t num_cells, const int *table,
                                                    int64_t *counts, const int64_t *pntr,
                                                    int *indx, const int *order)
{
  const int lane = threadIdx.x & 63;
  const int64_t row = int64_t(blockIdx.x) * 4 + (threadIdx.x >> 6);
  if (row >= num_cells)
    return;
  int rc[2 * D];
#pragma unroll
  for (int k = 0; k < 2 * D; k++)
    rc[k] = table[row * 2 * D + k];
  int64_t found = 0;
  int *out = (indx != nullptr) ? indx + pntr[row] : nullptr;
  for (int64_t c0 = 0; c0 < num_cells; c0 += 64)
  {
    const int64_t pos = c0 + lane;
    const int64_t col = (order != nullptr && pos < num_cells) ? order[pos] : pos;
    bool conn = false;
    if (pos < num_cells)
    {
      int cc[2 * D];
#pragma unroll
      for (int k = 0; k < 2 * D; k++)
        cc[k] = table[col * 2 * D + k];
      conn = connected_cells<D>(rc, cc);
    }
    const unsigned long long mask = __ballot(conn);
    if (out != nullptr && conn)
    {
      const unsigned long long below = mask & ((1ull << lane) - 1ull);
      out[found + __popcll(below)] = static_cast<int>(col);
    }
    found += __popcll(mask);
  }
  if (indx == nullptr && lane == 0)
    counts[row] = found;
}

template<int D>
static int scan_pass(int64_t num_cells, const int *table_dev, int64_t *counts,
                     const int64_t *pntr, int *indx, const int *order)
{
  const int64_t blocks = (num_cells + 3) / 4;
  pattern_scan<D><<<dim3(static_cast<unsigned>(blocks)), 256>>>(num_cells, table_dev, counts,
                                                                  pntr, indx, order);
  return check_launch("sparse pattern scan");
}

static int scan_dispatch(int d, int64_t num_cells, const int *table_dev, int64_t *counts,
                         const int64_t *pntr, int *indx, const int *order)
{
  switch (d)
  {
  case 1: return scan_pass<1>(num_cells, table_dev, counts, pntr, indx, order);
  case 2: return scan_pass<2>(num_cells, table_dev, counts, pntr, indx, order);
  case 3: return scan_pass<3>(num_cells, table_dev, counts, pntr, indx, order);
  case 4: return scan_pass<4>(num_cells, table_dev, counts, pntr, indx, order);
  case 5: return scan_pass<5>(num_cells, table_dev, counts, pntr, indx, order);
  case 6: return scan_pass<6>(num_cells, table_dev, counts, pntr, indx, order);
  default: return set_error(AKM_ERR_UNSUPPORTED, "sparse pattern: d = %d not in [1, 6]", d);
  }
}

void lexicographic_order(int d, int64_t num_cells, const int *elem, std::vector<int> &order)
{
  order.resize(num_cells);
  for (int64_t c = 0; c < num_cells; c++)
    order[c] = static_cast<int>(c);
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
    for (int k = 0; k < d; k++)
    {
      const int ea = elem[int64_t(a) * d + k], eb = elem[int64_t(b) * d + k];
      if (ea != eb)
        return ea < eb;
    }
    return false;
  });
}

int device_sparse_pattern(int d, int64_t num_cells, const int *table_host,
                          std::vector<int64_t> &pntr, dev_buffer<int> &indx_dev,
                          const int *order_host)
{
  if (num_cells <= 0)
    return set_error(AKM_ERR_INVALID, "sparse pattern: no cells");
  if (num_cells > (int64_t(1) << 31) - 2)
    return set_error(AKM_ERR_INVALID, "sparse pattern: too many cells");
  dev_buffer<int> table_dev, order_dev;
  AKM_TRY(table_dev.upload(table_host, num_cells * 2 * d));
  if (order_host != nullptr)
    AKM_TRY(order_dev.upload(order_host, num_cells));
  dev_buffer<int64_t> counts, pntr_dev;
  AKM_TRY(counts.allocate(num_cells));
  AKM_TRY(scan_dispatch(d, num_cells, table_dev.data(), counts.data(), nullptr, nullptr,
                        nullptr));
  std::vector<int64_t> cnt(num_cells);
  AKM_HIP(hipMemcpy(cnt.data(), counts.data(), sizeof(int64_t) * num_cells,
                    hipMemcpyDeviceToHost));
  pntr.assign(num_cells + 1, 0);
  for (int64_t i = 0; i < num_cells; i++)
    pntr[i + 1] = pntr[i] + cnt[i];
  AKM_TRY(pntr_dev.upload(pntr.data(), num_cells + 1));
  AKM_TRY(indx_dev.allocate(pntr[num_cells]));
  AKM_TRY(scan_dispatch(d, num_cells, table_dev.data(), nullptr, pntr_dev.data(),
                        indx_dev.data(), order_dev.data()));
  AKM_HIP(hipDeviceSynchronize());
  return AKM_SUCCESS;
}

} // namespace akm
