/*
 * Oracle: sparse-grid element table, 1D connectivity and the cell-pair
 * sparsity pattern of the kronmult operator.
 *
 * TEST INFRASTRUCTURE ONLY (the checker; never linked by the product).
 */
#include "asgard_oracle.h"

#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------- level tuples --- */

/* Lower index set generation, grouped by |l|_1, restating
 * permutations::generate_lower_index_level_sets (src/permutations.hpp:71-108)
 * with the selection criterion of get_lequal_multi
 * (src/permutations.cpp:36-55): l_k <= levels[k] and sum l_k <= max(levels).
 * Calls emit(level_sum, tuple) in generation order. */
typedef void (*tuple_sink)(void *ctx, int level_sum, const int *tuple);

static int inside_lequal(int d, const int *levels, int limit, const int *idx)
{
  int s = 0;
  for (int k = 0; k < d; k++)
  {
    if (idx[k] > levels[k])
      return 0;
    s += idx[k];
  }
  return s <= limit;
}

static void generate_level_sets(int d, const int *levels, int limit,
                                tuple_sink sink, void *ctx)
{
  int root[16];
  memset(root, 0, sizeof(root));
  int c = 0, level = 0, is_in = 1;
  while (is_in || c < d)
  {
    if (is_in)
    {
      sink(ctx, level, root);
      c = 0;
      root[c]++;
      level++;
    }
    else
    {
      for (int k = 0; k <= c; k++)
        root[k] = 0;
      c++;
      if (c == d)
        break;
      root[c]++;
      level = 0;
      for (int k = 0; k < d; k++)
        level += root[k];
    }
    is_in = inside_lequal(d, levels, limit, root);
  }
}

struct tuple_collect
{
  int d;
  int max_sum;
  int64_t *count_per_sum; /* [max_sum + 1] */
  int64_t *fill_pos;      /* write position per sum (NULL while counting) */
  int *out;
};

static void tuple_count_sink(void *vctx, int level_sum, const int *tuple)
{
  struct tuple_collect *ctx = (struct tuple_collect *)vctx;
  (void)tuple;
  ctx->count_per_sum[level_sum]++;
}

static void tuple_fill_sink(void *vctx, int level_sum, const int *tuple)
{
  struct tuple_collect *ctx = (struct tuple_collect *)vctx;
  int64_t pos = ctx->fill_pos[level_sum]++;
  memcpy(ctx->out + pos * ctx->d, tuple, sizeof(int) * ctx->d);
}

static int max_of(int d, const int *levels)
{
  int m = levels[0];
  for (int k = 1; k < d; k++)
    if (levels[k] > m)
      m = levels[k];
  return m;
}

int64_t orc_num_level_tuples(int d, const int *levels, int full_grid)
{
  if (d < 1 || d > 16)
    return -1;
  if (full_grid)
  {
    int64_t num = 1;
    for (int k = 0; k < d; k++)
      num *= (levels[k] + 1);
    return num;
  }
  int limit = max_of(d, levels);
  int64_t *cnt = (int64_t *)calloc((size_t)(limit * d + 1), sizeof(int64_t));
  struct tuple_collect ctx = {d, limit * d, cnt, NULL, NULL};
  generate_level_sets(d, levels, limit, tuple_count_sink, &ctx);
  int64_t total = 0;
  for (int s = 0; s <= limit * d; s++)
    total += cnt[s];
  free(cnt);
  return total;
}

int orc_level_tuples(int d, const int *levels, int full_grid, int *tuples)
{
  if (d < 1 || d > 16)
    return -1;
  if (full_grid)
  {
    /* permutations::get_max_multi, src/permutations.cpp:112-149, with
     * last_index_decreasing = false: digit 0 fastest. */
    int64_t num = orc_num_level_tuples(d, levels, 1);
    for (int64_t i = 0; i < num; i++)
    {
      int64_t t = i;
      for (int k = 0; k < d; k++)
      {
        tuples[i * d + k] = (int)(t % (levels[k] + 1));
        t /= (levels[k] + 1);
      }
    }
    return 0;
  }
  /* select_indexex(order_by_n = true, levels_in_reverse = true):
   * groups of equal |l|_1, the largest sum first, zero tuple last
   * (transp_data, src/permutations.hpp:131-166). */
  int limit = max_of(d, levels);
  int nsum = limit * d + 1;
  int64_t *cnt = (int64_t *)calloc((size_t)nsum, sizeof(int64_t));
  int64_t *pos = (int64_t *)calloc((size_t)nsum, sizeof(int64_t));
  struct tuple_collect ctx = {d, nsum - 1, cnt, NULL, NULL};
  generate_level_sets(d, levels, limit, tuple_count_sink, &ctx);
  int64_t start = 0;
  for (int s = nsum - 1; s >= 0; s--)
  {
    pos[s] = start;
    start += cnt[s];
  }
  ctx.fill_pos = pos;
  ctx.out = tuples;
  generate_level_sets(d, levels, limit, tuple_fill_sink, &ctx);
  free(cnt);
  free(pos);
  return 0;
}

static int64_t cells_of_tuple(int d, const int *tuple)
{
  int64_t c = 1;
  for (int k = 0; k < d; k++)
    c *= (tuple[k] <= 1) ? 1 : ((int64_t)1 << (tuple[k] - 1));
  return c;
}

int64_t orc_num_cells(int d, const int *levels, int full_grid)
{
  int64_t nt = orc_num_level_tuples(d, levels, full_grid);
  if (nt < 0)
    return -1;
  int *tuples = (int *)malloc(sizeof(int) * (size_t)(nt * d));
  orc_level_tuples(d, levels, full_grid, tuples);
  int64_t total = 0;
  for (int64_t i = 0; i < nt; i++)
    total += cells_of_tuple(d, tuples + i * d);
  free(tuples);
  return total;
}

int orc_element_table(int d, const int *levels, int full_grid, int max_level,
                      int *table, int64_t *ids)
{
  int64_t nt = orc_num_level_tuples(d, levels, full_grid);
  if (nt < 0)
    return -1;
  int *tuples = (int *)malloc(sizeof(int) * (size_t)(nt * d));
  orc_level_tuples(d, levels, full_grid, tuples);
  int64_t cell = 0;
  for (int64_t i = 0; i < nt; i++)
  {
    const int *lt = tuples + i * d;
    /* get_cell_index_set: dim 0 varies fastest, the last dim slowest. */
    int64_t ncell = cells_of_tuple(d, lt);
    for (int64_t j = 0; j < ncell; j++)
    {
      int *coords = table + cell * 2 * d;
      int64_t r = j;
      int64_t id = 0, stride = 1;
      for (int k = 0; k < d; k++)
      {
        int64_t nk = (lt[k] <= 1) ? 1 : ((int64_t)1 << (lt[k] - 1));
        int p = (int)(r % nk);
        r /= nk;
        coords[k] = lt[k];
        coords[d + k] = p;
        /* map_to_id, src/elements.cpp:59-84 */
        int64_t i1d = (lt[k] == 0) ? 0 : (((int64_t)1 << (lt[k] - 1)) + p);
        id += i1d * stride;
        stride *= ((int64_t)1 << max_level);
      }
      if (ids)
        ids[cell] = id;
      cell++;
    }
  }
  free(tuples);
  return 0;
}

/* ----------------------------------------------------------- connect_1d --- */

/* A growable int array. */
struct ivec
{
  int *v;
  int64_t size, cap;
};

static void ivec_push(struct ivec *a, int x)
{
  if (a->size == a->cap)
  {
    a->cap = a->cap ? 2 * a->cap : 64;
    a->v = (int *)realloc(a->v, sizeof(int) * (size_t)a->cap);
  }
  a->v[a->size++] = x;
}

/* connect_1d::build_connections, src/asgard_grid_1d.hpp:159-317. */
int orc_connect_1d(int levels, int mode, int *pntr_out, int *indx_out,
                   int *diag_out)
{
  if (levels < 1 || levels > 28)
    return -1;
  const int full = (mode == 1);
  int rows = 1 << levels;
  int *pntr = (int *)calloc((size_t)rows + 1, sizeof(int));
  int *diag = (int *)calloc((size_t)rows, sizeof(int));
  int *cpl = (int *)malloc(sizeof(int) * (size_t)(levels + 2));
  struct ivec indx = {NULL, 0, 0};

  cpl[0] = 1;
  cpl[1] = 1;
  for (int l = 2; l < levels + 2; l++)
    cpl[l] = 2 * cpl[l - 1];

  /* cells 0 and 1 connect to everything */
  pntr[1] = rows;
  pntr[2] = 2 * rows;
  for (int i = 0; i < rows; i++)
    ivec_push(&indx, i);
  for (int i = 0; i < rows; i++)
    ivec_push(&indx, i);
  diag[0] = 0;
  diag[1] = 1 + rows;

  for (int l = 2; l < levels + 1; l++)
  {
    int level_size = cpl[l];
    /* left edge cell */
    int i = level_size;
    ivec_push(&indx, 0);
    ivec_push(&indx, 1);
    for (int upl = 2; upl < l; upl++)
    {
      ivec_push(&indx, cpl[upl]);
      if (full)
        ivec_push(&indx, cpl[upl + 1] - 1);
    }
    diag[i] = (int)indx.size;
    ivec_push(&indx, i);
    if (full)
    {
      ivec_push(&indx, i + 1);
      if (l > 2)
        ivec_push(&indx, cpl[l + 1] - 1);
    }
    for (int downl = l + 1; downl < levels + 1; downl++)
    {
      int lstart = cpl[downl];
      for (int downp = 0; downp < cpl[downl - l + 1]; downp++)
        ivec_push(&indx, lstart + downp);
      if (full)
      {
        ivec_push(&indx, lstart + cpl[downl - l + 1]);
        ivec_push(&indx, cpl[downl + 1] - 1);
      }
    }
    pntr[i + 1] = (int)indx.size;

    /* middle cells */
    for (int p = 1; p < level_size - 1; p++)
    {
      i++;
      ivec_push(&indx, 0);
      ivec_push(&indx, 1);
      for (int upl = 2; upl < l; upl++)
      {
        int segment_size = cpl[l - upl + 1];
        int ancestor = p / segment_size;
        int edge = p - ancestor * segment_size;
        if (full && edge == 0)
          ivec_push(&indx, cpl[upl] + ancestor - 1);
        ivec_push(&indx, cpl[upl] + ancestor);
        if (full && edge == segment_size - 1)
          ivec_push(&indx, cpl[upl] + ancestor + 1);
      }
      if (full)
        ivec_push(&indx, i - 1);
      diag[i] = (int)indx.size;
      ivec_push(&indx, i);
      if (full)
        ivec_push(&indx, i + 1);
      int left_kid = p, num_kids = 1;
      for (int downl = l + 1; downl < levels + 1; downl++)
      {
        left_kid *= 2;
        num_kids *= 2;
        if (full)
          ivec_push(&indx, cpl[downl] + left_kid - 1);
        for (int j = left_kid; j < left_kid + num_kids; j++)
          ivec_push(&indx, cpl[downl] + j);
        if (full)
          ivec_push(&indx, cpl[downl] + left_kid + num_kids);
      }
      pntr[i + 1] = (int)indx.size;
    }

    /* right edge cell */
    i++;
    ivec_push(&indx, 0);
    ivec_push(&indx, 1);
    for (int upl = 2; upl < l; upl++)
    {
      if (full)
        ivec_push(&indx, cpl[upl]);
      ivec_push(&indx, cpl[upl + 1] - 1);
    }
    if (full)
    {
      if (l > 2)
        ivec_push(&indx, cpl[l]);
      ivec_push(&indx, i - 1);
    }
    diag[i] = (int)indx.size;
    ivec_push(&indx, i);
    for (int downl = l + 1; downl < levels + 1; downl++)
    {
      if (full)
        ivec_push(&indx, cpl[downl]);
      int lend = cpl[downl + 1] - 1;
      int lbegin = cpl[downl - l + 1] - 1;
      if (full)
        lbegin += 1;
      for (int downp = lbegin; downp > -1; downp--)
        ivec_push(&indx, lend - downp);
    }
    pntr[i + 1] = (int)indx.size;
  }

  int nconn = (int)indx.size;
  if (pntr_out)
    memcpy(pntr_out, pntr, sizeof(int) * ((size_t)rows + 1));
  if (indx_out)
    memcpy(indx_out, indx.v, sizeof(int) * (size_t)nconn);
  if (diag_out)
    memcpy(diag_out, diag, sizeof(int) * (size_t)rows);
  free(pntr);
  free(diag);
  free(cpl);
  free(indx.v);
  return nconn;
}

int orc_connect_1d_offset(const int *pntr, const int *indx, int row, int col)
{
  int sstart = pntr[row], send = pntr[row + 1] - 1;
  int current = (sstart + send) / 2;
  while (sstart <= send)
  {
    if (indx[current] < col)
      sstart = current + 1;
    else if (indx[current] > col)
      send = current - 1;
    else
      return current;
    current = (sstart + send) / 2;
  }
  return -1;
}

/* ------------------------------------------------------ check_connected --- */

int orc_check_connected_1d(int L1, int p1, int L2, int p2)
{
  /* src/asgard_kronmult_matrix.cpp:230-275, requires L1 <= L2 */
  if (L1 <= 1 || L2 <= 1)
    return 1;
  if ((p1 == 0 && p2 == ((1 << (L2 - 1)) - 1)) ||
      (p2 == 0 && p1 == ((1 << (L1 - 1)) - 1)))
    return 1;
  if (L1 == L2)
    return abs(p1 - p2) <= 1;
  int side = (p2 % 2 == 0) ? -1 : 1;
  while (L2 > L1)
  {
    if (p2 % 2 == 0)
      side = (side == -1) ? -1 : 0;
    else
      side = (side == 1) ? 1 : 0;
    L2--;
    p2 /= 2;
  }
  return (p2 == p1) || (p2 + side == p1);
}

int orc_check_connected(int d, const int *row, const int *col)
{
  /* src/asgard_kronmult_matrix.cpp:278-300 */
  for (int j = 0; j < d; j++)
  {
    if (row[j] <= col[j])
    {
      if (!orc_check_connected_1d(row[j], row[j + d], col[j], col[j + d]))
        return 0;
    }
    else
    {
      if (!orc_check_connected_1d(col[j], col[j + d], row[j], row[j + d]))
        return 0;
    }
  }
  return 1;
}

int64_t orc_sparse_pattern_count(int d, int64_t num_cells, const int *table,
                                 int64_t *pntr)
{
  int64_t *cnt = (int64_t *)calloc((size_t)num_cells, sizeof(int64_t));
#pragma omp parallel for schedule(dynamic, 16)
  for (int64_t row = 0; row < num_cells; row++)
  {
    const int *rc = table + 2 * d * row;
    int64_t c = 0;
    for (int64_t col = 0; col < num_cells; col++)
      if (orc_check_connected(d, rc, table + 2 * d * col))
        c++;
    cnt[row] = c;
  }
  int64_t nnz = 0;
  for (int64_t i = 0; i < num_cells; i++)
  {
    if (pntr)
      pntr[i] = nnz;
    nnz += cnt[i];
  }
  if (pntr)
    pntr[num_cells] = nnz;
  free(cnt);
  return nnz;
}

static int cell_1d(int level, int pos)
{
  return (level == 0) ? 0 : ((1 << (level - 1)) + pos);
}

void orc_cell_1d_index(int d, int64_t num_cells, const int *table, int *elem)
{
  for (int64_t c = 0; c < num_cells; c++)
    for (int k = 0; k < d; k++)
      elem[c * d + k] = cell_1d(table[c * 2 * d + k], table[c * 2 * d + d + k]);
}

int orc_sparse_pattern_fill(int d, int n, int T, int max_level,
                            int64_t num_cells, const int *table,
                            const int64_t *pntr, int *indx, int *iA)
{
  int nconn = orc_connect_1d(max_level, 1, NULL, NULL, NULL);
  if (nconn < 0)
    return -1;
  int *cp = (int *)malloc(sizeof(int) * ((size_t)(1 << max_level) + 1));
  int *ci = (int *)malloc(sizeof(int) * (size_t)nconn);
  orc_connect_1d(max_level, 1, cp, ci, NULL);
  const int n2 = n * n;
  const int block1d = d * T * n2;
  int bad = 0;
#pragma omp parallel for schedule(dynamic, 16) reduction(+ : bad)
  for (int64_t row = 0; row < num_cells; row++)
  {
    const int *rc = table + 2 * d * row;
    int64_t j = pntr[row];
    for (int64_t col = 0; col < num_cells; col++)
    {
      const int *cc = table + 2 * d * col;
      if (!orc_check_connected(d, rc, cc))
        continue;
      indx[j] = (int)col;
      if (iA)
      {
        /* compute_coefficient_offsets, src/asgard_kronmult_matrix.cpp:302-320 */
        int off[16];
        for (int k = 0; k < d; k++)
        {
          off[k] = orc_connect_1d_offset(cp, ci, cell_1d(rc[k], rc[d + k]),
                                         cell_1d(cc[k], cc[d + k]));
          if (off[k] < 0)
            bad++;
        }
        int *ia = iA + j * (int64_t)T * d;
        for (int t = 0; t < T; t++)
          for (int k = 0; k < d; k++)
            *ia++ = off[k] * block1d + (t * d + k) * n2;
      }
      j++;
    }
  }
  free(cp);
  free(ci);
  return bad ? -2 : 0;
}

void orc_build_vA(int d, int n, int T, int max_level, const double *const *ops,
                  int64_t lda, double *vA)
{
  int rows = 1 << max_level;
  int nconn = orc_connect_1d(max_level, 1, NULL, NULL, NULL);
  int *cp = (int *)malloc(sizeof(int) * ((size_t)rows + 1));
  int *ci = (int *)malloc(sizeof(int) * (size_t)nconn);
  orc_connect_1d(max_level, 1, cp, ci, NULL);
  double *pA = vA;
  for (int row = 0; row < rows; row++)
    for (int j = cp[row]; j < cp[row + 1]; j++)
    {
      int col = ci[j];
      for (int t = 0; t < T; t++)
        for (int k = 0; k < d; k++)
        {
          const double *op = ops[t * d + k];
          for (int i = 0; i < n; i++)
            for (int s = 0; s < n; s++)
              *pA++ = op[(int64_t)n * row + s + lda * ((int64_t)n * col + i)];
        }
    }
  free(cp);
  free(ci);
}

void orc_build_dense_terms(int d, int n, int T, int nb,
                           const double *const *ops, int64_t lda,
                           double *const *terms)
{
  for (int t = 0; t < T; t++)
  {
    double *pA = terms[t];
    for (int k = 0; k < d; k++)
    {
      const double *op = ops[t * d + k];
      for (int ocol = 0; ocol < nb; ocol++)
        for (int orow = 0; orow < nb; orow++)
          for (int i = 0; i < n; i++)
            for (int s = 0; s < n; s++)
              *pA++ = op[(int64_t)n * orow + s + lda * ((int64_t)n * ocol + i)];
    }
  }
}

int64_t orc_sparse_pattern_rows(int d, int n, int T, int max_level,
                                int64_t num_cells, const int *table,
                                const int *rows, int num_sel, int64_t *pntr,
                                int *indx, int *iA)
{
  /* the rows of make_kronmult_sparse (src/asgard_kronmult_matrix.cpp:406-466)
   * restricted to the selected row cells; indx/iA may be NULL to count */
  int nconn = orc_connect_1d(max_level, 1, NULL, NULL, NULL);
  if (nconn < 0)
    return -1;
  int *cp = (int *)malloc(sizeof(int) * ((size_t)(1 << max_level) + 1));
  int *ci = (int *)malloc(sizeof(int) * (size_t)nconn);
  orc_connect_1d(max_level, 1, cp, ci, NULL);
  const int n2 = n * n;
  const int block1d = d * T * n2;
  int64_t j = 0;
  int bad = 0;
  pntr[0] = 0;
  for (int s = 0; s < num_sel; s++)
  {
    const int *rc = table + 2 * d * (int64_t)rows[s];
    for (int64_t col = 0; col < num_cells; col++)
    {
      const int *cc = table + 2 * d * col;
      if (!orc_check_connected(d, rc, cc))
        continue;
      if (indx)
        indx[j] = (int)col;
      if (iA)
      {
        int off[16];
        for (int k = 0; k < d; k++)
        {
          off[k] = orc_connect_1d_offset(cp, ci, cell_1d(rc[k], rc[d + k]),
                                         cell_1d(cc[k], cc[d + k]));
          if (off[k] < 0)
            bad++;
        }
        int *ia = iA + j * (int64_t)T * d;
        for (int t = 0; t < T; t++)
          for (int k = 0; k < d; k++)
            *ia++ = off[k] * block1d + (t * d + k) * n2;
      }
      j++;
    }
    pntr[s + 1] = j;
  }
  free(cp);
  free(ci);
  return bad ? -2 : j;
}
