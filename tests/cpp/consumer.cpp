// A compiled consumer of the drop-in boundary: the reference's call surface
// (include/asgard_amd.hpp over the C ABI of include/asgard_amd.h) used the
// way time_advance and the solver tests use it (src/time_advance.cpp:432-452,
// src/solver_tests.cpp:27-213, src/asgard_kronmult_matrix_tests.cpp).
//
//   consumer <input> <output>
//
// <input> (written by tests/test_gpu_consumer.py): int32 d, n, L, T, cells,
// nnz, restart; float64 dt; table (cells x 2d int32); vA (float64, count
// int64 first); pntr (cells + 1 int32); indx (nnz int32); iA (nnz*T*d
// int32); x and b (cells*n^d float64 each).  <output>: named float64
// records, each a 32-byte name, an int64 count and the values.
#include "asgard_amd.hpp"

#include <hip/hip_runtime_api.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

namespace
{
struct reader
{
  FILE *f;
  template<typename T>
  T one()
  {
    T v{};
    if (std::fread(&v, sizeof(T), 1, f) != 1)
      throw std::runtime_error("short input");
    return v;
  }
  template<typename T>
  std::vector<T> many(size_t count)
  {
    std::vector<T> v(count);
    if (count > 0 && std::fread(v.data(), sizeof(T), count, f) != count)
      throw std::runtime_error("short input");
    return v;
  }
};

struct writer
{
  FILE *f;
  void put(const char *name, std::vector<double> const &v)
  {
    char buf[32] = {};
    std::strncpy(buf, name, sizeof(buf) - 1);
    const int64_t count = static_cast<int64_t>(v.size());
    std::fwrite(buf, 1, sizeof(buf), f);
    std::fwrite(&count, sizeof(count), 1, f);
    if (count > 0)
      std::fwrite(v.data(), sizeof(double), v.size(), f);
  }
};

void hip_check(hipError_t e)
{
  if (e != hipSuccess)
    throw std::runtime_error(hipGetErrorString(e));
}
} // namespace

int main(int argc, char **argv)
{
  using namespace asgard;
  if (argc != 3)
  {
    std::fprintf(stderr, "usage: consumer <input> <output>\n");
    return 2;
  }
  try
  {
    FILE *fin = std::fopen(argv[1], "rb");
    if (fin == nullptr)
      throw std::runtime_error("cannot open input");
    reader in{fin};
    const int d = in.one<int>(), n = in.one<int>(), L = in.one<int>(), T = in.one<int>();
    const int cells = in.one<int>(), nnz = in.one<int>(), restart = in.one<int>();
    const double dt = in.one<double>();
    auto table = in.many<int>(size_t(cells) * 2 * d);
    auto vA = in.many<double>(static_cast<size_t>(in.one<int64_t>()));
    auto pntr = in.many<int>(size_t(cells) + 1);
    auto indx = in.many<int>(size_t(nnz));
    auto iA = in.many<int>(size_t(nnz) * T * d);
    int N = 1;
    for (int k = 0; k < d; k++)
      N *= n;
    const size_t size = size_t(cells) * N;
    auto x = in.many<double>(size);
    auto b = in.many<double>(size);
    std::fclose(fin);

    FILE *fout = std::fopen(argv[2], "wb");
    if (fout == nullptr)
      throw std::runtime_error("cannot open output");
    writer out{fout};

    // make_kronmult_sparse equivalent, apply<resource::host>
    auto grid = kronmult_matrix<double>::sparse_grid(d, n, table, L, T, vA);
    std::vector<double> y(size, 0.0);
    grid.apply<resource::host>(1.0, x.data(), 0.0, y.data());
    out.put("grid_host", y);
    out.put("sizes", {double(grid.tensor_size()), double(grid.input_size()),
                      double(grid.output_size()), double(grid.num_pairs()),
                      double(grid.flops()),
                      double(kronmult_matrix<double>::compute_flops(d, n, T, grid.num_pairs()))});

    // the sparse CPU-format constructor (hpp:235-272), alpha/beta update
    kronmult_matrix<double> csr(d, n, cells, cells, T, std::vector<int>(pntr),
                                std::vector<int>(indx), std::vector<int>(iA),
                                std::vector<double>(vA));
    std::vector<double> y2(b);
    csr.apply<resource::host>(-0.5, x.data(), 2.0, y2.data());
    out.put("csr_axpby", y2);

    // apply<resource::device> on device vectors
    double *xd = nullptr, *yd = nullptr;
    hip_check(hipMalloc(reinterpret_cast<void **>(&xd), sizeof(double) * size));
    hip_check(hipMalloc(reinterpret_cast<void **>(&yd), sizeof(double) * size));
    hip_check(hipMemcpy(xd, x.data(), sizeof(double) * size, hipMemcpyHostToDevice));
    grid.apply<resource::device>(1.0, xd, 0.0, yd);
    std::vector<double> y3(size);
    hip_check(hipMemcpy(y3.data(), yd, sizeof(double) * size, hipMemcpyDeviceToHost));
    out.put("grid_device", y3);

    // matrix_list: make, apply, reset_coefficients (update only), clear
    matrix_list<double> ops;
    ops.make(matrix_entry::regular,
             [&]() { return kronmult_matrix<double>::sparse_grid(d, n, table, L, T, vA); });
    std::vector<double> y4(size, 0.0);
    ops.apply<resource::host>(matrix_entry::regular, 1.0, x.data(), 0.0, y4.data());
    int made = 0;
    ops.reset_coefficients(
        matrix_entry::regular,
        [&]() {
          made++;
          return kronmult_matrix<double>();
        },
        [&](kronmult_matrix<double> &m) {
          std::vector<double> twice(vA);
          for (auto &v : twice)
            v *= 2;
          m.update_stored_coefficients(std::move(twice));
        });
    std::vector<double> y5(size, 0.0);
    ops.apply<resource::host>(matrix_entry::regular, 1.0, x.data(), 0.0, y5.data());
    out.put("list_apply", y4);
    out.put("list_twice", y5);
    const bool unset_before = !ops[matrix_entry::imex_implicit];
    ops.clear(matrix_entry::regular);
    out.put("list_flags", {double(made), double(unset_before), double(bool(ops[regular]))});

    // the empty matrix (hpp:123-130): y = beta*y
    kronmult_matrix<double> empty(d, n, cells, cells);
    std::vector<double> y6(b);
    empty.apply<resource::host>(1.0, x.data(), 3.0, y6.data());
    out.put("empty_scale", y6);

    // simple_gmres_euler on host and on device vectors
    std::vector<double> xs(size, 0.0);
    auto info = solver::simple_gmres_euler<double, resource::host>(dt, grid, xs, b, restart,
                                                                   no_user_value, 1e-12);
    out.put("gmres_host", xs);
    out.put("gmres_info", {info.error, double(info.iterations)});
    struct device_vector
    {
      double *p;
      size_t n;
      double *data() const { return p; }
      size_t size() const { return n; }
    };
    double *bd = nullptr;
    hip_check(hipMalloc(reinterpret_cast<void **>(&bd), sizeof(double) * size));
    hip_check(hipMemcpy(bd, b.data(), sizeof(double) * size, hipMemcpyHostToDevice));
    hip_check(hipMemset(xd, 0, sizeof(double) * size));
    device_vector xdv{xd, size}, bdv{bd, size};
    auto dinfo = solver::simple_gmres_euler<double, resource::device>(dt, grid, xdv, bdv, restart,
                                                                      no_user_value, 1e-12);
    std::vector<double> xs2(size);
    hip_check(hipMemcpy(xs2.data(), xd, sizeof(double) * size, hipMemcpyDeviceToHost));
    out.put("gmres_device", xs2);
    out.put("gmres_device_info", {dinfo.error, double(dinfo.iterations)});

    // the reference's error behaviour: restart > n throws invalid_argument
    int threw = 0;
    try
    {
      std::vector<double> xz(size, 0.0);
      solver::simple_gmres_euler<double, resource::host>(dt, grid, xz, b, int(size) + 1, 1,
                                                         1e-12);
    }
    catch (std::invalid_argument const &)
    {
      threw = 1;
    }
    out.put("gmres_restart_throws", {double(threw)});

    // dense simple_gmres: the 3x3 gold of src/solver_tests.cpp:141-213
    // (column-major A), with the point Jacobi preconditioner
    const std::vector<double> A_gold = {3.383861628748717e+00, 3.210305545769361e+00,
                                        1.723479266939425e+00, 1.113343240310116e-02,
                                        3.412141162288144e+00, 1.710451084172946e+00,
                                        2.920740795411032e+00, 3.934494120167269e+00,
                                        4.450671104482062e+00};
    const std::vector<double> precond = {3.383861628748717e+00, 0, 0, 0, 3.412141162288144e+00,
                                         0, 0, 0, 4.450671104482062e+00};
    const std::vector<double> b_gold = {2.084406360034887e-01, 6.444769305362776e-01,
                                        3.687335330031937e-01};
    std::vector<double> xg(3, 0.0), xgp(3, 0.0);
    auto g1 = solver::simple_gmres(A_gold, xg, b_gold, std::vector<double>(), 3, 3,
                                   std::numeric_limits<double>::epsilon());
    auto g2 = solver::simple_gmres(A_gold, xgp, b_gold, precond, 3, 3,
                                   std::numeric_limits<double>::epsilon());
    out.put("dense_gmres", xg);
    out.put("dense_gmres_jacobi", xgp);
    out.put("dense_gmres_errors", {double(g1.error), double(g2.error)});

    hip_check(hipFree(xd));
    hip_check(hipFree(yd));
    hip_check(hipFree(bd));
    std::fclose(fout);
  }
  catch (std::exception const &e)
  {
    std::fprintf(stderr, "consumer: %s\n", e.what());
    return 1;
  }
  return 0;
}
