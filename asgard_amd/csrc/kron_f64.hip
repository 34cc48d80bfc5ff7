// Kronmult kernels instantiated for double.
#include "kron_impl.hpp"
#include "kron_mfma.hpp"

namespace akm
{
template<>
int launch_kronmult<double>(int D, int n, int mode, kron_args<double> const &a, int num_rows, int waves,
                        int force_generic, hipStream_t stream)
{
  if (a.family == 1 && !force_generic && mode == mode_grid && num_rows > 0)
  {
    if (D == 3 && n == 3 && launch_mfma_d3n3(a, num_rows, waves, stream))
      return check_launch("kronmult (matrix cores)");
    if (D == 6 && n == 2 && launch_mfma_d6n2(a, num_rows, waves, stream))
      return check_launch("kronmult (matrix cores)");
    if (D == 4 && n == 2 && launch_mfma_d4n2(a, num_rows, waves, stream))
      return check_launch("kronmult (matrix cores)");
  }
  return launch_kronmult_impl<double>(D, n, mode, a, num_rows, waves, force_generic, stream);
}

bool mfma_family_applies(int D, int n, int mode, int T, bool packed)
{
  return ((D == 3 && n == 3) || (D == 4 && n == 2) || (D == 6 && n == 2)) &&
         mode == mode_grid && T >= 1 && T <= mfma_max_terms && packed;
}

template<>
int launch_scale<double>(int64_t size, double beta, double *y, double gamma, const double *xg, hipStream_t stream)
{
  return launch_scale_impl<double>(size, beta, y, gamma, xg, stream);
}
} // namespace akm

namespace akm
{
int launch_rows_batch(int D, int n, int waves, const kron_args<double> *args,
                      const batch_system *sys, const int2 *work, int nwork, int grid,
                      const int *jcol, const int *active, int phase, hipStream_t stream)
{
  return launch_rows_batch_impl(D, n, waves, args, sys, work, nwork, grid, jcol, active, phase,
                                stream);
}
} // namespace akm
