// Kronmult kernels instantiated for double.
#include "kron_impl.hpp"

namespace akm
{
template<>
int launch_kronmult<double>(int D, int n, int mode, kron_args<double> const &a, int num_rows, int waves,
                        int force_generic, hipStream_t stream)
{
  return launch_kronmult_impl<double>(D, n, mode, a, num_rows, waves, force_generic, stream);
}

template<>
int launch_scale<double>(int64_t size, double beta, double *y, double gamma, const double *xg, hipStream_t stream)
{
  return launch_scale_impl<double>(size, beta, y, gamma, xg, stream);
}
} // namespace akm
