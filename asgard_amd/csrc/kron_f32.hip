// Kronmult kernels instantiated for float.
#include "kron_impl.hpp"

namespace akm
{
template<>
int launch_kronmult<float>(int D, int n, int mode, kron_args<float> const &a, int num_rows, int waves,
                        int force_generic, hipStream_t stream)
{
  return launch_kronmult_impl<float>(D, n, mode, a, num_rows, waves, force_generic, stream);
}

template<>
int launch_scale<float>(int64_t size, float beta, float *y, float gamma, const float *xg, hipStream_t stream)
{
  return launch_scale_impl<float>(size, beta, y, gamma, xg, stream);
}
} // namespace akm
