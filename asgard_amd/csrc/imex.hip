// One IMEX time step with every vector on the GPU (imex_advance,
// src/time_advance.cpp:760-1000, with imex_resrc = resource::device).
//
// Two SSP-RK2 stages, each an explicit apply of the explicit-term operator
// followed by an implicit backward-Euler solve of the implicit-term
// operator:
//   f1s = f0 + dt E f0;            solve (I - dt B) f1 = f1s      (guess f1s)
//   f2  = f1 + dt E f1;  f2s = (f0 + f2) / 2;
//                                  solve (I - dt/2 B) f = f2s     (guess f1)
// A caller hook runs before each explicit apply and each solve with the
// device vector about to be used, where the reference recomputes moments,
// regenerates coefficients and resets the operator matrices
// (calculate_moments, generate_all_coefficients, reset_coefficients); the
// hook updates the handles, e.g. through akm_update_operators.  Without
// the collision operator the solves are skipped, as in the reference's
// non-collision branch.
#include "akm_internal.hpp"
#include "matrix.hpp"

namespace akm
{
constexpr int imex_threads = 256;

//! y = a*x + b*y (y not read when b == 0)
__global__ void __launch_bounds__(imex_threads)
    imex_axpby(int64_t n, double a, const double *x, double b, double *y)
{
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x)
    y[i] = (b == 0) ? a * x[i] : a * x[i] + b * y[i];
}

static int imex_blocks(int64_t n)
{
  const int64_t b = (n + imex_threads - 1) / imex_threads;
  return static_cast<int>(b < 4096 ? (b > 0 ? b : 1) : 4096);
}

static int axpby(hipStream_t s, int64_t n, double a, const double *x, double b, double *y)
{
  imex_axpby<<<imex_blocks(n), imex_threads, 0, s>>>(n, a, x, b, y);
  return check_launch("imex vector update");
}
} // namespace akm

using namespace akm;

akm_status akm_imex_step(akm_matrix explicit_op, akm_matrix implicit_op, double dt, void *f,
                         int restart, int max_iter, double tol, int collision,
                         akm_imex_hook hook, void *user, double *errors, int *iterations)
{
  if (explicit_op == nullptr || implicit_op == nullptr || f == nullptr)
    return set_error(AKM_ERR_INVALID, "imex step: null argument");
  if (explicit_op->prec != AKM_F64 || implicit_op->prec != AKM_F64)
    return set_error(AKM_ERR_INVALID, "imex step: FP64 handles only");
  const int64_t n = implicit_op->output_size();
  if (explicit_op->rows != explicit_op->cols || implicit_op->rows != implicit_op->cols ||
      explicit_op->output_size() != n || n == 0)
    return set_error(AKM_ERR_INVALID, "imex step: the two operators must be square and of one size");
  akm_matrix_s &B = *implicit_op;
  hipStream_t s = B.stream;
  // f0, the explicit product, the stage-1 vector, the stage-1 solution
  AKM_TRY(B.imex_ws.allocate(4 * n));
  double *f0 = B.imex_ws.data(), *fx = f0 + n, *f1 = fx + n, *f1_out = f1 + n;
  double *fv = static_cast<double *>(f);
  const size_t bytes = size_t(n) * sizeof(double);
  int status = AKM_SUCCESS;
  auto call_hook = [&](int event, const double *v) -> int {
    if (hook == nullptr)
      return AKM_SUCCESS;
    AKM_HIP(hipStreamSynchronize(s));
    if (hook(user, event, v) != 0)
      return set_error(AKM_ERR_INVALID, "imex step: the hook failed at event %d", event);
    return AKM_SUCCESS;
  };
  auto explicit_apply = [&](const double *x) -> int {
    AKM_HIP(hipStreamSynchronize(s));
    AKM_TRY(apply_device(*explicit_op, 1.0, x, 0.0, fx, 0.0, nullptr));
    AKM_HIP(hipStreamSynchronize(explicit_op->stream));
    return AKM_SUCCESS;
  };
  auto solve = [&](int stage, double sdt, double *x, const double *rhs) -> int {
    double err = 0;
    int its = 0;
    const int rc = akm_gmres_euler(implicit_op, sdt, x, rhs, restart, max_iter, tol, 1, &err,
                                   &its);
    if (errors != nullptr)
      errors[stage] = err;
    if (iterations != nullptr)
      iterations[stage] = its;
    if (rc == AKM_ERR_NOT_CONVERGED)
      status = AKM_ERR_NOT_CONVERGED;
    else
      AKM_TRY(rc);
    return AKM_SUCCESS;
  };
  if (errors != nullptr)
    errors[0] = errors[1] = 0;
  if (iterations != nullptr)
    iterations[0] = iterations[1] = 0;

  // stage 1: f1s = f0 + dt E f0, then (I - dt B) f1 = f1s
  AKM_HIP(hipMemcpyAsync(f0, fv, bytes, hipMemcpyDeviceToDevice, s));
  AKM_TRY(call_hook(AKM_IMEX_EXPLICIT_1, fv));
  AKM_TRY(explicit_apply(fv));
  AKM_TRY(axpby(s, n, dt, fx, 1.0, fv));
  AKM_TRY(call_hook(AKM_IMEX_IMPLICIT_1, fv));
  AKM_HIP(hipMemcpyAsync(f1, fv, bytes, hipMemcpyDeviceToDevice, s));
  if (collision)
  {
    AKM_TRY(solve(0, dt, f1, fv));
    AKM_HIP(hipMemcpyAsync(f1_out, f1, bytes, hipMemcpyDeviceToDevice, s));
  }

  // stage 2: f2 = f1 + dt E f1, f2s = (f0 + f2) / 2, then (I - dt/2 B) f = f2s
  AKM_TRY(call_hook(AKM_IMEX_EXPLICIT_2, f1));
  AKM_TRY(explicit_apply(f1));
  AKM_TRY(axpby(s, n, dt, fx, 1.0, f1));
  AKM_TRY(axpby(s, n, 1.0, f0, 1.0, f1));
  AKM_TRY(axpby(s, n, 0.5, f1, 0.0, fv));
  AKM_TRY(call_hook(AKM_IMEX_IMPLICIT_2, fv));
  if (collision)
  {
    AKM_TRY(solve(1, 0.5 * dt, f1_out, fv));
    AKM_HIP(hipMemcpyAsync(fv, f1_out, bytes, hipMemcpyDeviceToDevice, s));
  }
  AKM_HIP(hipStreamSynchronize(s));
  return status;
}
