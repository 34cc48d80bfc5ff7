// Right preconditioners of the flexible GMRES: z = P(v) on device vectors.
// P may be nonlinear (the learned preconditioner normalises per cell), which
// flexible GMRES allows because it keeps every z_j.
#pragma once

#include "akm_internal.hpp"

struct akm_precond_s
{
  virtual ~akm_precond_s() = default;
  //! Length of the vectors the preconditioner acts on.
  virtual int64_t size() const = 0;
  //! z = P(v), FP64 device vectors, enqueued on `stream`.
  virtual int apply(const double *v, double *z, hipStream_t stream) = 0;
};
