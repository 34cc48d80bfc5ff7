// Internal helpers shared by the asgard_amd translation units: error
// reporting across the C ABI and owning device buffers.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <utility>

#include "../../include/asgard_amd.h"

namespace akm
{
//! Records a formatted message for akm_last_error() and returns `code`.
int set_error(int code, const char *fmt, ...) __attribute__((format(printf, 2, 3)));

//! Returns AKM_SUCCESS or records the pending HIP launch/runtime error.
int check_launch(const char *what);

#define AKM_HIP(call)                                                                              \
  do                                                                                               \
  {                                                                                                \
    hipError_t akm_e_ = (call);                                                                    \
    if (akm_e_ != hipSuccess)                                                                      \
      return ::akm::set_error(AKM_ERR_HIP, "%s failed: %s (%s:%d)", #call,                         \
                              hipGetErrorString(akm_e_), __FILE__, __LINE__);                      \
  } while (0)

#define AKM_TRY(call)                                                                              \
  do                                                                                               \
  {                                                                                                \
    int akm_rc_ = (call);                                                                          \
    if (akm_rc_ != AKM_SUCCESS)                                                                    \
      return akm_rc_;                                                                              \
  } while (0)

//! Owning device allocation (hipMalloc / hipFree).
template<typename T>
class dev_buffer
{
public:
  dev_buffer() = default;
  ~dev_buffer() { reset(); }
  dev_buffer(dev_buffer const &) = delete;
  dev_buffer &operator=(dev_buffer const &) = delete;
  dev_buffer(dev_buffer &&o) noexcept : ptr_(o.ptr_), size_(o.size_)
  {
    o.ptr_ = nullptr;
    o.size_ = 0;
  }
  dev_buffer &operator=(dev_buffer &&o) noexcept
  {
    if (this != &o)
    {
      reset();
      ptr_ = std::exchange(o.ptr_, nullptr);
      size_ = std::exchange(o.size_, 0);
    }
    return *this;
  }

  //! (Re)allocates `count` elements; contents undefined.
  int allocate(int64_t count)
  {
    if (count == size_ && ptr_ != nullptr)
      return AKM_SUCCESS;
    reset();
    if (count <= 0)
      return AKM_SUCCESS;
    hipError_t e = hipMalloc(reinterpret_cast<void **>(&ptr_), sizeof(T) * size_t(count));
    if (e != hipSuccess)
    {
      ptr_ = nullptr;
      return set_error(AKM_ERR_MEMORY, "hipMalloc of %lld bytes failed: %s",
                       static_cast<long long>(sizeof(T) * count), hipGetErrorString(e));
    }
    size_ = count;
    return AKM_SUCCESS;
  }
  //! Allocates and copies `count` host elements.
  int upload(const T *host, int64_t count, hipStream_t stream = nullptr)
  {
    AKM_TRY(allocate(count));
    if (count > 0)
    {
      AKM_HIP(hipMemcpyAsync(ptr_, host, sizeof(T) * size_t(count), hipMemcpyHostToDevice,
                             stream));
      AKM_HIP(hipStreamSynchronize(stream));
    }
    return AKM_SUCCESS;
  }
  void reset()
  {
    if (ptr_ != nullptr)
      (void)hipFree(ptr_);
    ptr_ = nullptr;
    size_ = 0;
  }
  T *data() const { return ptr_; }
  int64_t size() const { return size_; }

private:
  T *ptr_ = nullptr;
  int64_t size_ = 0;
};

template<typename P>
struct kron_args;

//! One workgroup of `waves` wavefronts per row (rows in a.row_order).
template<typename P>
int launch_kronmult(int D, int n, int mode, kron_args<P> const &a, int num_rows, int waves,
                    int force_generic, hipStream_t stream);
template<typename P>
int launch_scale(int64_t size, P beta, P *y, P gamma, const P *xg, hipStream_t stream);
//! True when the sum-factorised matrix-core kernel (kernel family 1) can run
//! an FP64 operator of this shape; `packed` = packed cell words present.
bool mfma_family_applies(int D, int n, int mode, int T, bool packed);

struct batch_system;
//! Batched sparse-grid FP64 kronmult over a (system, row) work list of
//! systems with the same (D, n); `grid` persistent workgroups.
int launch_rows_batch(int D, int n, int waves, const kron_args<double> *args,
                      const batch_system *sys, const int2 *work, int nwork, int grid,
                      const int *jcol, const int *active, int phase, hipStream_t stream);

} // namespace akm
