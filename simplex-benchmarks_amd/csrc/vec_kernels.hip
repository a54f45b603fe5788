// Element-wise vector kernels used by the operator entry points.
#include <hip/hip_runtime.h>

#include "op.h"

namespace smf {

namespace {
// inv_i = 1/d_i if |d_i| > 1e-10 else 1 (CG/cells_batched/application_hp.cc:1404-1407)
template <typename T>
__global__ void inverse_diagonal_kernel(int64_t n, T* __restrict__ d) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    const T v = d[i];
    d[i] = fabs((double)v) > 1e-10 ? T(1) / v : T(1);
  }
}
}  // namespace

void launch_invert_diagonal(int precision, int64_t n, void* d, hipStream_t s) {
  if (n <= 0) return;
  const dim3 grid((unsigned)((n + 255) / 256));
  if (precision == MF_F64)
    hipLaunchKernelGGL(inverse_diagonal_kernel<double>, grid, dim3(256), 0, s, n, static_cast<double*>(d));
  else
    hipLaunchKernelGGL(inverse_diagonal_kernel<float>, grid, dim3(256), 0, s, n, static_cast<float*>(d));
  SMF_HIP(hipGetLastError());
}

}  // namespace smf
