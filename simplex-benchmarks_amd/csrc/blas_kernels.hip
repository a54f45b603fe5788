// Vector, transfer and reduction kernels of the multigrid solver (gfx950).
//
// Reductions are deterministic: a fixed grid of 256 blocks writes one FP64
// partial sum per block (block-local tree in LDS), and a single-block kernel
// adds the partials in a fixed order -- two runs give bit-identical dots.
#include <hip/hip_runtime.h>

#include "blas.h"

namespace smf {

namespace {

constexpr int kBlock = 256;
constexpr int kRedBlocks = 1024;

inline unsigned grid_for(int64_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }

template <typename T>
__global__ void axpby_kernel(int64_t n, double a, const T* __restrict__ x, double b, T* __restrict__ y) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = T(a * (double)x[i] + b * (double)y[i]);
}

// y = a*x + b*y computed in T (keeps FP32 arithmetic in FP32)
template <typename T>
__global__ void axpby_t_kernel(int64_t n, T a, const T* __restrict__ x, T b, T* __restrict__ y) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = a * x[i] + b * y[i];
}

template <typename TO, typename TI>
__global__ void convert_kernel(int64_t n, const TI* __restrict__ x, TO* __restrict__ y) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = TO(x[i]);
}

template <typename T>
__global__ void scale_kernel(int64_t n, const T* __restrict__ d, const T* __restrict__ x, T* __restrict__ y) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = d[i] * x[i];
}

template <typename T>
__global__ void fill_kernel(int64_t n, T v, T* __restrict__ y) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = v;
}

// y = (sa S[ia]) x + S[ib] y (index < 0: 1), the arithmetic of axpby_kernel
__global__ void axpby_dev_kernel(int64_t n, const double* __restrict__ S, int ia, double sa,
                                 const double* __restrict__ x, int ib, double* __restrict__ y) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const double a = ia >= 0 ? sa * S[ia] : sa, b = ib >= 0 ? S[ib] : 1.0;
  if (i < n) y[i] = a * x[i] + b * y[i];
}

__global__ void krylov_scalars_kernel(double* S, int mode) {
  if (threadIdx.x != 0) return;
  if (mode == 0) {
    S[kS_ALPHA] = S[kS_RZ] / S[kS_PQ];
  } else {
    S[kS_BETA] = mode == 1 ? S[kS_RZN] / S[kS_RZ] : -S[kS_ALPHA] * S[kS_ZQ] / S[kS_RZ];
    S[kS_RZ] = S[kS_RZN];
  }
}

// partial[b] = sum over block b's grid-stride range of x*y (FP64)
template <typename T>
__global__ __launch_bounds__(kBlock) void dot_partial_kernel(int64_t n, const T* __restrict__ x,
                                                             const T* __restrict__ y,
                                                             double* __restrict__ partial) {
  __shared__ double s[kBlock];
  // four accumulators over strides i, i + g, i + 2g, i + 3g keep four loads
  // of each vector in flight; fixed grid and order: deterministic
  const int64_t g = (int64_t)gridDim.x * kBlock;
  int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
  for (; i + 3 * g < n; i += 4 * g) {
    a0 += (double)x[i] * (double)y[i];
    a1 += (double)x[i + g] * (double)y[i + g];
    a2 += (double)x[i + 2 * g] * (double)y[i + 2 * g];
    a3 += (double)x[i + 3 * g] * (double)y[i + 3 * g];
  }
  for (; i < n; i += g) a0 += (double)x[i] * (double)y[i];
  s[threadIdx.x] = (a0 + a1) + (a2 + a3);
  __syncthreads();
  for (int w = kBlock / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = s[0];
}

__global__ __launch_bounds__(kBlock) void reduce_final_kernel(int nparts, const double* __restrict__ partial,
                                                              double* __restrict__ out) {
  __shared__ double s[kBlock];
  double acc = 0.0;
  for (int i = threadIdx.x; i < nparts; i += kBlock) acc += partial[i];
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int w = kBlock / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = s[0];
}

// y = A x for a CSR matrix (one thread per row, fixed summation order);
// four entries per step: their column indices and values, then the four x
// gathers, are in flight together (clamped indices, masked sums)
template <typename T>
__global__ void csr_kernel(int64_t nrows, const int64_t* __restrict__ ptr, const int32_t* __restrict__ col,
                           const T* __restrict__ val, const T* __restrict__ x, T* __restrict__ y, int add) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nrows) return;
  const int64_t k0 = ptr[r], k1 = ptr[r + 1];
  const T y0 = add ? y[r] : T(0);
  T s = T(0);
  for (int64_t k = k0; k < k1; k += 4) {
    int32_t c[4];
    T v[4], xv[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int64_t kk = k + u < k1 ? k + u : k1 - 1;
      c[u] = col[kk];
      v[u] = val[kk];
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) xv[u] = x[c[u]];
#pragma unroll
    for (int u = 0; u < 4; ++u)
      if (k + u < k1) s += v[u] * xv[u];
  }
  y[r] = add ? y0 + s : s;
}

// Chebyshev step (three-term recurrence on the Jacobi-preconditioned
// residual): d = c1*d + c2*Dinv*(b - Ax); x += d
template <typename T>
__global__ void cheby_step_kernel(int64_t n, T c1, T c2, const T* __restrict__ dinv, const T* __restrict__ b,
                                  const T* __restrict__ ax, T* __restrict__ d, T* __restrict__ x, int first) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const T r = b[i] - ax[i];
  const T di = first ? c2 * dinv[i] * r : c1 * d[i] + c2 * dinv[i] * r;
  d[i] = di;
  x[i] += di;
}

// first Chebyshev step from x = 0: d = c2 Dinv b, x = d
template <typename T>
__global__ void cheby_zero_kernel(int64_t n, T c2, const T* __restrict__ dinv, const T* __restrict__ b,
                                  T* __restrict__ d, T* __restrict__ x) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const T di = c2 * dinv[i] * b[i];
  d[i] = di;
  x[i] = di;
}

// y = M x for a dense row-major n x n matrix (FP64): one wave per row, lanes
// stride the columns (coalesced), fixed-order butterfly sum (deterministic)
__global__ __launch_bounds__(256) void dense_gemv_kernel(int nf, int n, const double* __restrict__ M,
                                                         const double* __restrict__ x, double* __restrict__ y) {
  const int r = blockIdx.x * 4 + threadIdx.x / 64, lane = threadIdx.x % 64;
  if (r >= n) return;  // whole waves exit together
  if (r >= nf) {       // identity rows (Dirichlet DoFs of the coarsest level)
    if (lane == 0) y[r] = x[r];
    return;
  }
  const double* row = M + (int64_t)r * nf;
  // four independent accumulators keep four loads in flight per lane
  double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
  int c = lane;
  for (; c + 192 < nf; c += 256) {
    s0 += row[c] * x[c];
    s1 += row[c + 64] * x[c + 64];
    s2 += row[c + 128] * x[c + 128];
    s3 += row[c + 192] * x[c + 192];
  }
  for (; c < nf; c += 64) s0 += row[c] * x[c];
  double s = (s0 + s1) + (s2 + s3);
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  if (lane == 0) y[r] = s;
}

}  // namespace

template <typename T>
void blas_axpby(int64_t n, double a, const T* x, double b, T* y, hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(axpby_kernel<T>, dim3(grid_for(n)), dim3(kBlock), 0, s, n, a, x, b, y);
}
template <typename T>
void blas_axpby_t(int64_t n, T a, const T* x, T b, T* y, hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(axpby_t_kernel<T>, dim3(grid_for(n)), dim3(kBlock), 0, s, n, a, x, b, y);
}
template <typename TO, typename TI>
void blas_convert(int64_t n, const TI* x, TO* y, hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL((convert_kernel<TO, TI>), dim3(grid_for(n)), dim3(kBlock), 0, s, n, x, y);
}
template <typename T>
void blas_scale(int64_t n, const T* d, const T* x, T* y, hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(scale_kernel<T>, dim3(grid_for(n)), dim3(kBlock), 0, s, n, d, x, y);
}
template <typename T>
void blas_fill(int64_t n, T v, T* y, hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(fill_kernel<T>, dim3(grid_for(n)), dim3(kBlock), 0, s, n, v, y);
}
template <typename T>
void blas_dot(int64_t n, const T* x, const T* y, double* partial, double* out, hipStream_t s) {
  hipLaunchKernelGGL(dot_partial_kernel<T>, dim3(kRedBlocks), dim3(kBlock), 0, s, n, x, y, partial);
  hipLaunchKernelGGL(reduce_final_kernel, dim3(1), dim3(kBlock), 0, s, kRedBlocks, partial, out);
}
template <typename T>
void blas_csr(int64_t nrows, const int64_t* ptr, const int32_t* col, const T* val, const T* x, T* y, bool add,
              hipStream_t s) {
  if (nrows > 0)
    hipLaunchKernelGGL(csr_kernel<T>, dim3(grid_for(nrows)), dim3(kBlock), 0, s, nrows, ptr, col, val, x, y,
                       add ? 1 : 0);
}
template <typename T>
void blas_cheby_step(int64_t n, T c1, T c2, const T* dinv, const T* b, const T* ax, T* d, T* x, bool first,
                     hipStream_t s) {
  if (n > 0)
    hipLaunchKernelGGL(cheby_step_kernel<T>, dim3(grid_for(n)), dim3(kBlock), 0, s, n, c1, c2, dinv, b, ax, d, x,
                       first ? 1 : 0);
}
template <typename T>
void blas_cheby_zero(int64_t n, T c2, const T* dinv, const T* b, T* d, T* x, hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(cheby_zero_kernel<T>, dim3(grid_for(n)), dim3(kBlock), 0, s, n, c2, dinv, b, d, x);
}
void blas_axpby_dev(int64_t n, const double* S, int ia, double sa, const double* x, int ib, double* y, hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(axpby_dev_kernel, dim3(grid_for(n)), dim3(kBlock), 0, s, n, S, ia, sa, x, ib, y);
}
void blas_krylov_scalars(double* S, int mode, hipStream_t s) {
  hipLaunchKernelGGL(krylov_scalars_kernel, dim3(1), dim3(64), 0, s, S, mode);
}
void blas_dense_gemv(int nf, int n, const double* M, const double* x, double* y, hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(dense_gemv_kernel, dim3((n + 3) / 4), dim3(256), 0, s, nf, n, M, x, y);
}
int blas_reduction_parts() { return kRedBlocks; }

#define SMF_INST(T)                                                                                   \
  template void blas_cheby_zero<T>(int64_t, T, const T*, const T*, T*, T*, hipStream_t);           \
  template void blas_axpby<T>(int64_t, double, const T*, double, T*, hipStream_t);                 \
  template void blas_axpby_t<T>(int64_t, T, const T*, T, T*, hipStream_t);                        \
  template void blas_scale<T>(int64_t, const T*, const T*, T*, hipStream_t);                       \
  template void blas_fill<T>(int64_t, T, T*, hipStream_t);                                         \
  template void blas_dot<T>(int64_t, const T*, const T*, double*, double*, hipStream_t);           \
  template void blas_csr<T>(int64_t, const int64_t*, const int32_t*, const T*, const T*, T*, bool, \
                            hipStream_t);                                                          \
  template void blas_cheby_step<T>(int64_t, T, T, const T*, const T*, const T*, T*, T*, bool, hipStream_t);
SMF_INST(double)
SMF_INST(float)
template void blas_convert<float, double>(int64_t, const double*, float*, hipStream_t);
template void blas_convert<double, float>(int64_t, const float*, double*, hipStream_t);
template void blas_convert<double, double>(int64_t, const double*, double*, hipStream_t);
template void blas_convert<float, float>(int64_t, const float*, float*, hipStream_t);

}  // namespace smf
