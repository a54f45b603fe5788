// CSR product of an assembled system matrix (the matrix-based comparison
// path, CG/cells_batched/application.cc:1964-1983).
//
// CSR-stream scheme (Greathouse and Daga, "Efficient sparse matrix-vector
// multiplication on GPUs using the CSR storage format", SC 2014): one
// workgroup of 256 lanes takes a block of consecutive rows holding at most
// 2048 entries (row blocks fixed at assembly).  Lane t loads entries
// e0 + t + 256 u (u < 8) of the block's contiguous run, so every load
// instruction of a wavefront reads 64 consecutive columns and values; the
// eight column/value loads, then the eight x gathers, are in flight together.
// The products go to LDS; after one barrier, G lanes (G = 8, 4, 2, 1 for
// blocks of at most 32, 64, 128, 256 rows) sum each row: lane s of a group
// takes entries s, s + G, ... in order and a fixed xor butterfly combines the
// G partial sums.  G follows from the fixed row blocks, so the result does
// not depend on scheduling.  The matrix is streamed once with non-temporal
// loads, so that the cached lines are those of the gathered x.  Blocks are
// mapped XCD-contiguously: rows that are close in the DoF numbering, and the
// x entries they gather, share an L2.
#include <hip/hip_runtime.h>

#include "matrix.h"

namespace smf {

namespace {

constexpr int kSpmvThreads = 256;
constexpr int kPerLane = kSpmvBlockEntries / kSpmvThreads;
static_assert(kPerLane * kSpmvThreads == kSpmvBlockEntries, "row-block entries per lane");
static_assert(kSpmvBlockRows <= kSpmvThreads, "one lane per row of a block");

// ADD: y += A x (the prolongation of a multigrid level into its iterate)
template <typename T, bool ADD>
__global__ __launch_bounds__(kSpmvThreads) void spmv_stream_kernel(int64_t nblocks, const int32_t* __restrict__ blk,
                                                                   const int64_t* __restrict__ ptr,
                                                                   const int32_t* __restrict__ col,
                                                                   const T* __restrict__ val,
                                                                   const T* __restrict__ x, T* __restrict__ y) {
  __shared__ T prod[kSpmvBlockEntries];
  const int64_t b = xcd_block(blockIdx.x, nblocks);
  const int t = threadIdx.x;
  const int32_t r0 = blk[b], r1 = blk[b + 1];
  const int nrows = r1 - r0;  // 1..256
  const int G = nrows <= 32 ? 8 : (nrows <= 64 ? 4 : (nrows <= 128 ? 2 : 1));
  const int lr = t / G, sub = t % G;
  const bool has_row = lr < nrows;
  const int64_t rr = has_row ? r0 + lr : r0;
  const int64_t e0 = ptr[r0], e1 = ptr[r1];
  // this lane's row bounds are loaded together with the entries
  const int64_t k0 = ptr[rr], k1 = ptr[rr + 1];
  int32_t c[kPerLane];
  T v[kPerLane], xv[kPerLane];
  // entries past the block's run (or a block without entries) load nothing:
  // column 0 (x has at least one row) and value 0
#pragma unroll
  for (int u = 0; u < kPerLane; ++u) {
    const int64_t e = e0 + t + u * kSpmvThreads;
    if (e < e1) {
      c[u] = __builtin_nontemporal_load(col + e);
      v[u] = __builtin_nontemporal_load(val + e);
    } else {
      c[u] = 0;
      v[u] = T(0);
    }
  }
#pragma unroll
  for (int u = 0; u < kPerLane; ++u) xv[u] = x[c[u]];
#pragma unroll
  for (int u = 0; u < kPerLane; ++u)
    if (e0 + t + u * kSpmvThreads < e1) prod[t + u * kSpmvThreads] = v[u] * xv[u];
  __syncthreads();
  T s = T(0);
  if (has_row)
    for (int64_t k = k0 - e0 + sub; k < k1 - e0; k += G) s += prod[k];
  // every lane takes part (G is uniform over the workgroup)
  for (int m = 1; m < G; m <<= 1) s += __shfl_xor(s, m);
  if (has_row && sub == 0) y[rr] = ADD ? y[rr] + s : s;
}

}  // namespace

template <typename T>
void launch_spmv_t(const SysMatrix& A, T* dst, const T* src, bool add, hipStream_t s) {
  const int64_t nb = A.n_blocks();
  auto kern = add ? spmv_stream_kernel<T, true> : spmv_stream_kernel<T, false>;
  hipLaunchKernelGGL(kern, dim3((unsigned)nb), dim3(kSpmvThreads), 0, s, nb, A.d_blk.as<int32_t>(),
                     A.d_ptr.as<int64_t>(), A.d_col.as<int32_t>(), A.d_val.as<T>(), src, dst);
}

void launch_spmv(const SysMatrix& A, void* dst, const void* src, hipStream_t s, bool add) {
  if (!A.on_device) throw std::runtime_error("system matrix has no device data (host_only operator)");
  const int64_t nb = A.n_blocks();
  if (nb == 0) return;
  if (nb >= (int64_t(1) << 31)) throw std::runtime_error("system matrix: too many row blocks");
  if (A.precision == MF_F64)
    launch_spmv_t<double>(A, static_cast<double*>(dst), static_cast<const double*>(src), add, s);
  else
    launch_spmv_t<float>(A, static_cast<float*>(dst), static_cast<const float*>(src), add, s);
  SMF_HIP(hipGetLastError());
}

}  // namespace smf
