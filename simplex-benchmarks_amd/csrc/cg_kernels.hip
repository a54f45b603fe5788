// Continuous-Galerkin cell kernels for gfx950.
//
// One workgroup (256 threads, 4 wave64) processes one patch of up to
// 256*CPT cells of the hierarchical reordering:
//   1. the patch's DoF values are staged in LDS (owned DoFs are a contiguous
//      global range -> coalesced loads; ghost DoFs through an index list);
//   2. every thread evaluates CPT cells: gradients at the quadrature points
//      (interpolation with the reference shape-gradient matrix), the affine
//      quadrature-point operator w_q |det J| J^-1 J^-T, and the transposed
//      interpolation (integration) -- all in registers with FMA;
//   3. cell results are summed into an LDS accumulator colour by colour (cells
//      of one colour share no DoF), so the sums are deterministic;
//   4. owned DoFs are stored, ghost DoFs added to the value written by the
//      patch colours launched earlier.
// Patch colours are separate launches in a fixed order, so the whole vmult is
// bit-reproducible.
//
// Reference: do_cell_integral_masked_gather, CG/cells_batched/application.cc:
// 1236-1416 (masked gather 1272-1306, interpolation 1314-1319, affine D_e
// 1333-1360, integration 1381-1386, masked scatter-add 1388-1412, Dirichlet
// identity rows 1173-1175).
#include <hip/hip_runtime.h>

#include <stdexcept>

#include "op.h"

namespace smf {

namespace {

constexpr int kThreads = 256;
constexpr unsigned short kInvalid = 0xFFFF;

template <int P>
struct CGShape {
  static constexpr int NL = (P + 1) * (P + 2) * (P + 3) / 6;
  static constexpr int NQ = P == 1 ? 1 : (P == 2 ? 4 : 11);
};

// MODE 0: y = A_K u;  MODE 1: y_i = (A_K)_ii
template <int P, typename T, int CPT, int MODE>
__global__ __launch_bounds__(kThreads) void cg_patch_kernel(
    const int32_t* __restrict__ color_patches, int32_t patch_base,
    const int32_t* __restrict__ cell_start, const int32_t* __restrict__ own_start,
    const int32_t* __restrict__ ghost_start, const int32_t* __restrict__ ghost_idx,
    const int32_t* __restrict__ icol_ptr, const uint16_t* __restrict__ icol_off,
    const uint16_t* __restrict__ lidx, int64_t ncells_total, const T* __restrict__ metric,
    const T* __restrict__ shape, const T* __restrict__ src, T* __restrict__ dst) {
  constexpr int NL = CGShape<P>::NL, NQ = CGShape<P>::NQ;
  extern __shared__ unsigned char smem_raw[];
  T* su = reinterpret_cast<T*>(smem_raw);
  const int tid = threadIdx.x;
  const int patch = color_patches[patch_base + blockIdx.x];
  const int c0 = cell_start[patch], nc = cell_start[patch + 1] - c0;
  const int o0 = own_start[patch], no = own_start[patch + 1] - o0;
  const int g0 = ghost_start[patch], ng = ghost_start[patch + 1] - g0;
  const int nloc = no + ng;
  T* sy = su + nloc;

  // 1. stage the patch DoFs
  for (int j = tid; j < nloc; j += kThreads) {
    if (MODE == 0) su[j] = j < no ? src[o0 + j] : src[ghost_idx[g0 + j - no]];
    sy[j] = T(0);
  }
  __syncthreads();

  // shape tables: G[(i*NQ + q)*3 + d] then GW = w_q G at offset NL*NQ*3
  const T* G = shape;
  const T* GW = shape + NL * NQ * 3;

  // 2. cell evaluation
  T y[CPT][NL];
  uint16_t li[CPT][NL];
#pragma unroll
  for (int k = 0; k < CPT; ++k) {
    const int lc = tid + k * kThreads;
#pragma unroll
    for (int i = 0; i < NL; ++i) y[k][i] = T(0);
    if (lc < nc) {
      const int64_t cell = c0 + lc;
      T u[NL];
#pragma unroll
      for (int i = 0; i < NL; ++i) {
        li[k][i] = lidx[(int64_t)i * ncells_total + cell];
        if (MODE == 0) u[i] = li[k][i] != kInvalid ? su[li[k][i]] : T(0);
      }
      const T m0 = metric[0 * ncells_total + cell], m1 = metric[1 * ncells_total + cell],
              m2 = metric[2 * ncells_total + cell], m3 = metric[3 * ncells_total + cell],
              m4 = metric[4 * ncells_total + cell], m5 = metric[5 * ncells_total + cell];
      if (MODE == 0) {
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
          T gx = T(0), gy = T(0), gz = T(0);
#pragma unroll
          for (int i = 0; i < NL; ++i) {
            gx += G[(i * NQ + q) * 3 + 0] * u[i];
            gy += G[(i * NQ + q) * 3 + 1] * u[i];
            gz += G[(i * NQ + q) * 3 + 2] * u[i];
          }
          const T tx = m0 * gx + m1 * gy + m2 * gz;
          const T ty = m1 * gx + m3 * gy + m4 * gz;
          const T tz = m2 * gx + m4 * gy + m5 * gz;
#pragma unroll
          for (int i = 0; i < NL; ++i)
            y[k][i] += GW[(i * NQ + q) * 3 + 0] * tx + GW[(i * NQ + q) * 3 + 1] * ty +
                       GW[(i * NQ + q) * 3 + 2] * tz;
        }
      } else {
#pragma unroll
        for (int i = 0; i < NL; ++i)
#pragma unroll
          for (int q = 0; q < NQ; ++q) {
            const T gx = G[(i * NQ + q) * 3 + 0], gy = G[(i * NQ + q) * 3 + 1],
                    gz = G[(i * NQ + q) * 3 + 2];
            const T tx = m0 * gx + m1 * gy + m2 * gz;
            const T ty = m1 * gx + m3 * gy + m4 * gz;
            const T tz = m2 * gx + m4 * gy + m5 * gz;
            y[k][i] += GW[(i * NQ + q) * 3 + 0] * tx + GW[(i * NQ + q) * 3 + 1] * ty +
                       GW[(i * NQ + q) * 3 + 2] * tz;
          }
      }
    }
  }

  // 3. colour-ordered accumulation into LDS
  const int cb = icol_ptr[patch], ncol = icol_ptr[patch + 1] - cb - 1;
  for (int col = 0; col < ncol; ++col) {
    const int lo = icol_off[cb + col], hi = icol_off[cb + col + 1];
#pragma unroll
    for (int k = 0; k < CPT; ++k) {
      const int lc = tid + k * kThreads;
      if (lc >= lo && lc < hi) {
#pragma unroll
        for (int i = 0; i < NL; ++i)
          if (li[k][i] != kInvalid) sy[li[k][i]] += y[k][i];
      }
    }
    __syncthreads();
  }

  // 4. owned DoFs: store; ghost DoFs: add to what earlier colours wrote
  for (int j = tid; j < nloc; j += kThreads) {
    if (j < no)
      dst[o0 + j] = sy[j];
    else
      dst[ghost_idx[g0 + j - no]] += sy[j];
  }
}

// Dirichlet rows: dst = src (MODE 0) or 1 (MODE 1) on [n_free, n_dofs)
template <typename T, int MODE>
__global__ void cg_constrained_kernel(int64_t n0, int64_t n1, const T* __restrict__ src,
                                      T* __restrict__ dst) {
  const int64_t i = n0 + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n1) dst[i] = MODE == 0 ? src[i] : T(1);
}

template <int P, typename T, int MODE>
void launch_cg_t(const Operator& op, T* dst, const T* src, hipStream_t s) {
  constexpr int CPT = P == 1 ? 8 : (P == 2 ? 4 : 2);
  const CGPlan& pl = op.cg;
  if (pl.max_patch_cells > CPT * kThreads) throw std::runtime_error("patch larger than kernel capacity");
  const size_t shmem = 2 * (size_t)pl.max_local * sizeof(T);
  auto kern = cg_patch_kernel<P, T, CPT, MODE>;
  if (shmem > 65536) SMF_HIP(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shmem));
  for (int c = 0; c < pl.n_colors; ++c) {
    const int32_t base = pl.color_ptr[c], cnt = pl.color_ptr[c + 1] - base;
    if (cnt == 0) continue;
    hipLaunchKernelGGL(kern, dim3(cnt), dim3(kThreads), shmem, s, op.d_color_patches.as<int32_t>(),
                       base, op.d_cell_start.as<int32_t>(), op.d_own_start.as<int32_t>(),
                       op.d_ghost_start.as<int32_t>(), op.d_ghost_idx.as<int32_t>(),
                       op.d_icol_ptr.as<int32_t>(), op.d_icol_off.as<uint16_t>(),
                       op.d_lidx.as<uint16_t>(), pl.n_cells, op.d_metric.as<T>(),
                       op.d_shape.as<T>(), src, dst);
  }
  const int64_t ncon = pl.n_dofs - pl.n_free;
  if (ncon > 0)
    hipLaunchKernelGGL((cg_constrained_kernel<T, MODE>), dim3((ncon + 255) / 256), dim3(256), 0, s,
                       pl.n_free, pl.n_dofs, src, dst);
  SMF_HIP(hipGetLastError());
}

template <typename T>
void launch_cg_prec(const Operator& op, int mode, T* dst, const T* src, hipStream_t s) {
  switch (op.cg.p * 2 + mode) {
    case 2: launch_cg_t<1, T, 0>(op, dst, src, s); break;
    case 3: launch_cg_t<1, T, 1>(op, dst, src, s); break;
    case 4: launch_cg_t<2, T, 0>(op, dst, src, s); break;
    case 5: launch_cg_t<2, T, 1>(op, dst, src, s); break;
    case 6: launch_cg_t<3, T, 0>(op, dst, src, s); break;
    case 7: launch_cg_t<3, T, 1>(op, dst, src, s); break;
    default: throw std::runtime_error("unsupported CG degree");
  }
}

}  // namespace

void launch_cg(const Operator& op, int mode, void* dst, const void* src, hipStream_t s) {
  if (op.precision == MF_F64)
    launch_cg_prec<double>(op, mode, static_cast<double*>(dst), static_cast<const double*>(src), s);
  else
    launch_cg_prec<float>(op, mode, static_cast<float*>(dst), static_cast<const float*>(src), s);
}

int cg_quadrature_degree(int p) { return p == 1 ? 1 : (p == 2 ? 2 : 4); }

}  // namespace smf
