// Learned per-cell smoother kernels (gfx950): features, BF16 MFMA inference
// with FP32 accumulation, deterministic assembly, FP32 training (forward,
// backward with block-reduced weight gradients) and Adam.
//
// MFMA 16x16x16 bf16 (v_mfma_f32_16x16x16_bf16) operand layout, for lane l:
//   A (16 x 16, M x K): A[l % 16][(l / 16) * 4 + j], j = 0..3
//   B (16 x 16, K x N): B[(l / 16) * 4 + j][l % 16]
//   C (16 x 16, M x N): C[(l / 16) * 4 + j][l % 16]
// (checked on the device against a plain product by tests/test_gpu_mlp.py).
#include <hip/hip_runtime.h>

#include <cmath>

#include "mlp.h"

namespace smf {

namespace {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// float -> bf16 bits, round to nearest even (quiet NaNs kept)
__device__ inline short to_bf16(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (short)((u >> 16) | 0x40);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (short)(u >> 16);
}

typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef unsigned u32x2 __attribute__((ext_vector_type(2)));

// four floats -> four bf16 (round to nearest even, v_cvt_pk_bf16_f32)
__device__ inline bf16x4 pack_bf16(float a0, float a1, float a2, float a3) {
  const bf16x2_t lo = __builtin_convertvector(f32x2{a0, a1}, bf16x2_t);
  const bf16x2_t hi = __builtin_convertvector(f32x2{a2, a3}, bf16x2_t);
  const u32x2 w = {__builtin_bit_cast(unsigned, lo), __builtin_bit_cast(unsigned, hi)};
  return __builtin_bit_cast(bf16x4, w);
}

__device__ inline float gelu(float x) { return 0.5f * x * (1.0f + erff(x * 0.70710678118654752f)); }
__device__ inline float gelu_grad(float x) {
  return 0.5f * (1.0f + erff(x * 0.70710678118654752f)) + x * 0.3989422804014327f * expf(-0.5f * x * x);
}

// x[c][0..9] = r_K / s_K, x[c][10..19] = geometry features; s[c] = |r_K|_2;
// t[c][i] = e_K[i] / (s_K kappa_K) when e is given
__global__ void ls_features_kernel(int64_t ncells, const int32_t* __restrict__ cdofs, const float* __restrict__ geo,
                                   const float* __restrict__ kappa, const double* __restrict__ r,
                                   const double* __restrict__ e, float* __restrict__ x, float* __restrict__ t,
                                   float* __restrict__ scale) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncells) return;
  double rv[kMlpOut], ss = 0.0;
  for (int i = 0; i < kMlpOut; ++i) {
    const int32_t d = cdofs[i * ncells + c];
    rv[i] = d >= 0 ? r[d] : 0.0;
    ss += rv[i] * rv[i];
  }
  const double sn = sqrt(ss) + 1e-300;
  for (int i = 0; i < kMlpOut; ++i) x[c * kMlpIn + i] = (float)(rv[i] / sn);
  for (int i = 0; i < 10; ++i) x[c * kMlpIn + kMlpOut + i] = geo[i * ncells + c];
  if (scale) scale[c] = (float)sn;
  if (t) {
    const double den = sn * (double)kappa[c];
    for (int i = 0; i < kMlpOut; ++i) {
      const int32_t d = cdofs[i * ncells + c];
      t[c * kMlpOut + i] = d >= 0 ? (float)(e[d] / den) : 0.0f;
    }
  }
}

// GELU for inference: x Phi(x) with erf from Abramowitz & Stegun 7.1.26
// (|error| <= 1.5e-7, far below the BF16 rounding of the layer inputs)
__device__ inline float gelu_fast(float x) {
  const float z = fabsf(x) * 0.70710678118654752f;
  const float t = __builtin_amdgcn_rcpf(1.0f + 0.3275911f * z);
  const float poly = t * (0.254829592f + t * (-0.284496736f + t * (1.421413741f + t * (-1.453152027f +
                                                                                         t * 1.061405429f))));
  float e = 1.0f - poly * __builtin_amdgcn_exp2f(-1.4426950408889634f * z * z);
  e = x < 0.0f ? -e : e;
  return 0.5f * x * (1.0f + e);
}

// One wave per group of 64 cells: the group's features go to LDS (lane =
// cell; FUSED: formed here from the residual r - az, az may be null, and
// also written to x_out / scale_out when given; otherwise read from x), then
// four 16-cell tiles run the three MFMA layers.  Outputs are scaled by
// s_K kappa_K (FUSED, or when scale is given) or left raw.
struct InferArgs {
  int64_t ncells;
  const float* x;
  const int32_t* cdofs;
  const float* geo;
  const float* kappa;
  const double* r;
  const double* az;
  float* x_out;
  float* scale_out;
  const float* scale;
  const bf16x4* frag;
  const float* params;
  float* corr;
  double* direct;  // when set (DG: DoF = cell * 10 + node, one cell per DoF): the correction itself, FP64
};

template <bool FUSED>
__global__ __launch_bounds__(64) void ls_infer_kernel(InferArgs g) {
  __shared__ float hb[16][kMlpH + 1];
  __shared__ float xs[64][kMlpIn + 1];
  __shared__ float ss[64];
  const int l = threadIdx.x;
  const int row = l % 16, kq = (l / 16) * 4;
  const int64_t ncells = g.ncells;
  bf16x4 B1[8], B2[16], B3[4];
  for (int f = 0; f < 8; ++f) B1[f] = g.frag[(kFragL1 + f) * 64 + l];
  for (int f = 0; f < 16; ++f) B2[f] = g.frag[(kFragL2 + f) * 64 + l];
  for (int f = 0; f < 4; ++f) B3[f] = g.frag[(kFragL3 + f) * 64 + l];
  float bias1[4], bias2[4];
  for (int nb = 0; nb < 4; ++nb) {
    bias1[nb] = g.params[kOffB1 + nb * 16 + row];
    bias2[nb] = g.params[kOffB2 + nb * 16 + row];
  }
  const float bias3 = row < kMlpOut ? g.params[kOffB3 + row] : 0.0f;
  const int64_t ngroups = (ncells + 63) / 64;
  for (int64_t grp = blockIdx.x; grp < ngroups; grp += gridDim.x) {
    const int64_t cell = grp * 64 + l;
    const bool act = cell < ncells;
    if (FUSED) {
      // branch-free gathers: every load is issued before the first use
      // (clamped cell and DoF indices stay in bounds; masked afterwards)
      const int64_t cl = act ? cell : ncells - 1;
      int32_t dd[kMlpOut];
      double rv[kMlpOut], av[kMlpOut];
      float gv[10];
#pragma unroll
      for (int i = 0; i < kMlpOut; ++i) dd[i] = g.cdofs[i * ncells + cl];
#pragma unroll
      for (int i = 0; i < 10; ++i) gv[i] = g.geo[i * ncells + cl];
#pragma unroll
      for (int i = 0; i < kMlpOut; ++i) rv[i] = g.r[dd[i] >= 0 ? dd[i] : 0];
      if (g.az) {
#pragma unroll
        for (int i = 0; i < kMlpOut; ++i) av[i] = g.az[dd[i] >= 0 ? dd[i] : 0];
      } else {
#pragma unroll
        for (int i = 0; i < kMlpOut; ++i) av[i] = 0.0;
      }
      const float kap = g.kappa[cl];
      double sq = 0.0;
#pragma unroll
      for (int i = 0; i < kMlpOut; ++i) {
        rv[i] = act && dd[i] >= 0 ? rv[i] - av[i] : 0.0;
        sq += rv[i] * rv[i];
      }
      // 1/|r_K| from v_rsq_f64 (features are rounded to BF16); a zero
      // residual gives zero features
      const double inv = __builtin_amdgcn_rsq(sq > 1e-300 ? sq : 1e-300);
      const double sn = sq * inv;
#pragma unroll
      for (int i = 0; i < kMlpOut; ++i) xs[l][i] = (float)(rv[i] * inv);
#pragma unroll
      for (int i = 0; i < 10; ++i) xs[l][kMlpOut + i] = act ? gv[i] : 0.0f;
      ss[l] = act ? (float)sn * kap : 0.0f;
      if (act) {
        if (g.x_out)
          for (int k = 0; k < kMlpIn; ++k) g.x_out[cell * kMlpIn + k] = xs[l][k];
        if (g.scale_out) g.scale_out[cell] = (float)sn;
      }
    } else {
      for (int k = 0; k < kMlpIn; ++k) xs[l][k] = act ? g.x[cell * kMlpIn + k] : 0.0f;
      ss[l] = act && g.scale ? g.scale[cell] * g.kappa[cell] : 1.0f;
    }
    __syncthreads();
    for (int tq = 0; tq < 4; ++tq) {
      const int r0 = tq * 16;
      // layer 1: K = 20 padded to 32
      f32x4 acc[4];
      for (int nb = 0; nb < 4; ++nb) acc[nb] = f32x4{0, 0, 0, 0};
      for (int ks = 0; ks < 2; ++ks) {
        const int k = ks * 16 + kq;  // k..k+3 all below or all above kMlpIn (20 = 5 x 4)
        const float* xr = xs[r0 + row];
        const bf16x4 a = k < kMlpIn ? pack_bf16(xr[k], xr[k + 1], xr[k + 2], xr[k + 3]) : bf16x4{0, 0, 0, 0};
        for (int nb = 0; nb < 4; ++nb)
          acc[nb] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, B1[ks * 4 + nb], acc[nb], 0, 0, 0);
      }
      for (int nb = 0; nb < 4; ++nb)
        for (int j = 0; j < 4; ++j) hb[kq + j][nb * 16 + row] = gelu_fast(acc[nb][j] + bias1[nb]);
      __syncthreads();
      // layer 2
      for (int nb = 0; nb < 4; ++nb) acc[nb] = f32x4{0, 0, 0, 0};
      for (int ks = 0; ks < 4; ++ks) {
        const float* hr = &hb[row][ks * 16 + kq];
        const bf16x4 a = pack_bf16(hr[0], hr[1], hr[2], hr[3]);
        for (int nb = 0; nb < 4; ++nb)
          acc[nb] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, B2[ks * 4 + nb], acc[nb], 0, 0, 0);
      }
      __syncthreads();
      for (int nb = 0; nb < 4; ++nb)
        for (int j = 0; j < 4; ++j) hb[kq + j][nb * 16 + row] = gelu_fast(acc[nb][j] + bias2[nb]);
      __syncthreads();
      // layer 3: N = 10 padded to 16
      f32x4 o = f32x4{0, 0, 0, 0};
      for (int ks = 0; ks < 4; ++ks) {
        const float* hr = &hb[row][ks * 16 + kq];
        const bf16x4 a = pack_bf16(hr[0], hr[1], hr[2], hr[3]);
        o = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, B3[ks], o, 0, 0, 0);
      }
      __syncthreads();
      if (row < kMlpOut)
        for (int j = 0; j < 4; ++j) {
          const int64_t cj = grp * 64 + r0 + kq + j;
          if (cj < ncells) {
            const float v = (o[j] + bias3) * ss[r0 + kq + j];
            if (g.direct) g.direct[cj * kMlpOut + row] = (double)v;
            else g.corr[cj * kMlpOut + row] = v;
          }
        }
    }
    __syncthreads();
  }
}

// out[d] = mean of the cell corrections at DoF d (0 where no cell slot:
// Dirichlet).  Eight slot indices, then eight values, are loaded per step
// with clamped indices (two round trips per eight slots); the sum runs in
// slot order.
__global__ void ls_assemble_kernel(int64_t ndofs, const int64_t* __restrict__ ptr, const int32_t* __restrict__ slot,
                                   const float* __restrict__ corr, double* __restrict__ out) {
  const int64_t d = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (d >= ndofs) return;
  const int32_t k0 = (int32_t)ptr[d], k1 = (int32_t)ptr[d + 1];
  float s = 0.0f;
  for (int32_t k = k0; k < k1; k += 8) {
    int32_t sl[8];
    float v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) sl[u] = slot[min(k + u, k1 - 1)];
#pragma unroll
    for (int u = 0; u < 8; ++u) v[u] = corr[sl[u]];
#pragma unroll
    for (int u = 0; u < 8; ++u)
      if (k + u < k1) s += v[u];
  }
  out[d] = k1 > k0 ? (double)(s * __builtin_amdgcn_rcpf((float)(k1 - k0))) : 0.0;
}

// FP32 training.  The output gradient is 2 (y - t) / n (mean-squared loss
// against targets t) or given per cell (dyin).  A block of 64 threads takes 64-cell chunks; each thread
// runs one cell's forward and backward pass with its activations in padded
// LDS rows (lane stride kLd: conflict-free) and the weights broadcast from
// LDS, then every thread accumulates its own subset of the weight gradient
// over the chunk's 64 cells into the block's slice of gpart (fixed order:
// deterministic).  Loss = sum over cells of |y - t|^2 / n.
constexpr int kTB = 64, kLd = kMlpH + 1;
__global__ __launch_bounds__(kTB) void ls_grad_kernel(int64_t n, const float* __restrict__ x, const float* __restrict__ t,
                                                      const float* __restrict__ dyin,
                                                      const float* __restrict__ params, float* __restrict__ gpart,
                                                      float* __restrict__ lpart) {
  __shared__ float W[kMlpParams];
  __shared__ float sx[kTB][kMlpIn + 1], sdy[kTB][kMlpOut + 1];
  __shared__ float P[kTB][kLd], Q[kTB][kLd], R[kTB][kLd], D2[kTB][kLd];  // a1->d1, h1, a2->h2, d2
  __shared__ float sl[kTB];
  const int tid = threadIdx.x;
  for (int i = tid; i < kMlpParams; i += kTB) W[i] = params[i];
  float* gb = gpart + (int64_t)blockIdx.x * kMlpParams;
  for (int e = tid; e < kMlpParams; e += kTB) gb[e] = 0.0f;
  __syncthreads();
  float loss = 0.0f;
  const float inv_n = 1.0f / (float)n;
  for (int64_t base = (int64_t)blockIdx.x * kTB; base < n; base += (int64_t)gridDim.x * kTB) {
    const int64_t c = base + tid;
    const bool act = c < n;
    for (int k = 0; k < kMlpIn; ++k) sx[tid][k] = act ? x[c * kMlpIn + k] : 0.0f;
    for (int o = 0; o < kMlpH; ++o) {
      float s = W[kOffB1 + o];
#pragma unroll
      for (int k = 0; k < kMlpIn; ++k) s += W[kOffW1 + o * kMlpIn + k] * sx[tid][k];
      P[tid][o] = s;
      Q[tid][o] = gelu(s);
    }
    for (int o = 0; o < kMlpH; ++o) {
      float s = W[kOffB2 + o];
#pragma unroll 16
      for (int k = 0; k < kMlpH; ++k) s += W[kOffW2 + o * kMlpH + k] * Q[tid][k];
      R[tid][o] = s;
    }
    for (int o = 0; o < kMlpOut; ++o) {
      float s = W[kOffB3 + o];
#pragma unroll 16
      for (int k = 0; k < kMlpH; ++k) s += W[kOffW3 + o * kMlpH + k] * gelu(R[tid][k]);
      if (dyin) {
        sdy[tid][o] = act ? dyin[c * kMlpOut + o] : 0.0f;
      } else {
        const float diff = act ? s - t[c * kMlpOut + o] : 0.0f;
        loss += diff * diff;
        sdy[tid][o] = 2.0f * diff * inv_n;
      }
    }
    // d2 = (W3^T dy) .* gelu'(a2); R becomes h2
    for (int k = 0; k < kMlpH; ++k) {
      float s = 0.0f;
#pragma unroll
      for (int o = 0; o < kMlpOut; ++o) s += W[kOffW3 + o * kMlpH + k] * sdy[tid][o];
      const float a2 = R[tid][k];
      D2[tid][k] = s * gelu_grad(a2);
      R[tid][k] = gelu(a2);
    }
    // d1 = (W2^T d2) .* gelu'(a1); P becomes d1
    for (int k = 0; k < kMlpH; ++k) {
      float s = 0.0f;
#pragma unroll 16
      for (int o = 0; o < kMlpH; ++o) s += W[kOffW2 + o * kMlpH + k] * D2[tid][o];
      P[tid][k] = s * gelu_grad(P[tid][k]);
    }
    __syncthreads();
    for (int e = tid; e < kMlpParams; e += kTB) {
      float s = 0.0f;
      if (e < kOffB1) {
        const int o = e / kMlpIn, k = e % kMlpIn;
        for (int cc = 0; cc < kTB; ++cc) s += P[cc][o] * sx[cc][k];
      } else if (e < kOffW2) {
        const int o = e - kOffB1;
        for (int cc = 0; cc < kTB; ++cc) s += P[cc][o];
      } else if (e < kOffB2) {
        const int o = (e - kOffW2) / kMlpH, k = (e - kOffW2) % kMlpH;
        for (int cc = 0; cc < kTB; ++cc) s += D2[cc][o] * Q[cc][k];
      } else if (e < kOffW3) {
        const int o = e - kOffB2;
        for (int cc = 0; cc < kTB; ++cc) s += D2[cc][o];
      } else if (e < kOffB3) {
        const int o = (e - kOffW3) / kMlpH, k = (e - kOffW3) % kMlpH;
        for (int cc = 0; cc < kTB; ++cc) s += sdy[cc][o] * R[cc][k];
      } else {
        const int o = e - kOffB3;
        for (int cc = 0; cc < kTB; ++cc) s += sdy[cc][o];
      }
      gb[e] += s;
    }
    __syncthreads();
  }
  // block loss (fixed-order tree)
  sl[tid] = loss;
  __syncthreads();
  for (int w = kTB / 2; w > 0; w >>= 1) {
    if (tid < w) sl[tid] += sl[tid + w];
    __syncthreads();
  }
  if (tid == 0) lpart[blockIdx.x] = sl[0] * inv_n;
}

__global__ void ls_reduce_kernel(int nblocks, const float* __restrict__ gpart, const float* __restrict__ lpart,
                                 float* __restrict__ grad, float* __restrict__ loss) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e < kMlpParams) {
    float s = 0.0f;
    for (int b = 0; b < nblocks; ++b) s += gpart[(int64_t)b * kMlpParams + e];
    grad[e] = s;
  }
  if (e == 0) {
    float s = 0.0f;
    for (int b = 0; b < nblocks; ++b) s += lpart[b];
    *loss = s;
  }
}

// Adam (Kingma & Ba 2015) with bias correction
__global__ void ls_adam_kernel(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                               float* __restrict__ v, float lr, float b1, float b2, float eps, float c1, float c2) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= kMlpParams) return;
  const float gi = g[e];
  const float mi = b1 * m[e] + (1.0f - b1) * gi;
  const float vi = b2 * v[e] + (1.0f - b2) * gi * gi;
  m[e] = mi;
  v[e] = vi;
  p[e] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// BF16 B-operand fragments: B[k][n] = W[n][k] of each layer
__global__ void ls_pack_kernel(const float* __restrict__ p, bf16x4* __restrict__ frag) {
  const int f = blockIdx.x, l = threadIdx.x;
  int woff, nin, nout, ks, nb;
  if (f < kFragL2) {
    woff = kOffW1; nin = kMlpIn; nout = kMlpH; ks = (f - kFragL1) / 4; nb = (f - kFragL1) % 4;
  } else if (f < kFragL3) {
    woff = kOffW2; nin = kMlpH; nout = kMlpH; ks = (f - kFragL2) / 4; nb = (f - kFragL2) % 4;
  } else {
    woff = kOffW3; nin = kMlpH; nout = kMlpOut; ks = f - kFragL3; nb = 0;
  }
  const int n = nb * 16 + l % 16;
  bf16x4 b;
  for (int j = 0; j < 4; ++j) {
    const int k = ks * 16 + (l / 16) * 4 + j;
    b[j] = to_bf16(n < nout && k < nin ? p[woff + n * nin + k] : 0.0f);
  }
  frag[f * 64 + l] = b;
}

// dL/dy of a loss given as a function of the assembled correction c:
// dy[K][i] = coef * s_K kappa_K * gc[d] / mult(d) for free DoF d = dof(K, i)
__global__ void ls_dy_kernel(int64_t ncells, const int32_t* __restrict__ cdofs, const int64_t* __restrict__ ptr,
                             const float* __restrict__ scale, const float* __restrict__ kappa,
                             const double* __restrict__ gc, double coef, float* __restrict__ dy) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncells) return;
  const double f = coef * (double)scale[c] * (double)kappa[c];
  for (int i = 0; i < kMlpOut; ++i) {
    const int32_t d = cdofs[i * ncells + c];
    dy[c * kMlpOut + i] = d >= 0 ? (float)(f * gc[d] / (double)(ptr[d + 1] - ptr[d])) : 0.0f;
  }
}

// Partial sums of c.(r - az) and c.ac over a fixed block decomposition
// (kRed blocks x 256 threads, tree order): deterministic.
constexpr int kRed = 256;
__global__ __launch_bounds__(256) void ls_dot2_kernel(int64_t n, const double* __restrict__ c,
                                                      const double* __restrict__ r, const double* __restrict__ az,
                                                      const double* __restrict__ ac, double* __restrict__ part) {
  __shared__ double s0[256], s1[256];
  double a = 0.0, q = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)kRed * 256) {
    a += c[i] * (r[i] - az[i]);
    q += c[i] * ac[i];
  }
  s0[threadIdx.x] = a;
  s1[threadIdx.x] = q;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      s0[threadIdx.x] += s0[threadIdx.x + w];
      s1[threadIdx.x] += s1[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = s0[0];
    part[2 * blockIdx.x + 1] = s1[0];
  }
}

// Every block reduces the kRed partial pairs in the same fixed order to
// num = c.r', den = c.A c; omega = num / den is the energy-optimal step along
// c.  z += omega c, az += omega ac.  An unusable step (non-finite, or den <=
// 0) applies omega = 0; block 0 writes sc = {num, den, omega, unusable}.
__global__ __launch_bounds__(256) void ls_omega_kernel(int64_t n, const double* __restrict__ part,
                                                       double* __restrict__ sc, const double* __restrict__ c,
                                                       const double* __restrict__ ac, double* __restrict__ z,
                                                       double* __restrict__ az) {
  __shared__ double s0[kRed], s1[kRed];
  s0[threadIdx.x] = part[2 * threadIdx.x];
  s1[threadIdx.x] = part[2 * threadIdx.x + 1];
  __syncthreads();
  for (int w = kRed / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      s0[threadIdx.x] += s0[threadIdx.x + w];
      s1[threadIdx.x] += s1[threadIdx.x + w];
    }
    __syncthreads();
  }
  const double num = s0[0], den = s1[0];
  double w = num / den;
  const bool ok = den > 0.0 && isfinite(w);
  if (!ok) w = 0.0;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    sc[0] = num;
    sc[1] = den;
    sc[2] = w;
    sc[3] = ok ? 0.0 : 1.0;
  }
  if (!ok) return;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
    z[i] += w * c[i];
    az[i] += w * ac[i];
  }
}

}  // namespace

void ls_launch_features(int64_t ncells, const int32_t* cdofs, const float* geo, const float* kappa, const double* r,
                        const double* e, float* x, float* t, float* scale, hipStream_t s) {
  if (ncells > 0)
    hipLaunchKernelGGL(ls_features_kernel, dim3((unsigned)((ncells + 255) / 256)), dim3(256), 0, s, ncells, cdofs,
                       geo, kappa, r, e, x, t, scale);
  SMF_HIP(hipGetLastError());
}

void ls_launch_infer(int64_t ncells, const float* x, const void* frag, const float* params, const float* scale,
                     const float* kappa, float* corr, hipStream_t s) {
  InferArgs g{};
  g.ncells = ncells;
  g.x = x;
  g.scale = scale;
  g.kappa = kappa;
  g.frag = static_cast<const bf16x4*>(frag);
  g.params = params;
  g.corr = corr;
  const int64_t ngroups = (ncells + 63) / 64;
  if (ncells > 0)
    hipLaunchKernelGGL(ls_infer_kernel<false>, dim3((unsigned)std::min<int64_t>(ngroups, 16384)), dim3(64), 0, s, g);
  SMF_HIP(hipGetLastError());
}

void ls_launch_infer_fused(int64_t ncells, const int32_t* cdofs, const float* geo, const float* kappa,
                           const double* r, const double* az, float* x_out, float* scale_out, const void* frag,
                           const float* params, float* corr, hipStream_t s, double* direct) {
  InferArgs g{};
  g.ncells = ncells;
  g.cdofs = cdofs;
  g.geo = geo;
  g.kappa = kappa;
  g.r = r;
  g.az = az;
  g.x_out = x_out;
  g.scale_out = scale_out;
  g.frag = static_cast<const bf16x4*>(frag);
  g.params = params;
  g.corr = corr;
  g.direct = direct;
  const int64_t ngroups = (ncells + 63) / 64;
  if (ncells > 0)
    hipLaunchKernelGGL(ls_infer_kernel<true>, dim3((unsigned)std::min<int64_t>(ngroups, 16384)), dim3(64), 0, s, g);
  SMF_HIP(hipGetLastError());
}

void ls_launch_assemble(int64_t ndofs, const int64_t* ptr, const int32_t* slot, const float* corr, double* out,
                        hipStream_t s) {
  if (ndofs > 0)
    hipLaunchKernelGGL(ls_assemble_kernel, dim3((unsigned)((ndofs + 255) / 256)), dim3(256), 0, s, ndofs, ptr, slot,
                       corr, out);
  SMF_HIP(hipGetLastError());
}

int ls_grad_blocks() { return 256; }

void ls_launch_grad(int64_t n, const float* x, const float* t, const float* dy, const float* params, float* gpart,
                    float* lpart, float* grad, float* loss, hipStream_t s) {
  const int nb = ls_grad_blocks();
  hipLaunchKernelGGL(ls_grad_kernel, dim3(nb), dim3(kTB), 0, s, n, x, t, dy, params, gpart, lpart);
  hipLaunchKernelGGL(ls_reduce_kernel, dim3((kMlpParams + 255) / 256), dim3(256), 0, s, nb, gpart, lpart, grad, loss);
  SMF_HIP(hipGetLastError());
}

void ls_launch_adam(float* p, const float* g, float* m, float* v, double lr, int64_t step, hipStream_t s) {
  const double b1 = 0.9, b2 = 0.999;
  const float c1 = (float)(1.0 - std::pow(b1, (double)step)), c2 = (float)(1.0 - std::pow(b2, (double)step));
  hipLaunchKernelGGL(ls_adam_kernel, dim3((kMlpParams + 255) / 256), dim3(256), 0, s, p, g, m, v, (float)lr,
                     (float)b1, (float)b2, 1e-8f, c1, c2);
  SMF_HIP(hipGetLastError());
}

void ls_launch_pack(const float* p, void* frag, hipStream_t s) {
  hipLaunchKernelGGL(ls_pack_kernel, dim3(kFrags), dim3(64), 0, s, p, static_cast<bf16x4*>(frag));
  SMF_HIP(hipGetLastError());
}

}  // namespace smf

namespace smf {

void ls_launch_dy(int64_t ncells, const int32_t* cdofs, const int64_t* ptr, const float* scale, const float* kappa,
                  const double* gc, double coef, float* dy, hipStream_t s) {
  if (ncells > 0)
    hipLaunchKernelGGL(ls_dy_kernel, dim3((unsigned)((ncells + 255) / 256)), dim3(256), 0, s, ncells, cdofs, ptr,
                       scale, kappa, gc, coef, dy);
  SMF_HIP(hipGetLastError());
}

int ls_dot2_parts() { return kRed; }

void ls_launch_omega(int64_t n, double* part, double* sc, const double* c, const double* r, const double* az,
                     const double* ac, double* z, double* azout, hipStream_t s) {
  hipLaunchKernelGGL(ls_dot2_kernel, dim3(kRed), dim3(256), 0, s, n, c, r, az, ac, part);
  const unsigned grid = (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 2048));
  hipLaunchKernelGGL(ls_omega_kernel, dim3(grid), dim3(256), 0, s, n, part, sc, c, ac, z, azout);
  SMF_HIP(hipGetLastError());
}

}  // namespace smf
