// Discontinuous-Galerkin SIPG kernels for gfx950 (scalar and
// components-batched vector / Helmholtz-like operators).
//
// One lane per cell; a workgroup (4 wave64) shares the reference tables in
// LDS.  A lane evaluates the cell integral and all four face integrals of its
// cell and writes only its own DoFs, so every DoF is written once, with no
// atomics and no colouring: the result is bit-reproducible.  Interior faces
// are evaluated by both adjacent cells (each computing its own test side);
// the neighbour's trace uses the same canonical face parametrisation
// (ascending global vertex order), so its table is indexed by its local face
// number only.
//
// Reference: vmult_manual, DG/cells_batched/application.cc:1318-1332 with the
// cell integral (1493-1735), the SIPG flux of interior faces (jump, average
// normal derivative, sigma = (p+1)(p+3)/3 max(tau-,tau+), 2104-2187) and the
// mirror-principle boundary flux (2688-2701); components batched as in
// DG/components_batched/application.cc:715-1535; Helmholtz mass term as in
// DG/components_batched/application_helmholtz.cc:1025-1111.
#pragma once
#include <hip/hip_runtime.h>

#include <stdexcept>

#include "dg_cell.h"
#include "epilogue.h"
#include "op.h"

namespace smf {

namespace {

constexpr int kThreads = 256;

template <int P, bool MASS>
struct DGShape {
  static constexpr int NL = (P + 1) * (P + 2) * (P + 3) / 6;
  // cell rule: stiffness exact (2p-2) or, with a mass term, 2p
  static constexpr int NQ = MASS ? (P == 1 ? 4 : (P == 2 ? 11 : 24)) : (P == 1 ? 1 : (P == 2 ? 4 : 11));
  static constexpr int NQF = P == 1 ? 3 : (P == 2 ? 6 : 12);
};

__device__ __constant__ double kNhatD[4][3] = {{1, 1, 1}, {-1, 0, 0}, {0, -1, 0}, {0, 0, -1}};

// MODE 0: dst = A src; MODE 1: dst = diag(A)
template <int P, typename T, int NC, bool MASS, int MODE>
__global__ __launch_bounds__(kThreads) void dg_cell_kernel(
    int64_t ncells, const T* __restrict__ geo, const int32_t* __restrict__ nbr,
    const uint32_t* __restrict__ nface, const T* __restrict__ cell_tab,
    const T* __restrict__ face_tab, T factor, T mass, T lap, const T* __restrict__ src,
    T* __restrict__ dst) {
  using S = DGShape<P, MASS>;
  constexpr int NL = S::NL, NQ = S::NQ, NQF = S::NQF;
  constexpr int CT = NQ * NL * 4 + NQ;            // cell table size
  constexpr int FS = NQF * NL * 4;                // per-face stride
  constexpr int FT = 4 * FS + NQF;                // face table size
  extern __shared__ unsigned char smem_raw[];
  T* sct = reinterpret_cast<T*>(smem_raw);
  T* sft = sct + CT;
  for (int j = threadIdx.x; j < CT; j += kThreads) sct[j] = cell_tab[j];
  for (int j = threadIdx.x; j < FT; j += kThreads) sft[j] = face_tab[j];
  __syncthreads();

  const int64_t k = xcd_block(blockIdx.x, gridDim.x) * kThreads + threadIdx.x;
  if (k >= ncells) return;
  const T* CG = sct;
  const T* CW = sct + NQ * NL * 3;
  const T* CV = CW + NQ;
  const T* FW = sft + 4 * FS;

  T jit[9];
#pragma unroll
  for (int t = 0; t < 9; ++t) jit[t] = geo[t * ncells + k];
  const T adet = geo[9 * ncells + k], tau = geo[10 * ncells + k];
  const T M0 = adet * (jit[0] * jit[0] + jit[3] * jit[3] + jit[6] * jit[6]);
  const T M1 = adet * (jit[0] * jit[1] + jit[3] * jit[4] + jit[6] * jit[7]);
  const T M2 = adet * (jit[0] * jit[2] + jit[3] * jit[5] + jit[6] * jit[8]);
  const T M3 = adet * (jit[1] * jit[1] + jit[4] * jit[4] + jit[7] * jit[7]);
  const T M4 = adet * (jit[1] * jit[2] + jit[4] * jit[5] + jit[7] * jit[8]);
  const T M5 = adet * (jit[2] * jit[2] + jit[5] * jit[5] + jit[8] * jit[8]);
  const uint32_t nf = nface[k];

#pragma unroll 1
  for (int c = 0; c < NC; ++c) {
    const int64_t base = (k * NC + c) * NL;
    T u[NL], y[NL];
#pragma unroll
    for (int i = 0; i < NL; ++i) {
      u[i] = MODE == 0 ? src[base + i] : T(0);
      y[i] = T(0);
    }
    // ---- cell integral ----
#pragma unroll 1
    for (int q = 0; q < NQ; ++q) {
      const T* G = CG + q * NL * 3;
      const T* V = CV + q * NL;
      const T w = CW[q];
      if (MODE == 0) {
        T gx = T(0), gy = T(0), gz = T(0), uv = T(0);
#pragma unroll
        for (int i = 0; i < NL; ++i) {
          gx += G[3 * i] * u[i];
          gy += G[3 * i + 1] * u[i];
          gz += G[3 * i + 2] * u[i];
          if (MASS) uv += V[i] * u[i];
        }
        const T wl = w * lap;
        const T tx = wl * (M0 * gx + M1 * gy + M2 * gz);
        const T ty = wl * (M1 * gx + M3 * gy + M4 * gz);
        const T tz = wl * (M2 * gx + M4 * gy + M5 * gz);
        const T tv = MASS ? w * mass * adet * uv : T(0);
#pragma unroll
        for (int i = 0; i < NL; ++i) {
          y[i] += G[3 * i] * tx + G[3 * i + 1] * ty + G[3 * i + 2] * tz;
          if (MASS) y[i] += V[i] * tv;
        }
      } else {
#pragma unroll
        for (int i = 0; i < NL; ++i) {
          const T gx = G[3 * i], gy = G[3 * i + 1], gz = G[3 * i + 2];
          y[i] += w * lap * (gx * (M0 * gx + M1 * gy + M2 * gz) + gy * (M1 * gx + M3 * gy + M4 * gz) +
                             gz * (M2 * gx + M4 * gy + M5 * gz));
          if (MASS) y[i] += w * mass * adet * V[i] * V[i];
        }
      }
    }
    // ---- face integrals ----
#pragma unroll 1
    for (int f = 0; f < 4; ++f) {
      const T* FV = sft + f * FS;
      const T* FG = FV + NQF * NL;
      T nj0 = jit[0] * (T)kNhatD[f][0] + jit[1] * (T)kNhatD[f][1] + jit[2] * (T)kNhatD[f][2];
      T nj1 = jit[3] * (T)kNhatD[f][0] + jit[4] * (T)kNhatD[f][1] + jit[5] * (T)kNhatD[f][2];
      T nj2 = jit[6] * (T)kNhatD[f][0] + jit[7] * (T)kNhatD[f][1] + jit[8] * (T)kNhatD[f][2];
      const T len = sqrt(nj0 * nj0 + nj1 * nj1 + nj2 * nj2);
      const T il = T(1) / len;
      const T n0 = nj0 * il, n1 = nj1 * il, n2 = nj2 * il;
      const T area2 = adet * len * lap;
      const T nJ0 = jit[0] * n0 + jit[3] * n1 + jit[6] * n2;
      const T nJ1 = jit[1] * n0 + jit[4] * n1 + jit[7] * n2;
      const T nJ2 = jit[2] * n0 + jit[5] * n1 + jit[8] * n2;
      const int32_t kb = nbr[f * ncells + k];
      const bool interior = kb >= 0;
      T sigma = factor * tau;
      T nb0 = T(0), nb1 = T(0), nb2 = T(0);
      T ub[NL];
      const T* FVb = sft;
      if (interior) {
        const int fb = (nf >> (8 * f)) & 3;
        FVb = sft + fb * FS;
        T jb[9];
#pragma unroll
        for (int t = 0; t < 9; ++t) jb[t] = geo[t * ncells + kb];
        nb0 = jb[0] * n0 + jb[3] * n1 + jb[6] * n2;
        nb1 = jb[1] * n0 + jb[4] * n1 + jb[7] * n2;
        nb2 = jb[2] * n0 + jb[5] * n1 + jb[8] * n2;
        const T taub = geo[10 * ncells + kb];
        sigma = factor * (tau > taub ? tau : taub);
        if (MODE == 0) {
#pragma unroll
          for (int i = 0; i < NL; ++i) ub[i] = src[(kb * NC + c) * NL + i];
        }
      }
      const T* FGb = FVb + NQF * NL;
#pragma unroll 1
      for (int q = 0; q < NQF; ++q) {
        const T* V = FV + q * NL;
        const T* G = FG + q * NL * 3;
        const T jxw = FW[q] * area2;
        if (MODE == 0) {
          T uv = T(0), dn = T(0);
#pragma unroll
          for (int i = 0; i < NL; ++i) {
            uv += V[i] * u[i];
            dn += (G[3 * i] * nJ0 + G[3 * i + 1] * nJ1 + G[3 * i + 2] * nJ2) * u[i];
          }
          T up = -uv, dnp = dn;
          if (interior) {
            const T* Vb = FVb + q * NL;
            const T* Gb = FGb + q * NL * 3;
            up = T(0);
            dnp = T(0);
#pragma unroll
            for (int i = 0; i < NL; ++i) {
              up += Vb[i] * ub[i];
              dnp += (Gb[3 * i] * nb0 + Gb[3 * i + 1] * nb1 + Gb[3 * i + 2] * nb2) * ub[i];
            }
          }
          const T jump = uv - up, avg = T(0.5) * (dn + dnp);
          const T tv = (sigma * jump - avg) * jxw, tg = T(-0.5) * jump * jxw;
#pragma unroll
          for (int i = 0; i < NL; ++i)
            y[i] += V[i] * tv + (G[3 * i] * nJ0 + G[3 * i + 1] * nJ1 + G[3 * i + 2] * nJ2) * tg;
        } else {
          const T s = interior ? T(1) : T(2);
#pragma unroll
          for (int i = 0; i < NL; ++i) {
            const T vi = V[i];
            const T di = G[3 * i] * nJ0 + G[3 * i + 1] * nJ1 + G[3 * i + 2] * nJ2;
            const T jump = s * vi, avg = T(0.5) * s * di;
            y[i] += jxw * ((sigma * jump - avg) * vi - T(0.5) * jump * di);
          }
        }
      }
    }
#pragma unroll
    for (int i = 0; i < NL; ++i) dst[base + i] = y[i];
  }
}

// Curved cells: one lane per cell, the reference tables (with the rule
// points) in LDS.  The P2 geometry map's Jacobian is affine in the reference
// coordinates, J(xi) = C + sum_f xi_f D_f (27 coefficients per cell), so each
// quadrature point forms J, its cofactor matrix and det J in registers:
//   cell: g <- w_q |det J| J^-1 J^-T g = (w_q / |det J|) adj (adj^T g), the
//         general D_e branch (DG/cells_batched/application.cc:1650-1664);
//   face: area element |cof J N_f|, outward normal n = sgn(det J) cof J N_f /
//         |cof J N_f|, and per side J^-1 n at that side's reference point (the
//         same physical point: canonical face parametrisation), the curvilinear
//         flux of :2191-2238 with sigma = factor max(tau-, tau+) per face.
constexpr int kCurvedThreads = 64;
template <int P>
struct DGCurvedShape {
  static constexpr int NL = (P + 1) * (P + 2) * (P + 3) / 6;
  static constexpr int NQ = P == 1 ? 4 : (P == 2 ? 11 : 24);   // cell rule of degree 2p
  static constexpr int NQF = P == 1 ? 3 : (P == 2 ? 6 : 12);   // face rule of degree 2p
  static constexpr int CT = NQ * NL * 4 + NQ + 3 * NQ;          // cell table + points
  static constexpr int FS = NQF * NL * 4;                       // per-face stride
  static constexpr int FT = 4 * FS + NQF + 12 * NQF;            // face tables + weights + points
};

// cofactor matrix C of J(xi) (C[d][e] = cofactor of J[d][e]; adj J = C^T) and det J
template <typename T>
__device__ __forceinline__ void curved_cof(const T* jc, T x0, T x1, T x2, T C[3][3], T& det) {
  T J[3][3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    J[d][0] = jc[d] + x0 * jc[9 + d] + x1 * jc[18 + d] + x2 * jc[21 + d];
    J[d][1] = jc[3 + d] + x1 * jc[12 + d] + x0 * jc[18 + d] + x2 * jc[24 + d];
    J[d][2] = jc[6 + d] + x2 * jc[15 + d] + x0 * jc[21 + d] + x1 * jc[24 + d];
  }
  C[0][0] = J[1][1] * J[2][2] - J[1][2] * J[2][1];
  C[0][1] = J[1][2] * J[2][0] - J[1][0] * J[2][2];
  C[0][2] = J[1][0] * J[2][1] - J[1][1] * J[2][0];
  C[1][0] = J[0][2] * J[2][1] - J[0][1] * J[2][2];
  C[1][1] = J[0][0] * J[2][2] - J[0][2] * J[2][0];
  C[1][2] = J[0][1] * J[2][0] - J[0][0] * J[2][1];
  C[2][0] = J[0][1] * J[1][2] - J[0][2] * J[1][1];
  C[2][1] = J[0][2] * J[1][0] - J[0][0] * J[1][2];
  C[2][2] = J[0][0] * J[1][1] - J[0][1] * J[1][0];
  det = J[0][0] * C[0][0] + J[0][1] * C[0][1] + J[0][2] * C[0][2];
}

template <int P, typename T, int NC, bool MASS, int MODE>
__global__ __launch_bounds__(kCurvedThreads) void dg_curved_kernel(
    int64_t ncells, const T* __restrict__ geo, const T* __restrict__ jgeo, const int32_t* __restrict__ nbr,
    const uint32_t* __restrict__ nface, const T* __restrict__ cell_tab, const T* __restrict__ face_tab, T factor,
    T mass, T lap, const T* __restrict__ src, T* __restrict__ dst) {
  using S = DGCurvedShape<P>;
  constexpr int NL = S::NL, NQ = S::NQ, NQF = S::NQF, FS = S::FS;
  extern __shared__ unsigned char smem_raw[];
  T* sct = reinterpret_cast<T*>(smem_raw);
  T* sft = sct + S::CT;
  for (int j = threadIdx.x; j < S::CT; j += kCurvedThreads) sct[j] = cell_tab[j];
  for (int j = threadIdx.x; j < S::FT; j += kCurvedThreads) sft[j] = face_tab[j];
  __syncthreads();
  const int64_t k = xcd_block(blockIdx.x, gridDim.x) * kCurvedThreads + threadIdx.x;
  if (k >= ncells) return;
  const T* CG = sct;
  const T* CW = sct + NQ * NL * 3;
  const T* CV = CW + NQ;
  const T* CP = CV + NQ * NL;
  const T* FW = sft + 4 * FS;
  const T* FP = FW + NQF;
  T jc[27];
#pragma unroll
  for (int t = 0; t < 27; ++t) jc[t] = jgeo[(int64_t)t * ncells + k];
  const T tau = geo[10 * ncells + k];
  const uint32_t nf = nface[k];

#pragma unroll 1
  for (int c = 0; c < NC; ++c) {
    const int64_t base = (k * NC + c) * NL;
    T u[NL], y[NL];
#pragma unroll
    for (int i = 0; i < NL; ++i) {
      u[i] = MODE == 0 ? src[base + i] : T(0);
      y[i] = T(0);
    }
    // ---- cell integral ----
#pragma unroll 1
    for (int q = 0; q < NQ; ++q) {
      T C[3][3], det;
      curved_cof(jc, CP[3 * q], CP[3 * q + 1], CP[3 * q + 2], C, det);
      const T adet = det < T(0) ? -det : det;
      const T* G = CG + q * NL * 3;
      const T* V = CV + q * NL;
      const T sc = CW[q] * lap / adet;
      if (MODE == 0) {
        T gx = T(0), gy = T(0), gz = T(0), uv = T(0);
#pragma unroll
        for (int i = 0; i < NL; ++i) {
          gx += G[3 * i] * u[i];
          gy += G[3 * i + 1] * u[i];
          gz += G[3 * i + 2] * u[i];
          if (MASS) uv += V[i] * u[i];
        }
        // t = (w / |det|) adj (adj^T g), adj = C^T
        const T h0 = C[0][0] * gx + C[0][1] * gy + C[0][2] * gz;
        const T h1 = C[1][0] * gx + C[1][1] * gy + C[1][2] * gz;
        const T h2 = C[2][0] * gx + C[2][1] * gy + C[2][2] * gz;
        const T tx = sc * (C[0][0] * h0 + C[1][0] * h1 + C[2][0] * h2);
        const T ty = sc * (C[0][1] * h0 + C[1][1] * h1 + C[2][1] * h2);
        const T tz = sc * (C[0][2] * h0 + C[1][2] * h1 + C[2][2] * h2);
        const T tv = MASS ? CW[q] * mass * adet * uv : T(0);
#pragma unroll
        for (int i = 0; i < NL; ++i) {
          y[i] += G[3 * i] * tx + G[3 * i + 1] * ty + G[3 * i + 2] * tz;
          if (MASS) y[i] += V[i] * tv;
        }
      } else {
#pragma unroll
        for (int i = 0; i < NL; ++i) {
          const T gx = G[3 * i], gy = G[3 * i + 1], gz = G[3 * i + 2];
          const T h0 = C[0][0] * gx + C[0][1] * gy + C[0][2] * gz;
          const T h1 = C[1][0] * gx + C[1][1] * gy + C[1][2] * gz;
          const T h2 = C[2][0] * gx + C[2][1] * gy + C[2][2] * gz;
          y[i] += sc * (h0 * h0 + h1 * h1 + h2 * h2);
          if (MASS) y[i] += CW[q] * mass * adet * V[i] * V[i];
        }
      }
    }
    // ---- face integrals ----
#pragma unroll 1
    for (int f = 0; f < 4; ++f) {
      const T N0 = T(kNhatD[f][0]), N1 = T(kNhatD[f][1]), N2 = T(kNhatD[f][2]);
      const int32_t kb = nbr[f * ncells + k];
      const bool interior = kb >= 0;
      const int fb = interior ? (int)((nf >> (8 * f)) & 3) : 0;
      const int64_t kn = interior ? kb : k;
      const T taub = geo[10 * ncells + kn];
      const T sigma = factor * (interior ? (tau > taub ? tau : taub) : tau);
      T jcb[27], ub[NL];
      if (MODE == 0 && interior) {
#pragma unroll
        for (int t = 0; t < 27; ++t) jcb[t] = jgeo[(int64_t)t * ncells + kn];
#pragma unroll
        for (int i = 0; i < NL; ++i) ub[i] = src[(kn * NC + c) * NL + i];
      }
      const T* FV = sft + f * FS;
      const T* FG = FV + NQF * NL;
      const T* FVb = sft + fb * FS;
      const T* FGb = FVb + NQF * NL;
#pragma unroll 1
      for (int q = 0; q < NQF; ++q) {
        const T* xi = FP + (f * NQF + q) * 3;
        T C[3][3], det;
        curved_cof(jc, xi[0], xi[1], xi[2], C, det);
        // area vector cof J N_f; outward unit normal; J^-1 n = C^T n / det
        const T a0 = C[0][0] * N0 + C[0][1] * N1 + C[0][2] * N2;
        const T a1 = C[1][0] * N0 + C[1][1] * N1 + C[1][2] * N2;
        const T a2 = C[2][0] * N0 + C[2][1] * N1 + C[2][2] * N2;
        const T la = sqrt(a0 * a0 + a1 * a1 + a2 * a2);
        const T jxw = FW[q] * la * lap;
        const T sn = (det < T(0) ? T(-1) : T(1)) / la;
        const T n0 = sn * a0, n1 = sn * a1, n2 = sn * a2;
        const T idet = T(1) / det;
        const T nJ0 = (C[0][0] * n0 + C[1][0] * n1 + C[2][0] * n2) * idet;
        const T nJ1 = (C[0][1] * n0 + C[1][1] * n1 + C[2][1] * n2) * idet;
        const T nJ2 = (C[0][2] * n0 + C[1][2] * n1 + C[2][2] * n2) * idet;
        const T* V = FV + q * NL;
        const T* G = FG + q * NL * 3;
        if (MODE == 0) {
          T uv = T(0), dn = T(0);
#pragma unroll
          for (int i = 0; i < NL; ++i) {
            uv += V[i] * u[i];
            dn += (G[3 * i] * nJ0 + G[3 * i + 1] * nJ1 + G[3 * i + 2] * nJ2) * u[i];
          }
          T up = -uv, dnp = dn;  // mirror principle on the boundary
          if (interior) {
            const T* xb = FP + (fb * NQF + q) * 3;
            T Cb[3][3], detb;
            curved_cof(jcb, xb[0], xb[1], xb[2], Cb, detb);
            const T idb = T(1) / detb;
            const T b0 = (Cb[0][0] * n0 + Cb[1][0] * n1 + Cb[2][0] * n2) * idb;
            const T b1 = (Cb[0][1] * n0 + Cb[1][1] * n1 + Cb[2][1] * n2) * idb;
            const T b2 = (Cb[0][2] * n0 + Cb[1][2] * n1 + Cb[2][2] * n2) * idb;
            const T* Vb = FVb + q * NL;
            const T* Gb = FGb + q * NL * 3;
            up = T(0);
            dnp = T(0);
#pragma unroll
            for (int i = 0; i < NL; ++i) {
              up += Vb[i] * ub[i];
              dnp += (Gb[3 * i] * b0 + Gb[3 * i + 1] * b1 + Gb[3 * i + 2] * b2) * ub[i];
            }
          }
          const T jump = uv - up, avg = T(0.5) * (dn + dnp);
          const T tv = (sigma * jump - avg) * jxw, tg = T(-0.5) * jump * jxw;
#pragma unroll
          for (int i = 0; i < NL; ++i)
            y[i] += V[i] * tv + (G[3 * i] * nJ0 + G[3 * i + 1] * nJ1 + G[3 * i + 2] * nJ2) * tg;
        } else {
          const T s = interior ? T(1) : T(2);
#pragma unroll
          for (int i = 0; i < NL; ++i) {
            const T vi = V[i];
            const T di = G[3 * i] * nJ0 + G[3 * i + 1] * nJ1 + G[3 * i + 2] * nJ2;
            const T jump = s * vi, avg = T(0.5) * s * di;
            y[i] += jxw * ((sigma * jump - avg) * vi - T(0.5) * jump * di);
          }
        }
      }
    }
#pragma unroll
    for (int i = 0; i < NL; ++i) dst[base + i] = y[i];
  }
}

template <int P, typename T, int NC, bool MASS, int MODE>
void launch_dg_curved(const Operator& op, T* dst, const T* src, hipStream_t s) {
  using S = DGCurvedShape<P>;
  const DGPlan& pl = op.dg;
  if (pl.n_q != S::NQ || pl.n_qf != S::NQF) throw std::runtime_error("curved DG: rule sizes do not match the kernel");
  if ((int64_t)pl.cell_tab.size() != S::CT || (int64_t)pl.face_tab.size() != S::FT)
    throw std::runtime_error("curved DG: table sizes do not match the kernel");
  const size_t shmem = (size_t)(S::CT + S::FT) * sizeof(T);
  auto kern = dg_curved_kernel<P, T, NC, MASS, MODE>;
  if (shmem > 65536)
    SMF_HIP(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shmem));
  if (pl.n_cells > 0)
    hipLaunchKernelGGL(kern, dim3((unsigned)((pl.n_cells + kCurvedThreads - 1) / kCurvedThreads)),
                       dim3(kCurvedThreads), shmem, s, pl.n_cells, op.d_dg_geo.as<T>(), op.d_dg_jgeo.as<T>(),
                       op.d_dg_nbr.as<int32_t>(), op.d_dg_nface.as<uint32_t>(), op.d_dg_cell_tab.as<T>(),
                       op.d_dg_face_tab.as<T>(), (T)pl.penalty_factor, (T)pl.mass, (T)pl.laplace, src, dst);
  SMF_HIP(hipGetLastError());
}

// One lane per cell, compile-time tables (see dg_cell.h).
template <int P, typename T, int NC, bool MASS, int MODE>
__global__ __launch_bounds__(kThreads) void dg_cell_kernel_v2(int64_t ncells, const T* __restrict__ geo,
                                                               const int32_t* __restrict__ nbr,
                                                               const uint32_t* __restrict__ nface, T factor, T mass,
                                                               T lap, const T* __restrict__ src,
                                                               T* __restrict__ dst, int pkind) {
  const int64_t k = xcd_block(blockIdx.x, gridDim.x) * kThreads + threadIdx.x;
  if (k >= ncells) return;
  dgt::apply_cell<P, T, NC, MASS, MODE>(k, ncells, geo, nbr, nface, factor, mass, lap, src, dst, pkind);
}

// Four lanes per cell, lane f = face f (dgt::apply_cell_lane): a quarter of
// the registers of the one-lane form, so four times the waves in flight to
// cover the neighbour gathers.  The lanes' rows meet in LDS and are summed
// in lane order (deterministic); each lane then stores every fourth row.
constexpr int kQuadCells = 64;  // cells per 256-thread block
// Stage the block's own values in LDS (coalesced loads; the neighbour's
// values too when it is one of the block's cells: 69% of the interior faces
// in the hierarchical order).  Measured at 64^3 (same run, staged against
// not): P2 FP64 0.177 / 0.190 ms, P2 FP32 0.099 / 0.108, Helmholtz P2 32^3
// 0.066 / 0.068, P1 FP64 0.098 / 0.100, P3 FP32 0.227 / 0.247; slower for
// P3 FP64 (0.395 / 0.380, the extra 10 KB of LDS per block) and P1 FP32
// (0.046 / 0.044).
// The four lanes' rows of each cell (ybuf[cell][lane][slot], written by
// lane_component) summed in lane order, then written back into the same
// buffer as the block's result rows, cell after cell ([cell][node], the
// order of dst), so that the caller's stores are contiguous runs instead of
// 32 B pieces 8 NL B apart.  Every thread of the block calls it (barriers).
// Same runs at 64^3 against the direct stores: DG SIPG P2 FP64 0.1748 /
// 0.1771 ms; DG P2 MG-CG at 48^3 (the FP32 epilogue kernel) 15.4 / 15.7 ms.
template <int P, typename T, int W>
__device__ inline T* quad_rows_to_block(T (&ybuf)[kQuadCells][4][W], int cl, int f) {
  constexpr int NL = dgt::nloc(P);
  static_assert(W >= NL, "row buffer too short");
  __syncthreads();
  T sum[(NL + 3) / 4];
#pragma unroll
  for (int t = 0; t < (NL + 3) / 4; ++t)
    sum[t] = f + 4 * t < NL ? ((ybuf[cl][0][dgt::lane_slot4<P>(0, t, f)] + ybuf[cl][1][dgt::lane_slot4<P>(1, t, f)]) +
                               ybuf[cl][2][dgt::lane_slot4<P>(2, t, f)]) +
                                  ybuf[cl][3][dgt::lane_slot4<P>(3, t, f)]
                            : T(0);
  __syncthreads();
  T* const yblk = &ybuf[0][0][0];
#pragma unroll
  for (int t = 0; t < (NL + 3) / 4; ++t)
    if (f + 4 * t < NL) yblk[cl * NL + f + 4 * t] = sum[t];
  __syncthreads();
  return yblk;
}
template <int P, typename T>
constexpr bool kQuadStage = P == 2 || (P == 1 && sizeof(T) == 8) || (P == 3 && sizeof(T) == 4);
template <int P, typename T, int NC, bool MASS, int MODE, bool ST = false, bool NT = false>
__global__ __launch_bounds__(4 * kQuadCells) void dg_quad_kernel(int64_t ncells, const T* __restrict__ bgeo,
                                                                  const T* __restrict__ fgeo,
                                                                  const int32_t* __restrict__ nbrf, T mass, T lap,
                                                                  const T* __restrict__ src, T* __restrict__ dst) {
  constexpr int NL = dgt::nloc(P);
  __shared__ T ybuf[kQuadCells][4][NL + 1];
  const int cl = threadIdx.x >> 2, f = threadIdx.x & 3;
  // one component per block, the components of a cell group in consecutive
  // logical blocks (same XCD, so the second and third read the cells'
  // geometry and neighbours from L2); no loop over components, whose
  // invariant table constants the compiler would keep in registers
  const int64_t blk = xcd_block(blockIdx.x, gridDim.x);
  const int c = NC == 1 ? 0 : (int)(blk % NC);
  const int64_t k0 = (NC == 1 ? blk : blk / NC) * kQuadCells;
  const int64_t k = k0 + cl;
  const bool active = k < ncells;
  const int64_t kc = active ? k : ncells - 1;  // every lane reaches the barrier
  // ST: the block's own values (component c) staged in LDS with coalesced loads
  __shared__ T sblk[ST ? kQuadCells * NL : 1];
  const int ncb = (int)(ncells - k0 < kQuadCells ? ncells - k0 : kQuadCells);
  if (ST) {
    for (int j = threadIdx.x; j < ncb * NL; j += 4 * kQuadCells)
      sblk[j] = src[((k0 + j / NL) * NC + c) * NL + j % NL];
    __syncthreads();
  }
  // NT (scalar FP64 products whose data exceeds the Infinity Cache, chosen
  // at launch): the geometry, face records and neighbour words are loaded and
  // dst stored non-temporally, as they are touched once per product, and the
  // caches are left to the gathered neighbour values
  static_assert(!NT || (NC == 1 && sizeof(T) == 8), "non-temporal variant: scalar FP64 products");
  const dgt::LaneGeo<T> g = dgt::lane_geometry<P, T, NT>(kc, f, ncells, bgeo, fgeo, nbrf, lap);
  dgt::lane_component<P, T, NC, MASS, MODE, ST>(g, kc, f, c, mass, lap, src, ybuf[cl][f], sblk, k0, ncb);
  if (sizeof(T) == 8) {
    T* const yblk = quad_rows_to_block<P>(ybuf, cl, f);
    for (int j = threadIdx.x; j < ncb * NL; j += 4 * kQuadCells) {
      if (NT) __builtin_nontemporal_store(yblk[j], dst + k0 * NL + j);
      else dst[NC == 1 ? k0 * NL + j : ((k0 + j / NL) * NC + c) * NL + j % NL] = yblk[j];
    }
    return;
  }
  // FP32: the rows stored straight from the lane sums (the block-ordered
  // store measured slower in FP32: Helmholtz P2 32^3 0.0383 against 0.0360 ms)
  __syncthreads();
  if (active)
#pragma unroll
    for (int t = 0; t < (NL + 3) / 4; ++t)
      if (f + 4 * t < NL)
        dst[(k * NC + c) * NL + f + 4 * t] =
            ((ybuf[cl][0][dgt::lane_slot4<P>(0, t, f)] + ybuf[cl][1][dgt::lane_slot4<P>(1, t, f)]) +
             ybuf[cl][2][dgt::lane_slot4<P>(2, t, f)]) +
            ybuf[cl][3][dgt::lane_slot4<P>(3, t, f)];
}

// The four-lane kernel for scalar SIPG with a RowEpilogue on every row
// instead of the store (the Chebyshev step or the residual of a multigrid
// level).
template <int P, typename T>
__global__ __launch_bounds__(4 * kQuadCells) void dg_quad_ep_kernel(int64_t ncells, const T* __restrict__ bgeo,
                                                                     const T* __restrict__ fgeo,
                                                                     const int32_t* __restrict__ nbrf,
                                                                     const T* __restrict__ src, RowEpilogue<T> ep) {
  constexpr int NL = dgt::nloc(P);
  __shared__ T ybuf[kQuadCells][4][NL + 1];
  const int cl = threadIdx.x >> 2, f = threadIdx.x & 3;
  const int64_t k0 = xcd_block(blockIdx.x, gridDim.x) * kQuadCells;
  const int64_t k = k0 + cl;
  const bool active = k < ncells;
  const int64_t kc = active ? k : ncells - 1;
  constexpr bool ST = kQuadStage<P, T>;
  __shared__ T sblk[ST ? kQuadCells * NL : 1];
  const int ncb = (int)(ncells - k0 < kQuadCells ? ncells - k0 : kQuadCells);
  if (ST) {
    for (int j = threadIdx.x; j < ncb * NL; j += 4 * kQuadCells) sblk[j] = src[k0 * NL + j];
    __syncthreads();
  }
  const dgt::LaneGeo<T> g = dgt::lane_geometry<P, T>(kc, f, ncells, bgeo, fgeo, nbrf, T(1));
  dgt::lane_component<P, T, 1, false, 0, ST>(g, kc, f, 0, T(0), T(1), src, ybuf[cl][f], sblk, k0, ncb);
  T* const yblk = quad_rows_to_block<P>(ybuf, cl, f);
  for (int j = threadIdx.x; j < ncb * NL; j += 4 * kQuadCells) apply_epilogue(ep, k0 * NL + j, yblk[j]);
}

// The non-temporal variant of the four-lane kernel for scalar FP64 products
// whose geometry, face records, neighbour words, src and dst exceed the
// 256 MB Infinity Cache (at 64^3 P2 566 MB; at 48^3 239 MB, whose data the
// cache keeps between the products of a solve).  Same runs at 64^3 against
// normal loads and stores: P1 0.0890 / 0.0900 against 0.0928 / 0.0949 ms, P2
// 0.1684 / 0.1702 against 0.1738 / 0.1749 ms, P3 0.394 / 0.390 against
// 0.408 / 0.406 ms.  FP32 products keep normal loads: P1 FP32 at 64^3 (239
// MB) stays in the cache from one product to the next and measured 0.051
// against 0.039 ms with the hints.
template <int P, typename T, int NC, bool MASS, int MODE, bool ST>
void launch_quad(const Operator& op, T* dst, const T* src, hipStream_t s) {
  const DGPlan& pl = op.dg;
  const dim3 grid((unsigned)((pl.n_cells + kQuadCells - 1) / kQuadCells * NC)), block(4 * kQuadCells);
  if constexpr (NC == 1 && sizeof(T) == 8 && MODE == 0) {
    const double bytes = (double)pl.n_cells * (7 * sizeof(T) + 4 * (dgt::kFaceRec * sizeof(T) + 4)) +
                         2.0 * sizeof(T) * op.m();
    if (bytes > 256.0 * (1 << 20)) {
      hipLaunchKernelGGL((dg_quad_kernel<P, T, NC, MASS, MODE, ST, true>), grid, block, 0, s, pl.n_cells,
                         op.d_dg_bgeo.as<T>(), op.d_dg_fgeo.as<T>(), op.d_dg_nbrf.as<int32_t>(), (T)pl.mass,
                         (T)pl.laplace, src, dst);
      return;
    }
  }
  hipLaunchKernelGGL((dg_quad_kernel<P, T, NC, MASS, MODE, ST>), grid, block, 0, s, pl.n_cells, op.d_dg_bgeo.as<T>(),
                     op.d_dg_fgeo.as<T>(), op.d_dg_nbrf.as<int32_t>(), (T)pl.mass, (T)pl.laplace, src, dst);
}

template <int P, typename T, int NC, bool MASS, int MODE>
void launch_dg_t(const Operator& op, T* dst, const T* src, hipStream_t s) {
  const DGPlan& pl = op.dg;
  if (pl.curved) return launch_dg_curved<P, T, NC, MASS, MODE>(op, dst, src, s);
  // auto: four lanes per cell (the products at every degree, the diagonal
  // from P2 up; at P1 64^3 FP64 0.098-0.100 against 0.115 ms for one lane per
  // cell, FP32 0.044 against 0.072 ms)
  if (MODE == 0 && (op.opt.dg_kernel == 4 || (op.opt.dg_kernel == 0 && kQuadStage<P, T>))) {
    if (pl.n_cells > 0) launch_quad<P, T, NC, MASS, MODE, true>(op, dst, src, s);
    SMF_HIP(hipGetLastError());
    return;
  }
  if (op.opt.dg_kernel == 3 || op.opt.dg_kernel == 4 || (op.opt.dg_kernel == 0 && (P >= 2 || MODE == 0))) {
    if (pl.n_cells > 0) launch_quad<P, T, NC, MASS, MODE, false>(op, dst, src, s);
    SMF_HIP(hipGetLastError());
    return;
  }
  // compile-time-table kernel except P3 FP64, which exceeds the register file
  const bool v2 = op.opt.dg_kernel == 2 || (op.opt.dg_kernel == 0 && (P <= 2 || sizeof(T) == 4));
  if (v2) {
    if (pl.n_cells > 0)
      hipLaunchKernelGGL((dg_cell_kernel_v2<P, T, NC, MASS, MODE>),
                         dim3((unsigned)((pl.n_cells + kThreads - 1) / kThreads)), dim3(kThreads), 0, s, pl.n_cells,
                         op.d_dg_geo.as<T>(), op.d_dg_nbr.as<int32_t>(), op.d_dg_nface.as<uint32_t>(),
                         (T)pl.penalty_factor, (T)pl.mass, (T)pl.laplace, src, dst, pl.penalty_kind);
    SMF_HIP(hipGetLastError());
    return;
  }
  using S = DGShape<P, MASS>;
  if (pl.n_q != S::NQ || pl.n_qf != S::NQF) throw std::runtime_error("DG table size mismatch");
  if (pl.penalty_kind != 0) throw std::runtime_error("penalty_kind 1 needs dg_kernel 0, 2 or 3");
  const size_t shmem = (size_t)(S::NQ * S::NL * 4 + S::NQ + 4 * S::NQF * S::NL * 4 + S::NQF) * sizeof(T);
  auto kern = dg_cell_kernel<P, T, NC, MASS, MODE>;
  if (shmem > 65536)
    SMF_HIP(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shmem));
  if (pl.n_cells > 0)
    hipLaunchKernelGGL(kern, dim3((unsigned)((pl.n_cells + kThreads - 1) / kThreads)), dim3(kThreads), shmem, s,
                       pl.n_cells, op.d_dg_geo.as<T>(), op.d_dg_nbr.as<int32_t>(), op.d_dg_nface.as<uint32_t>(),
                       op.d_dg_cell_tab.as<T>(), op.d_dg_face_tab.as<T>(), (T)pl.penalty_factor, (T)pl.mass,
                       (T)pl.laplace, src, dst);
  SMF_HIP(hipGetLastError());
}

template <int P, typename T, int MODE>
void launch_dg_p(const Operator& op, T* dst, const T* src, hipStream_t s) {
  if (op.kind == MF_DG_SIPG) launch_dg_t<P, T, 1, false, MODE>(op, dst, src, s);
  else if (op.kind == MF_DG_VECTOR) launch_dg_t<P, T, 3, false, MODE>(op, dst, src, s);
  else launch_dg_t<P, T, 3, true, MODE>(op, dst, src, s);
}

template <int P, typename T>
void launch_dg_ep_t(const Operator& op, const T* x, RowEpilogue<T> ep, hipStream_t s) {
  const DGPlan& pl = op.dg;
  if (pl.n_cells > 0)
    hipLaunchKernelGGL((dg_quad_ep_kernel<P, T>), dim3((unsigned)((pl.n_cells + kQuadCells - 1) / kQuadCells)),
                       dim3(4 * kQuadCells), 0, s, pl.n_cells, op.d_dg_bgeo.as<T>(), op.d_dg_fgeo.as<T>(),
                       op.d_dg_nbrf.as<int32_t>(), x, ep);
  SMF_HIP(hipGetLastError());
}

}  // namespace

// One degree and precision of the DG operators (every kind, mode and kernel
// variant); each is instantiated in a translation unit of its own
// (dg_p<P>_f<bits>.hip) so that the six compile in parallel.
template <int P, typename T>
void launch_dg_degree(const Operator& op, int mode, T* dst, const T* src, hipStream_t s) {
  if (mode == 0) launch_dg_p<P, T, 0>(op, dst, src, s);
  else launch_dg_p<P, T, 1>(op, dst, src, s);
}
template <int P, typename T>
void launch_dg_ep_degree(const Operator& op, const T* x, RowEpilogue<T> ep, hipStream_t s) {
  launch_dg_ep_t<P, T>(op, x, ep, s);
}

}  // namespace smf
