// Continuous-Galerkin cell kernels for gfx950.
//
// One workgroup of W wavefronts (W = 1..4, 64 W lanes) processes one patch of
// cells of the hierarchical reordering, in rounds of 64 W cells (one cell per
// lane):
//   1. the patch's DoF values are staged in LDS (owned DoFs are a contiguous
//      global range -> coalesced loads; ghost DoFs through an index list);
//   2. each lane evaluates its cell: gradients at the quadrature points
//      (interpolation with the reference shape-gradient matrix), the affine
//      quadrature-point operator w_q |det J| J^-1 J^-T and the transposed
//      interpolation (integration), in registers with FP64/FP32 FMA;
//   3. the round's cell vectors are added into an LDS accumulator: FP64 with
//      ds_add_f64 one wave after the other, in a wave order fixed per round
//      (W - 1 barriers per round); FP32 colour by colour with read-add-write
//      and a barrier after each colour (cells of one colour share no DoF);
//   4. owned DoFs are stored; ghost DoFs (owned by another patch) are written
//      as partial sums that a fix-up kernel adds to the owner's value in patch
//      order.  No floating-point atomics reach global memory: the result is
//      bit-reproducible.
// All patches run in one launch.
//
// Reference: do_cell_integral_masked_gather, CG/cells_batched/application.cc:
// 1236-1416 (masked gather 1272-1306, interpolation 1314-1319, affine D_e
// 1333-1360, integration 1381-1386, masked scatter-add 1388-1412, Dirichlet
// identity rows 1173-1175).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <stdexcept>
#include <string>

#include "epilogue.h"
#include "op.h"
#include "shape_tables.h"

namespace smf {

namespace {

constexpr int kWave = 64;
constexpr int kMaxThreads = 4 * kWave;

template <int P>
struct CGShape {
  static constexpr int NL = (P + 1) * (P + 2) * (P + 3) / 6;
  static constexpr int NQ = P == 1 ? 1 : (P == 2 ? 4 : 11);
  // curved cells: the rule of degree 2p
  static constexpr int NQC = P == 1 ? 4 : (P == 2 ? 11 : 24);
};

// P2 cell: y = A_K u in barycentric form at the symmetric 4-point rule
// (point m: lambda_m = a, the other three b, a + 3b = 1, weight 1/24 each).
// d u / d lambda_a = u_a (4 lambda_a - 1) + 4 sum_{b != a} u_ab lambda_b is
// linear in lambda, so at point m it is P_a + c N_am with
//   P_a = (4b - 1) u_a + 4b S_a,  S_a = sum_{b != a} u_ab,  c = 4(a - b),
//   N_am = u_a (a = m) or u_am (a != m),
// and the reference gradient is (D_1 - D_0, D_2 - D_0, D_3 - D_0).  The
// transposed step uses tau(m) = (-(t_0 + t_1 + t_2), t_0, t_1, t_2) of the
// flux t = w M g and T_a = sum_m tau_a(m):
//   y_v  = (4b - 1) T_v + c tau_v(v),
//   y_vw = 4b (T_v + T_w) + c (tau_v(w) + tau_w(v)).
// 138 FMA-class operations against 320 for the gradient-table form of the
// same quadrature (equal up to rounding).  Node order: vertices 0-3, then
// edges (01) (02) (03) (12) (13) (23).
template <typename T>
__device__ __forceinline__ void p2_cell_apply(const T* u, T m0, T m1, T m2, T m3, T m4, T m5, T* y) {
  constexpr double kB = q_tet_d2_pts[3][0], kA = q_tet_d2_pts[2][0], kW = q_tet_d2_w[0];
  static_assert(q_tet_d2_pts[0][2] == kA && q_tet_d2_pts[1][1] == kA && q_tet_d2_pts[0][0] == kB,
                "4-point rule with lambda_{3-q} = a at point q");
  static_assert(q_tet_d2_w[1] == kW && q_tet_d2_w[2] == kW && q_tet_d2_w[3] == kW, "equal weights");
  const T c4b1 = T(4 * kB - 1), c4b = T(4 * kB), c = T(4 * (kA - kB));
  // edge value of vertex pair (a, b), a != b
  constexpr int E[4][4] = {{-1, 4, 5, 6}, {4, -1, 7, 8}, {5, 7, -1, 9}, {6, 8, 9, -1}};
  T P[4];
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    int e[3], k = 0;
#pragma unroll
    for (int b = 0; b < 4; ++b)
      if (b != a) e[k++] = E[a][b];
    P[a] = c4b1 * u[a] + c4b * ((u[e[0]] + u[e[1]]) + u[e[2]]);
  }
  const T w = T(kW);
  const T M0 = w * m0, M1 = w * m1, M2 = w * m2, M3 = w * m3, M4 = w * m4, M5 = w * m5;
  T tau[4][4];  // tau[m][a]
#pragma unroll
  for (int m = 0; m < 4; ++m) {
    T D[4];
#pragma unroll
    for (int a = 0; a < 4; ++a) D[a] = P[a] + c * u[a == m ? a : E[a][m]];
    const T gx = D[1] - D[0], gy = D[2] - D[0], gz = D[3] - D[0];
    const T tx = M0 * gx + M1 * gy + M2 * gz;
    const T ty = M1 * gx + M3 * gy + M4 * gz;
    const T tz = M2 * gx + M4 * gy + M5 * gz;
    tau[m][0] = -((tx + ty) + tz);
    tau[m][1] = tx;
    tau[m][2] = ty;
    tau[m][3] = tz;
  }
  T Ts[4];
#pragma unroll
  for (int a = 0; a < 4; ++a) Ts[a] = ((tau[0][a] + tau[1][a]) + tau[2][a]) + tau[3][a];
#pragma unroll
  for (int v = 0; v < 4; ++v) y[v] = c4b1 * Ts[v] + c * tau[v][v];
#pragma unroll
  for (int v = 0; v < 4; ++v)
#pragma unroll
    for (int x = v + 1; x < 4; ++x) y[E[v][x]] = c4b * (Ts[v] + Ts[x]) + c * (tau[x][v] + tau[v][x]);
}

// Interface rows left pending by the previous step of a ChebyRun: while
// staging, ghost DoF g (interface row t = g - n_int) gets
//   d' = c1 ds[t] + c2 dinv[g] (b[g] - sum of part[fix_ptr[t]..fix_ptr[t+1]))
//   (d' = c2 dinv[g] (...) when first),  x' = xs[t] + d',
// the fix-up's epilogue in the same order and form; the first patch of the
// partial list stores x', d' to xw[t], dw[t].
template <typename T>
struct Pending {
  int on = 0, first = 0;
  T c1 = T(0), c2 = T(0);
  int64_t n_int = 0;
  const int32_t* fix_ptr = nullptr;
  const T* part = nullptr;
  const T* xs = nullptr;
  const T* ds = nullptr;
  const T* b = nullptr;
  const T* dinv = nullptr;
  T* xw = nullptr;
  T* dw = nullptr;
};

// sum of the partials part[k0..k1) in order, four loads in flight per step
// (most interface DoFs have two to four partials)
template <typename T>
__device__ __forceinline__ T sum_partials(const T* __restrict__ part, int32_t k0, int32_t k1) {
  T s = T(0);
  for (int32_t k = k0; k < k1; k += 4) {
    T v[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) v[u] = part[min(k + u, k1 - 1)];
#pragma unroll
    for (int u = 0; u < 4; ++u)
      if (k + u < k1) s = k + u == k0 ? v[u] : s + v[u];
  }
  return s;
}

// MODE 0: y = A_K u;  MODE 1: y_i = (A_K)_ii
// shape: G[(q*NL + i)*3 + d] = d phi_i / d xhat_d at point q, then w[q].
// ACC: how a round's cell vectors reach the LDS accumulator --
//   0: plain LDS read-add-write, one pass per colour (ds_add_f32 is slow on
//      gfx950);
//   1: ds_add (fast for f64 on gfx950), one pass per colour;
//   2: ds_add, one pass per wave: the waves of the workgroup add their whole
//      round in turn, in an order fixed per round (0..W-1 on even rounds,
//      W-1..0 on odd ones, so one wave adds twice in a row across the round
//      boundary and a round needs W-1 barriers).  Same-address lanes of one
//      ds_add instruction are applied by the LDS unit in its own fixed order,
//      so every DoF's sum has one order for given tables: bit-reproducible,
//      but not the cell order of ACC 0/1.
// NC components (components-batched vector Laplacian,
// CG/components_batched/application.cc:1024-1216) share the index tables and
// the geometry of a cell; component c of scalar DoF d is row c*cstride + d.
// FEAT: bit 0 = pending interface staging (ChebyRun steps), bit 1 = per-cell
// Nitsche boundary terms; both compile out of the plain operator kernels
// (they cost registers: P2 FP32 64 -> 120 VGPRs with both).  Bit 16:
// the packed indices and the metric, read once per product, are loaded
// non-temporally (FP64 products whose data exceeds the Infinity Cache)
template <int P, typename T, int MODE, int ACC, int NC, int FEAT>
__global__ __launch_bounds__(kMaxThreads) void cg_wave_kernel(
    const int32_t* __restrict__ cell_start, const int32_t* __restrict__ own_start,
    const int32_t* __restrict__ ghost_start, const int32_t* __restrict__ ghost_idx,
    const int32_t* __restrict__ pslot, const int32_t* __restrict__ gfix, const uint8_t* __restrict__ ccol,
    const uint16_t* __restrict__ lidx, const uint32_t* __restrict__ lidx12,
    int64_t ncells_total,
    const T* __restrict__ metric, const T* __restrict__ shape, const T* src,
    T* __restrict__ dst, T* __restrict__ partial, int64_t cstride, int64_t pstride, RowEpilogue<T> ep,
    Pending<T> pg, const int32_t* __restrict__ bcell, const T* __restrict__ bmat) {
  constexpr int NL = CGShape<P>::NL, NQ = CGShape<P>::NQ;
  extern __shared__ unsigned char smem_raw[];
  T* sbase = reinterpret_cast<T*>(smem_raw);
  const int lane = threadIdx.x;  // cell slot within a round
  const int nt = blockDim.x;     // 64 W
  const int patch = (int)xcd_block(blockIdx.x, gridDim.x);
  const int c0 = cell_start[patch], nc = cell_start[patch + 1] - c0;
  const int o0 = own_start[patch], no = own_start[patch + 1] - o0;
  const int g0 = ghost_start[patch], ng = ghost_start[patch + 1] - g0;
  const int nloc = no + ng;
  // per component: su (staged values) then sy (accumulator), nloc+1 each;
  // slot nloc is a dummy for constrained DoFs (index 0xFFFF clamps to it):
  // it reads as zero and absorbs their contributions, so no branches
  const int lstride = 2 * (nloc + 1);

  // 1. stage the patch DoFs.  Owned DoFs are the contiguous global range
  //    [o0, o0 + no) (coalesced); ghost DoFs come through ghost_idx.  Loads
  //    go in batches of SB per thread, all in flight before the first LDS
  //    write (clamped indices keep the spare loads in bounds).
  constexpr int SB = 8;
  for (int cc = 0; cc < NC; ++cc) {
    T* su = sbase + cc * lstride;
    T* sy = su + nloc + 1;
    const T* sc = src + cc * cstride;
    for (int j = lane; j <= nloc; j += nt) sy[j] = T(0);
    if (MODE == 0) {
      for (int j0 = 0; j0 < no; j0 += SB * nt) {
        T v[SB];
#pragma unroll
        for (int u = 0; u < SB; ++u) v[u] = sc[o0 + min(j0 + u * nt + lane, no - 1)];
#pragma unroll
        for (int u = 0; u < SB; ++u)
          if (j0 + u * nt + lane < no) su[j0 + u * nt + lane] = v[u];
      }
      if ((FEAT & 1) && pg.on) {
        // ghosts are interface DoFs whose previous step is still pending
        constexpr int GB = 4;
        for (int j0 = 0; j0 < ng; j0 += GB * nt) {
          // ghost_idx, pslot and gfix give the partial list [k0, k1) and
          // whether this patch is its first (two dependent round trips)
          int32_t gi[GB], sl[GB], gf[GB], k0[GB], k1[GB];
          T xv[GB], dv[GB], bv[GB], iv[GB];
#pragma unroll
          for (int u = 0; u < GB; ++u) {
            const int jj = min(j0 + u * nt + lane, ng - 1);
            gi[u] = ghost_idx[g0 + jj];
            sl[u] = pslot[g0 + jj];
            gf[u] = gfix[g0 + jj];
          }
#pragma unroll
          for (int u = 0; u < GB; ++u) {
            const int64_t t = gi[u] - pg.n_int;
            k0[u] = sl[u] - (gf[u] >> 16);
            k1[u] = k0[u] + (gf[u] & 0xFFFF);
            xv[u] = pg.xs[t];
            dv[u] = pg.first ? T(0) : pg.ds[t];
            bv[u] = pg.b[gi[u]];
            iv[u] = pg.dinv[gi[u]];
          }
#pragma unroll
          for (int u = 0; u < GB; ++u) {
            const T r = bv[u] - sum_partials(pg.part, k0[u], k1[u]);
            const T di = pg.first ? pg.c2 * iv[u] * r : pg.c1 * dv[u] + pg.c2 * iv[u] * r;
            const T xn = xv[u] + di;
            const int j = j0 + u * nt + lane;
            if (j < ng) {
              su[no + j] = xn;
              if ((gf[u] >> 16) == 0) {
                pg.xw[gi[u] - pg.n_int] = xn;
                pg.dw[gi[u] - pg.n_int] = di;
              }
            }
          }
        }
      } else {
        for (int j0 = 0; j0 < ng; j0 += SB * nt) {
          int32_t gi[SB];
          T v[SB];
#pragma unroll
          for (int u = 0; u < SB; ++u) gi[u] = ghost_idx[g0 + min(j0 + u * nt + lane, ng - 1)];
#pragma unroll
          for (int u = 0; u < SB; ++u) v[u] = sc[gi[u]];
#pragma unroll
          for (int u = 0; u < SB; ++u)
            if (j0 + u * nt + lane < ng) su[no + j0 + u * nt + lane] = v[u];
        }
      }
      if (lane == 0) su[nloc] = T(0);
    }
  }
  __syncthreads();

  // per-round cell data, loaded one round ahead (software pipelining); the
  // 12-bit packed indices (FEAT bit 8) are kept as words and decoded when
  // the round starts, so the loads stay in flight across the round before
  constexpr int NW = (FEAT & 8) ? (NL * 12 + 31) / 32 : 1;
  uint16_t li_n[NL];
  uint32_t lw_n[NW];
  T m_n[6];
  int col_n = 0, bc_n = -1;
  auto load_round = [&](int r0) {
    const int64_t cell = c0 + min(r0 + lane, nc - 1);
    if constexpr ((FEAT & 8) != 0) {
#pragma unroll
      for (int w = 0; w < NW; ++w) {
        if constexpr ((FEAT & 16) != 0) lw_n[w] = __builtin_nontemporal_load(lidx12 + (int64_t)w * ncells_total + cell);
        else lw_n[w] = lidx12[(int64_t)w * ncells_total + cell];
      }
    } else {
#pragma unroll
      for (int i = 0; i < NL; ++i) li_n[i] = lidx[(int64_t)i * ncells_total + cell];
    }
    if constexpr ((FEAT & 4) == 0) {
#pragma unroll
      for (int t = 0; t < 6; ++t) {
        if constexpr ((FEAT & 16) != 0) m_n[t] = __builtin_nontemporal_load(metric + t * ncells_total + cell);
        else m_n[t] = metric[t * ncells_total + cell];
      }
    }
    if constexpr ((FEAT & 8) == 0) col_n = ccol[cell];  // packed: decoded at the round start
    if ((FEAT & 2) && bcell) bc_n = bcell[cell];
  };
  if (nc > 0) load_round(0);
  for (int r0 = 0; r0 < nc; r0 += nt) {
    const int lc = r0 + lane;
    const bool active = lc < nc;
    uint16_t li[NL];
#pragma unroll
    for (int i = 0; i < NL; ++i) {
      if constexpr ((FEAT & 8) != 0) {
        // index i: bits 12 i .. 12 i + 11; 0xFFF (constrained) clamps to the dummy slot
        const int b = 12 * i, w = b >> 5, sh = b & 31;
        uint32_t v = lw_n[w] >> sh;
        if (sh > 20) v |= lw_n[w + 1] << (32 - sh);
        li[i] = min((int)(v & 0xFFFu), nloc);
      } else {
        li[i] = min((int)li_n[i], nloc);
      }
    }
    const T m0 = m_n[0], m1 = m_n[1], m2 = m_n[2], m3 = m_n[3], m4 = m_n[4], m5 = m_n[5];
    // packed: the colour follows the indices (bits 12 NL .. 12 NL + 7)
    int mycol = col_n;
    if constexpr ((FEAT & 8) != 0) mycol = (int)((lw_n[(NL * 12) >> 5] >> ((NL * 12) & 31)) & 0xFFu);
    const int mybc = bc_n;
    if (r0 + nt < nc) load_round(r0 + nt);
#pragma unroll 1
    for (int cc = 0; cc < NC; ++cc) {
    T* su = sbase + cc * lstride;
    T* sy = su + nloc + 1;
    T y[NL];
#pragma unroll
    for (int i = 0; i < NL; ++i) y[i] = T(0);
    // Nitsche boundary terms of the cell (Chebyshev runs of Nitsche levels)
    auto add_boundary = [&](const T* uu) {
      if (!(FEAT & 2) || mybc < 0) return;
      const T* B = bmat + (int64_t)mybc * NL * NL;
#pragma unroll
      for (int i = 0; i < NL; ++i) {
        T sb = T(0);
#pragma unroll
        for (int j = 0; j < NL; ++j) sb += B[i * NL + j] * uu[j];
        y[i] += sb;
      }
    };
    // P1/P2: compile-time tables (literal constants after unrolling); P3
    // has too many distinct values for scalar registers and reads the table
    // from memory one quadrature point at a time.  Curved cells (FEAT bit 4)
    // read the gradient table of the degree-2p rule the same way (P2: the
    // barycentric form at the point instead) and form
    // the metric w_q |det J_q| J_q^-1 J_q^-T per point in registers from the
    // cell's 27 coefficients of J(xi) = C + sum_f xi_f D_f (a P2 map has an
    // affine Jacobian): 216 B per cell instead of 48 B per point (the
    // general D_e branch, CG/cells_batched/application.cc:1361-1376).
    if constexpr ((FEAT & 4) != 0) {
      constexpr int NQC = CGShape<P>::NQC;
      const T* W = shape + NQC * NL * 3;  // weights, then the points (xi_0, xi_1, xi_2) per point
      const T* XI = W + NQC;
      const T* jg = metric + (c0 + min(lc, nc - 1));
      T jc[27];
#pragma unroll
      for (int t = 0; t < 27; ++t) jc[t] = jg[(int64_t)t * ncells_total];
      T u[NL];
      if (MODE == 0) {
#pragma unroll
        for (int i = 0; i < NL; ++i) u[i] = su[li[i]];
      }
#pragma unroll 1
      for (int q = 0; q < NQC; ++q) {
        const T x0 = XI[3 * q], x1 = XI[3 * q + 1], x2 = XI[3 * q + 2];
        // J[d][e] = C_e + xi_e E_e + sum_{f != e} xi_f F_ef
        T J[3][3];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          J[d][0] = jc[d] + x0 * jc[9 + d] + x1 * jc[18 + d] + x2 * jc[21 + d];
          J[d][1] = jc[3 + d] + x1 * jc[12 + d] + x0 * jc[18 + d] + x2 * jc[24 + d];
          J[d][2] = jc[6 + d] + x2 * jc[15 + d] + x0 * jc[21 + d] + x1 * jc[24 + d];
        }
        // adj = det J^-1; metric w |det J| J^-1 J^-T = (w / det) adj adj^T
        const T a00 = J[1][1] * J[2][2] - J[1][2] * J[2][1], a01 = J[0][2] * J[2][1] - J[0][1] * J[2][2],
                a02 = J[0][1] * J[1][2] - J[0][2] * J[1][1], a10 = J[1][2] * J[2][0] - J[1][0] * J[2][2],
                a11 = J[0][0] * J[2][2] - J[0][2] * J[2][0], a12 = J[0][2] * J[1][0] - J[0][0] * J[1][2],
                a20 = J[1][0] * J[2][1] - J[1][1] * J[2][0], a21 = J[0][1] * J[2][0] - J[0][0] * J[2][1],
                a22 = J[0][0] * J[1][1] - J[0][1] * J[1][0];
        const T det = J[0][0] * a00 + J[0][1] * a10 + J[0][2] * a20;
        const T sc = W[q] / det;
        // flux t = (w / det) adj (adj^T g) (vmult: the metric is not formed)
        auto flux = [&](T gx, T gy, T gz, T& tx, T& ty, T& tz) {
          const T h0 = a00 * gx + a10 * gy + a20 * gz;
          const T h1 = a01 * gx + a11 * gy + a21 * gz;
          const T h2 = a02 * gx + a12 * gy + a22 * gz;
          tx = sc * (a00 * h0 + a01 * h1 + a02 * h2);
          ty = sc * (a10 * h0 + a11 * h1 + a12 * h2);
          tz = sc * (a20 * h0 + a21 * h1 + a22 * h2);
        };
        const T* G = shape + q * NL * 3;
        if (P == 2 && MODE == 0) {
          // barycentric P2 at a general point (lambda = (1 - x0 - x1 - x2, x0, x1, x2)):
          // D_a = u_a (4 l_a - 1) + sum_{b != a} u_ab 4 l_b, reference gradient
          // g_e = D_{e+1} - D_0; transposed with tau = (-(t0 + t1 + t2), t0, t1, t2):
          // y_a += tau_a (4 l_a - 1), y_ab += tau_a 4 l_b + tau_b 4 l_a
          const T l4[4] = {T(4) * (T(1) - x0 - x1 - x2), T(4) * x0, T(4) * x1, T(4) * x2};
          constexpr int E[4][4] = {{-1, 4, 5, 6}, {4, -1, 7, 8}, {5, 7, -1, 9}, {6, 8, 9, -1}};  // node of edge (a, b)
          T D[4];
#pragma unroll
          for (int a = 0; a < 4; ++a) {
            T s = u[a] * (l4[a] - T(1));
#pragma unroll
            for (int b = 0; b < 4; ++b)
              if (b != a) s += u[E[a][b]] * l4[b];
            D[a] = s;
          }
          T tx, ty, tz;
          flux(D[1] - D[0], D[2] - D[0], D[3] - D[0], tx, ty, tz);
          const T tau[4] = {-(tx + ty + tz), tx, ty, tz};
#pragma unroll
          for (int a = 0; a < 4; ++a) y[a] += tau[a] * (l4[a] - T(1));
#pragma unroll
          for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = a + 1; b < 4; ++b) y[E[a][b]] += tau[a] * l4[b] + tau[b] * l4[a];
        } else if (MODE == 0) {
          T gx = T(0), gy = T(0), gz = T(0);
#pragma unroll
          for (int i = 0; i < NL; ++i) {
            gx += G[i * 3 + 0] * u[i];
            gy += G[i * 3 + 1] * u[i];
            gz += G[i * 3 + 2] * u[i];
          }
          T tx, ty, tz;
          flux(gx, gy, gz, tx, ty, tz);
#pragma unroll
          for (int i = 0; i < NL; ++i) y[i] += G[i * 3 + 0] * tx + G[i * 3 + 1] * ty + G[i * 3 + 2] * tz;
        } else {
          // diagonal: the metric (w / det) adj adj^T
          const T mq0 = sc * (a00 * a00 + a01 * a01 + a02 * a02), mq1 = sc * (a00 * a10 + a01 * a11 + a02 * a12),
                  mq2 = sc * (a00 * a20 + a01 * a21 + a02 * a22), mq3 = sc * (a10 * a10 + a11 * a11 + a12 * a12),
                  mq4 = sc * (a10 * a20 + a11 * a21 + a12 * a22), mq5 = sc * (a20 * a20 + a21 * a21 + a22 * a22);
#pragma unroll
          for (int i = 0; i < NL; ++i) {
            const T gx = G[i * 3 + 0], gy = G[i * 3 + 1], gz = G[i * 3 + 2];
            y[i] += gx * (mq0 * gx + mq1 * gy + mq2 * gz) + gy * (mq1 * gx + mq3 * gy + mq4 * gz) +
                    gz * (mq2 * gx + mq4 * gy + mq5 * gz);
          }
        }
      }
    } else if constexpr (P == 2 && MODE == 0) {
      T u[NL];
#pragma unroll
      for (int i = 0; i < NL; ++i) u[i] = su[li[i]];
      p2_cell_apply<T>(u, m0, m1, m2, m3, m4, m5, y);
      add_boundary(u);
    } else if constexpr (P <= 2) {
      constexpr tab::CellGrad<P> TB = tab::make_cell_grad<P>();
      if (MODE == 0) {
        T u[NL];
#pragma unroll
        for (int i = 0; i < NL; ++i) u[i] = su[li[i]];
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
          T gx = T(0), gy = T(0), gz = T(0);
#pragma unroll
          for (int i = 0; i < NL; ++i) {
            gx += T(TB.g[q][i][0]) * u[i];
            gy += T(TB.g[q][i][1]) * u[i];
            gz += T(TB.g[q][i][2]) * u[i];
          }
          const T w = T(TB.w[q]);
          const T tx = w * (m0 * gx + m1 * gy + m2 * gz);
          const T ty = w * (m1 * gx + m3 * gy + m4 * gz);
          const T tz = w * (m2 * gx + m4 * gy + m5 * gz);
#pragma unroll
          for (int i = 0; i < NL; ++i)
            y[i] += T(TB.g[q][i][0]) * tx + T(TB.g[q][i][1]) * ty + T(TB.g[q][i][2]) * tz;
        }
        add_boundary(u);
      } else {
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
          const T w = T(TB.w[q]);
#pragma unroll
          for (int i = 0; i < NL; ++i) {
            const T gx = T(TB.g[q][i][0]), gy = T(TB.g[q][i][1]), gz = T(TB.g[q][i][2]);
            y[i] += w * (gx * (m0 * gx + m1 * gy + m2 * gz) + gy * (m1 * gx + m3 * gy + m4 * gz) +
                         gz * (m2 * gx + m4 * gy + m5 * gz));
          }
        }
      }
    } else {
      const T* W = shape + NQ * NL * 3;
      if (MODE == 0) {
        T u[NL];
#pragma unroll
        for (int i = 0; i < NL; ++i) u[i] = su[li[i]];
#pragma unroll 1
        for (int q = 0; q < NQ; ++q) {
          const T* G = shape + q * NL * 3;
          T gx = T(0), gy = T(0), gz = T(0);
#pragma unroll
          for (int i = 0; i < NL; ++i) {
            gx += G[i * 3 + 0] * u[i];
            gy += G[i * 3 + 1] * u[i];
            gz += G[i * 3 + 2] * u[i];
          }
          const T w = W[q];
          const T tx = w * (m0 * gx + m1 * gy + m2 * gz);
          const T ty = w * (m1 * gx + m3 * gy + m4 * gz);
          const T tz = w * (m2 * gx + m4 * gy + m5 * gz);
#pragma unroll
          for (int i = 0; i < NL; ++i) y[i] += G[i * 3 + 0] * tx + G[i * 3 + 1] * ty + G[i * 3 + 2] * tz;
        }
        add_boundary(u);
      } else {
#pragma unroll 1
        for (int q = 0; q < NQ; ++q) {
          const T* G = shape + q * NL * 3;
          const T w = W[q];
#pragma unroll
          for (int i = 0; i < NL; ++i) {
            const T gx = G[i * 3 + 0], gy = G[i * 3 + 1], gz = G[i * 3 + 2];
            y[i] += w * (gx * (m0 * gx + m1 * gy + m2 * gz) + gy * (m1 * gx + m3 * gy + m4 * gz) +
                         gz * (m2 * gx + m4 * gy + m5 * gz));
          }
        }
      }
    }
    // 3. accumulation.  ACC 2: wave-ordered passes (see above)
    if constexpr (ACC == 2) {
      const int nw = nt / kWave, wave = lane / kWave;
      const bool fwd = ((r0 / nt) & 1) == 0;
      for (int k = 0; k < nw; ++k) {
        if (wave == (fwd ? k : nw - 1 - k) && active) {
#pragma unroll
          for (int i = 0; i < NL; ++i) atomicAdd(&sy[li[i]], y[i]);
        }
        if (k + 1 < nw) __syncthreads();
      }
      continue;
    }
    // ACC 0/1: colour-ordered accumulation (colours ascend with the cell
    // index); the colour range of the round is uniform over the workgroup
    const int clo = ccol[c0 + r0];
    const int last = min(r0 + nt, nc) - 1;
    const int chi = ccol[c0 + last];
    for (int c = clo; c <= chi; ++c) {
      if (active && mycol == c) {
        if (ACC == 1) {
#pragma unroll
          for (int i = 0; i < NL; ++i) atomicAdd(&sy[li[i]], y[i]);
        } else {
          // distinct addresses within a colour except the dummy slot, whose
          // value is never read
          T v[NL];
#pragma unroll
          for (int i = 0; i < NL; ++i) v[i] = sy[li[i]];
#pragma unroll
          for (int i = 0; i < NL; ++i) sy[li[i]] = v[i] + y[i];
        }
      }
      if (nt > kWave) __syncthreads();  // one wave alone keeps its LDS order
    }
    }  // components
  }
  __syncthreads();

  // 4. owned DoFs: store (or the row epilogue: owned DoFs are read by this
  //    patch only); ghost DoFs: partial sums for the fix-up
  for (int cc = 0; cc < NC; ++cc) {
    const T* sy = sbase + cc * lstride + nloc + 1;
    T* dc = dst + cc * cstride;
    T* pc = partial + cc * pstride;
    if (ep.kind == 0) {
      for (int j = lane; j < no; j += nt) dc[o0 + j] = sy[j];
    } else {
      epilogue_rows<T, 4>(ep, o0, no, lane, nt, sy);
    }
    for (int j = lane; j < ng; j += nt) pc[pslot[g0 + j]] = sy[no + j];
  }
}

// Interface DoF n_int + t = sum of its patch partials partial[k], k in
// [fix_ptr[t], fix_ptr[t+1]) (contiguous, patch order); Dirichlet rows
// dst = src (MODE 0) or 1 (MODE 1).  Every load that does not depend on
// fix_ptr (the rows' epilogue inputs, the Dirichlet source) is issued first.
template <typename T, int MODE>
__global__ void cg_fixup_kernel(int64_t n_int, int64_t n_iface, const int32_t* __restrict__ fix_ptr,
                                const T* __restrict__ partial, int64_t n0, int64_t n1, const T* src,
                                T* __restrict__ dst, int nc, int64_t cstride, int64_t pstride, RowEpilogue<T> ep) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool iface = t < n_iface, dir = n0 + t < n1;
  const int64_t row = n_int + (iface ? t : 0), j = dir ? n0 + t : n0;
  for (int c = 0; c < nc; ++c) {
    const T* pc = partial + c * pstride;
    EpilogueIn<T> in_i{}, in_d{};
    T ax_d = T(1);
    if (dir && MODE == 0) ax_d = src[c * cstride + j];
    if (ep.kind != 0) {
      if (iface) in_i = epilogue_load(ep, row);
      if (dir) in_d = epilogue_load(ep, j);
    }
    if (iface) {
      const T s = sum_partials(pc, fix_ptr[t], fix_ptr[t + 1]);
      if (ep.kind == 0) dst[c * cstride + row] = s;
      else epilogue_finish(ep, row, s, in_i);
    }
    if (dir) {
      if (ep.kind == 0) dst[c * cstride + j] = ax_d;
      else epilogue_finish(ep, j, ax_d, in_d);
    }
  }
}

// partial_out/pg/fixup: a ChebyRun step (its own partial buffer, pending
// interface rows of the previous step, no fix-up launch)
template <int P, typename T, int MODE, int NC, int FEAT>
void launch_cg_k(const Operator& op, T* dst, const T* src, hipStream_t s, RowEpilogue<T> ep, Pending<T> pg,
                 T* partial_out, bool fixup, bool cell_boundary) {
  const CGPlan& pl = op.cg;
  const size_t shmem = NC * 2 * ((size_t)pl.max_local + 1) * sizeof(T);
  // accumulate option: 1 atomics per colour, 2 read-add-write per colour,
  // 3 atomics per wave; 0: FP64 atomics per wave (P2 FP64 128^3 0.348 ->
  // 0.319 ms, vector P2 64^3 0.168 -> 0.122 ms against atomics per colour),
  // FP32 read-add-write per colour (FP32 atomics are 3x slower)
  const int acc = op.opt.accumulate == 3 || (op.opt.accumulate == 0 && sizeof(T) == 8) ? 2
                  : op.opt.accumulate == 1 ? 1 : 0;
  auto kern = acc == 2 ? cg_wave_kernel<P, T, MODE, 2, NC, FEAT>
              : acc == 1 ? cg_wave_kernel<P, T, MODE, 1, NC, FEAT> : cg_wave_kernel<P, T, MODE, 0, NC, FEAT>;
  if (shmem > 160 * 1024)
    throw std::runtime_error("CG patch needs " + std::to_string(shmem) + " bytes of LDS (max 160 KiB): use smaller patch_cells");
  if (shmem > 65536)
    SMF_HIP(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shmem));
  if (pl.waves < 1 || pl.waves > kMaxThreads / kWave) throw std::runtime_error("CG plan: 1 to 4 waves per patch");
  if (pl.n_patches > 0)
    hipLaunchKernelGGL(kern, dim3((unsigned)pl.n_patches), dim3(kWave * pl.waves), shmem, s,
                       op.d_cell_start.as<int32_t>(), op.d_own_start.as<int32_t>(),
                       op.d_ghost_start.as<int32_t>(), op.d_ghost_idx.as<int32_t>(), op.d_pslot.as<int32_t>(),
                       op.d_gfix.as<int32_t>(), op.d_ccol.as<uint8_t>(), op.d_lidx.as<uint16_t>(),
                       op.d_lidx12.as<uint32_t>(), pl.n_cells,
                       op.d_metric.as<T>(), op.d_shape.as<T>(), src, dst,
                       partial_out ? partial_out : op.d_partial.as<T>(), pl.n_dofs,
                       std::max<int64_t>(pl.n_ghost_total, 1), ep, pg,
                       cell_boundary ? op.d_bcell.as<int32_t>() : nullptr,
                       cell_boundary ? op.d_bmat.as<T>() : nullptr);
  if (!fixup) {
    SMF_HIP(hipGetLastError());
    return;
  }
  const int64_t nif = pl.n_free - pl.n_int;
  const int64_t ncon = pl.n_dofs - pl.n_free;
  const int64_t nthreads = std::max(nif, ncon);
  if (nthreads > 0)
    hipLaunchKernelGGL((cg_fixup_kernel<T, MODE>), dim3((unsigned)((nthreads + 255) / 256)), dim3(256),
                       0, s, pl.n_int, nif, op.d_fix_ptr.as<int32_t>(), op.d_partial.as<T>(), pl.n_free, pl.n_dofs,
                       src, dst, NC, pl.n_dofs, std::max<int64_t>(pl.n_ghost_total, 1), ep);
  SMF_HIP(hipGetLastError());
}

template <int P, typename T, int MODE, int NC>
void launch_cg_t(const Operator& op, T* dst, const T* src, hipStream_t s, RowEpilogue<T> ep = {},
                 Pending<T> pg = {}, T* partial_out = nullptr, bool fixup = true, bool cell_boundary = false) {
  // 12-bit packed local indices when the plan has them (P >= 2)
  if constexpr (P >= 2) {
    if (op.cg.lidx12_words > 0 && !op.cg.curved) {
      if (op.cg.lidx12_words != (CGShape<P>::NL * 12 + 31) / 32)
        throw std::runtime_error("CG kernel: packed index width does not match the degree");
      if constexpr (MODE == 0 && NC == 1) {
        if (partial_out && cell_boundary)
          return launch_cg_k<P, T, MODE, NC, 11>(op, dst, src, s, ep, pg, partial_out, fixup, true);
        if (partial_out) return launch_cg_k<P, T, MODE, NC, 9>(op, dst, src, s, ep, pg, partial_out, fixup, false);
      }
      if (pg.on || cell_boundary) throw std::runtime_error("CG kernel: pending staging / cell boundary need a Chebyshev run");
      if constexpr (MODE == 0 && NC == 1 && sizeof(T) == 8) {
        // indices, metric, src and dst beyond the 256 MB Infinity Cache
        const CGPlan& pl = op.cg;
        const double bytes = (double)pl.n_cells * (4.0 * op.cg.lidx12_words + 6.0 * sizeof(T)) +
                             2.0 * sizeof(T) * pl.n_dofs;
        if (ep.kind == 0 && bytes > 256.0 * (1 << 20))
          return launch_cg_k<P, T, MODE, NC, 8 | 16>(op, dst, src, s, ep, pg, partial_out, fixup, false);
      }
      return launch_cg_k<P, T, MODE, NC, 8>(op, dst, src, s, ep, pg, partial_out, fixup, false);
    }
  }
  if (op.cg.curved) {
    if (pg.on || cell_boundary || partial_out) throw std::runtime_error("CG kernel: curved cells take no Chebyshev run");
    if (op.cg.n_q != CGShape<P>::NQC) throw std::runtime_error("CG kernel: curved-cell rule does not match the kernel");
    return launch_cg_k<P, T, MODE, NC, 4>(op, dst, src, s, ep, pg, partial_out, fixup, false);
  }
  if constexpr (MODE == 0 && NC == 1) {
    if (partial_out && cell_boundary)
      return launch_cg_k<P, T, MODE, NC, 3>(op, dst, src, s, ep, pg, partial_out, fixup, true);
    if (partial_out) return launch_cg_k<P, T, MODE, NC, 1>(op, dst, src, s, ep, pg, partial_out, fixup, false);
  }
  if (pg.on || cell_boundary) throw std::runtime_error("CG kernel: pending staging / cell boundary need a Chebyshev run");
  launch_cg_k<P, T, MODE, NC, 0>(op, dst, src, s, ep, pg, partial_out, fixup, false);
}

// Nitsche boundary rows: dst[row] += sum_k val[k] src[col[k]] (MODE 0) or
// dst[row] += B[row][row] (MODE 1).  G consecutive lanes share a row: lane s
// sums entries s, s + G, ... (coalesced column/value reads), then a
// fixed-order butterfly combines the G partial sums.  Every row has one
// writer, so the result is deterministic.
template <typename T, int MODE, bool STORE, int G>
__global__ __launch_bounds__(256) void cg_boundary_kernel(int64_t nrows, const int32_t* __restrict__ rows,
                                                          const int32_t* __restrict__ ptr,
                                                          const int32_t* __restrict__ col,
                                                          const T* __restrict__ val, const T* __restrict__ src,
                                                          T* __restrict__ dst) {
  const int64_t r = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / G;
  const int sub = threadIdx.x % G;
  const bool ok = r < nrows;  // no early exit: every lane takes part in the butterfly
  T s = T(0);
  if (ok) {
    const int32_t row = rows[r], k1 = ptr[r + 1];
    for (int32_t k = ptr[r] + sub; k < k1; k += G) {
      if (MODE == 0) s += val[k] * src[col[k]];
      else if (col[k] == row) s += val[k];
    }
  }
#pragma unroll
  for (int off = G / 2; off > 0; off >>= 1) s += __shfl_xor(s, off, G);
  if (ok && sub == 0) {
    const int32_t row = rows[r];
    if (STORE) dst[row] = s;
    else dst[row] += s;
  }
}

// lanes per row from the mean row length (about 4 entries per lane)
template <typename T, int MODE, bool STORE>
void launch_boundary_rows(const Operator& op, T* dst, const T* src, hipStream_t s) {
  const int64_t nr = (int64_t)op.cg.bnd_rows.size();
  if (nr == 0) return;
  const int64_t mean = (int64_t)op.cg.bnd_col.size() / nr;
  const int g = mean >= 48 ? 16 : (mean >= 20 ? 8 : 4);
  const unsigned blocks = (unsigned)((nr * g + 255) / 256);
  auto args = [&](auto kern) {
    hipLaunchKernelGGL(kern, dim3(blocks), dim3(256), 0, s, nr, op.d_bnd_rows.as<int32_t>(),
                       op.d_bnd_ptr.as<int32_t>(), op.d_bnd_col.as<int32_t>(), op.d_bnd_val.as<T>(), src, dst);
  };
  if (g == 16) args(cg_boundary_kernel<T, MODE, STORE, 16>);
  else if (g == 8) args(cg_boundary_kernel<T, MODE, STORE, 8>);
  else args(cg_boundary_kernel<T, MODE, STORE, 4>);
  SMF_HIP(hipGetLastError());
}

template <typename T, int MODE>
void launch_boundary_t(const Operator& op, T* dst, const T* src, hipStream_t s) {
  launch_boundary_rows<T, MODE, false>(op, dst, src, s);
}

template <typename T, int NC>
void launch_cg_nc(const Operator& op, int mode, T* dst, const T* src, hipStream_t s) {
  switch (op.cg.p * 2 + mode) {
    case 2: launch_cg_t<1, T, 0, NC>(op, dst, src, s); break;
    case 3: launch_cg_t<1, T, 1, NC>(op, dst, src, s); break;
    case 4: launch_cg_t<2, T, 0, NC>(op, dst, src, s); break;
    case 5: launch_cg_t<2, T, 1, NC>(op, dst, src, s); break;
    case 6: launch_cg_t<3, T, 0, NC>(op, dst, src, s); break;
    case 7: launch_cg_t<3, T, 1, NC>(op, dst, src, s); break;
    default: throw std::runtime_error("unsupported CG degree");
  }
}

template <typename T>
void launch_cg_prec(const Operator& op, int mode, T* dst, const T* src, hipStream_t s) {
  if (op.ncomp == 3) launch_cg_nc<T, 3>(op, mode, dst, src, s);
  else launch_cg_nc<T, 1>(op, mode, dst, src, s);
  if (op.cg.nitsche) {
    if (mode == 0) launch_boundary_t<T, 0>(op, dst, src, s);
    else launch_boundary_t<T, 1>(op, dst, src, s);
  }
}

}  // namespace

void launch_cg(const Operator& op, int mode, void* dst, const void* src, hipStream_t s) {
  if (op.precision == MF_F64)
    launch_cg_prec<double>(op, mode, static_cast<double*>(dst), static_cast<const double*>(src), s);
  else
    launch_cg_prec<float>(op, mode, static_cast<float*>(dst), static_cast<const float*>(src), s);
}

template <typename T>
void launch_cg_epilogue(const Operator& op, const T* x, RowEpilogue<T> ep, hipStream_t s) {
  if (op.is_dg() || op.ncomp != 1 || ep.kind == 0) throw std::runtime_error("CG epilogue: scalar CG, kind 1 or 2");
  if ((op.precision == MF_F64) != (sizeof(T) == 8)) throw std::runtime_error("CG epilogue: precision mismatch");
  if (op.cg.nitsche) {
    // boundary row sums first, read by the epilogues
    launch_boundary_rows<T, 0, true>(op, op.d_bnd_dense.as<T>(), x, s);
    ep.bnd = op.d_bnd_dense.as<T>();
  }
  switch (op.cg.p) {
    case 1: launch_cg_t<1, T, 0, 1>(op, nullptr, x, s, ep); break;
    case 2: launch_cg_t<2, T, 0, 1>(op, nullptr, x, s, ep); break;
    case 3: launch_cg_t<3, T, 0, 1>(op, nullptr, x, s, ep); break;
    default: throw std::runtime_error("unsupported CG degree");
  }
}
template void launch_cg_epilogue<double>(const Operator&, const double*, RowEpilogue<double>, hipStream_t);
template void launch_cg_epilogue<float>(const Operator&, const float*, RowEpilogue<float>, hipStream_t);

namespace {

template <typename T>
struct ChebySteps {
  int n = 0;
  ChebyStep<T> s[kMaxChebySteps];
};

// Close of a ChebyRun: interface row t gets the last step's pending update
// (pg, written to x, d at n_int + t); Dirichlet row j (A row = identity)
// runs every step of the run: d = c1 d + c2 dinv (b - x), x += d.
template <typename T>
__global__ void cg_cheby_close_kernel(int64_t n_iface, Pending<T> pg, int64_t n0, int64_t n1, T* __restrict__ x,
                                      T* __restrict__ d, ChebySteps<T> st) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n_iface) {
    const int64_t row = pg.n_int + t;
    const T xv = pg.xs[t], dv = pg.first ? T(0) : pg.ds[t], bv = pg.b[row], iv = pg.dinv[row];
    const T r = bv - sum_partials(pg.part, pg.fix_ptr[t], pg.fix_ptr[t + 1]);
    const T di = pg.first ? pg.c2 * iv * r : pg.c1 * dv + pg.c2 * iv * r;
    x[row] = xv + di;
    d[row] = di;
  }
  const int64_t j = n0 + t;
  if (j < n1) {
    T xv = x[j], dv = d[j];
    const T bv = pg.b[j], iv = pg.dinv[j];
    for (int k = 0; k < st.n; ++k) {
      const T r = bv - xv;
      const T di = st.s[k].first ? st.s[k].c2 * iv * r : st.s[k].c1 * dv + st.s[k].c2 * iv * r;
      dv = di;
      xv = xv + di;
    }
    x[j] = xv;
    d[j] = dv;
  }
}

// pending update of run step k (interface rows of step k, applied by step k + 1 or the close)
template <typename T>
Pending<T> pending_of(const Operator& op, const ChebyRun<T>& run, int k) {
  Pending<T> pg;
  pg.on = 1;
  pg.first = run.steps[k].first;
  pg.c1 = run.steps[k].c1;
  pg.c2 = run.steps[k].c2;
  pg.n_int = op.cg.n_int;
  pg.fix_ptr = op.d_fix_ptr.as<int32_t>();
  pg.part = run.part[k % 2];
  pg.xs = k == 0 ? run.x + op.cg.n_int : run.xi[k % 2];
  pg.ds = k == 0 ? run.d + op.cg.n_int : run.di[k % 2];
  pg.b = run.b;
  pg.dinv = run.dinv;
  return pg;
}

}  // namespace

bool cg_has_cheby_run(const Operator& op) { return !op.is_dg() && op.ncomp == 1 && !op.cg.curved; }

template <typename T>
void launch_cg_cheby_step(const Operator& op, ChebyRun<T>& run, T c1, T c2, int first, hipStream_t s) {
  if (!cg_has_cheby_run(op)) throw std::runtime_error("Chebyshev run: scalar CG operator");
  if ((op.precision == MF_F64) != (sizeof(T) == 8)) throw std::runtime_error("Chebyshev run: precision mismatch");
  const int k = run.n;
  if (k >= kMaxChebySteps) throw std::runtime_error("Chebyshev run: too many steps");
  RowEpilogue<T> ep;
  ep.kind = 1;
  ep.first = first;
  ep.c1 = c1;
  ep.c2 = c2;
  ep.b = run.b;
  ep.dinv = run.dinv;
  ep.d = run.d;
  ep.x = run.x;
  Pending<T> pg;
  if (k > 0) {
    pg = pending_of(op, run, k - 1);
    pg.xw = run.xi[k % 2];
    pg.dw = run.di[k % 2];
  }
  // Nitsche rows: the boundary terms are added per cell inside the kernel,
  // which stages the pending interface values they read
  const bool nb = op.cg.nitsche;
  switch (op.cg.p) {
    case 1: launch_cg_t<1, T, 0, 1>(op, nullptr, run.x, s, ep, pg, run.part[k % 2], false, nb); break;
    case 2: launch_cg_t<2, T, 0, 1>(op, nullptr, run.x, s, ep, pg, run.part[k % 2], false, nb); break;
    case 3: launch_cg_t<3, T, 0, 1>(op, nullptr, run.x, s, ep, pg, run.part[k % 2], false, nb); break;
    default: throw std::runtime_error("unsupported CG degree");
  }
  run.steps[k] = ChebyStep<T>{c1, c2, first};
  run.n = k + 1;
}

template <typename T>
void launch_cg_cheby_close(const Operator& op, ChebyRun<T>& run, hipStream_t s) {
  if (run.n == 0) return;
  const CGPlan& pl = op.cg;
  const Pending<T> pg = pending_of(op, run, run.n - 1);
  ChebySteps<T> st;
  st.n = run.n;
  for (int k = 0; k < run.n; ++k) st.s[k] = run.steps[k];
  const int64_t nif = pl.n_free - pl.n_int, ncon = pl.n_dofs - pl.n_free;
  const int64_t nthreads = std::max(nif, ncon);
  if (nthreads > 0)
    hipLaunchKernelGGL(cg_cheby_close_kernel<T>, dim3((unsigned)((nthreads + 255) / 256)), dim3(256), 0, s, nif, pg,
                       pl.n_free, pl.n_dofs, run.x, run.d, st);
  SMF_HIP(hipGetLastError());
  run.n = 0;
}

template void launch_cg_cheby_step<double>(const Operator&, ChebyRun<double>&, double, double, int, hipStream_t);
template void launch_cg_cheby_step<float>(const Operator&, ChebyRun<float>&, float, float, int, hipStream_t);
template void launch_cg_cheby_close<double>(const Operator&, ChebyRun<double>&, hipStream_t);
template void launch_cg_cheby_close<float>(const Operator&, ChebyRun<float>&, hipStream_t);

void launch_dg(const Operator& op, int mode, void* dst, const void* src, hipStream_t s);

void launch_op(const Operator& op, int mode, void* dst, const void* src, hipStream_t s) {
  if (op.is_dg()) launch_dg(op, mode, dst, src, s);
  else launch_cg(op, mode, dst, src, s);
}

int cg_quadrature_degree(int p) { return p == 1 ? 1 : (p == 2 ? 2 : 4); }

}  // namespace smf
