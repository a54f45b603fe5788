// Per-cell SIPG algorithm shared by the HIP kernel and its host execution.
//
// A cell K computes its cell integral and, for each local face f, the SIPG
// flux with the neighbour K' across f, testing only K's basis functions.
//   * Canonical faces: every cell lists its vertices in ascending global id,
//     so face f of K and face f' of K' are parametrised identically.
//   * Neighbour in "face 3" frame: K''s local nodes are renumbered so that
//     f' becomes local face 3 (vertices (kFV[f'] ascending, f')); then every
//     neighbour evaluation uses the same compile-time face-3 tables, and only
//     the DoF gather offsets and the barycentric normal components depend on
//     f' (small selects).
//   * Normal derivatives through the gradient field: grad u of a P_p function
//     lies in (P_{p-1})^3, so grad u . nJ on a face is the P_{p-1} face
//     interpolant of its values at the P_{p-1} face nodes (3 for P2).
// Reference: SIPG face flux of DG/cells_batched/application.cc:2104-2187
// (interior) and 2688-2701 (boundary), cell integral 1493-1735, penalty
// 1193-1245; components batched per DG/components_batched/application.cc;
// mass term of application_helmholtz.cc:1025-1111.
#pragma once
#include <cmath>
#include <cstdint>

#include "shape_tables.h"

#ifdef __HIPCC__
#define SMF_HD __host__ __device__
// the cell bodies are inlined into every kernel that calls them (a shared
// out-of-line copy costs the kernels occupancy)
#define SMF_HDI __host__ __device__ __attribute__((always_inline))
#else
#define SMF_HD
#define SMF_HDI
#endif

namespace smf {
namespace dgt {

using tab::alpha;
using tab::basis;
using tab::kFV;
using tab::nloc;

template <int P>
struct FaceRule;
template <>
struct FaceRule<1> {
  static constexpr int NQF = 3;
  static constexpr const double (*pts)[2] = q_tri_d2_pts;
  static constexpr const double* w = q_tri_d2_w;
};
template <>
struct FaceRule<2> {
  static constexpr int NQF = 6;
  static constexpr const double (*pts)[2] = q_tri_d4_pts;
  static constexpr const double* w = q_tri_d4_w;
};
template <>
struct FaceRule<3> {
  static constexpr int NQF = 12;
  static constexpr const double (*pts)[2] = q_tri_d6_pts;
  static constexpr const double* w = q_tri_d6_w;
};

// cell rule: stiffness only -> degree 2p-2; with a mass term -> degree 2p
template <int P, bool MASS>
struct DGCellRule {
  static constexpr int NQ = tab::CellRule<P>::NQ;
  static constexpr const double (*pts)[3] = tab::CellRule<P>::pts;
  static constexpr const double* w = tab::CellRule<P>::w;
};
template <>
struct DGCellRule<1, true> {
  static constexpr int NQ = 4;
  static constexpr const double (*pts)[3] = q_tet_d2_pts;
  static constexpr const double* w = q_tet_d2_w;
};
template <>
struct DGCellRule<2, true> {
  static constexpr int NQ = 11;
  static constexpr const double (*pts)[3] = q_tet_d4_pts;
  static constexpr const double* w = q_tet_d4_w;
};
template <>
struct DGCellRule<3, true> {
  static constexpr int NQ = 24;
  static constexpr const double (*pts)[3] = q_tet_d6_pts;
  static constexpr const double* w = q_tet_d6_w;
};

template <int P, bool MASS>
struct DGCellTab {
  static constexpr int NL = nloc(P), NQ = DGCellRule<P, MASS>::NQ;
  double g[NQ][NL][3];
  double v[NQ][NL];
  double w[NQ];
};

template <int P, bool MASS>
constexpr DGCellTab<P, MASS> make_cell_tab() {
  DGCellTab<P, MASS> t{};
  for (int q = 0; q < DGCellRule<P, MASS>::NQ; ++q) {
    t.w[q] = DGCellRule<P, MASS>::w[q];
    for (int i = 0; i < nloc(P); ++i) {
      t.v[q][i] = basis(P, i, DGCellRule<P, MASS>::pts[q], -1);
      for (int d = 0; d < 3; ++d) t.g[q][i][d] = basis(P, i, DGCellRule<P, MASS>::pts[q], d);
    }
  }
  return t;
}

// Face-3 mass matrices of the lane frame on the reference triangle (face
// weights, sum 1/2): Mvv[k][l] = sum_q w_q v_k v_l over the P_p face nodes,
// Mvg[k][j] = sum_q w_q v_k psi_j with the P_{p-1} gradient nodes.  The face
// is affine, so the SIPG face integrals are these matrices times the face
// Jacobian: exact, like the quadrature they replace.
template <int P>
struct DGFaceMass {
  static constexpr int NFN = (P + 1) * (P + 2) / 2, NFG = P == 1 ? 1 : P * (P + 1) / 2;
  double vv[NFN][NFN];
  double vg[NFN][NFG];
};

template <int P>
struct DGFaceTab {
  static constexpr int NL = nloc(P), NQF = FaceRule<P>::NQF;
  static constexpr int NFN = (P + 1) * (P + 2) / 2;          // P_p nodes on a face
  static constexpr int NFG = P == 1 ? 1 : P * (P + 1) / 2;   // P_{p-1} nodes on a face
  int fn[4][NFN];            // cell-local index of the face nodes of face f
  double v[4][NQF][NFN];     // their values at the face points
  double dg[4][NFG][NL][3];  // reference gradients of all basis functions at the P_{p-1} face nodes
  double psi[4][NQF][NFG];   // P_{p-1} basis of those nodes at the face points
  double w[NQF];             // face weights (sum 1/2)
  int perm[4][NL];           // neighbour frame: new node i <- old node perm[f'][i]
  int sigma[4][4];           // new vertex j <- old vertex sigma[f'][j]
};

constexpr void face_point(int f, double s, double t, double* x) {
  const double V[4][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
  const int a = kFV[f][0], b = kFV[f][1], c = kFV[f][2];
  for (int d = 0; d < 3; ++d) x[d] = V[a][d] + s * (V[b][d] - V[a][d]) + t * (V[c][d] - V[a][d]);
}

template <int P>
constexpr DGFaceTab<P> make_face_tab() {
  DGFaceTab<P> t{};
  constexpr int NL = nloc(P), NQF = FaceRule<P>::NQF;
  for (int q = 0; q < NQF; ++q) t.w[q] = FaceRule<P>::w[q];
  for (int f = 0; f < 4; ++f) {
    int n = 0;
    for (int i = 0; i < NL; ++i)
      if (alpha(P, i).a[f] == 0) t.fn[f][n++] = i;
    for (int q = 0; q < NQF; ++q) {
      double x[3] = {0, 0, 0};
      face_point(f, FaceRule<P>::pts[q][0], FaceRule<P>::pts[q][1], x);
      for (int k = 0; k < DGFaceTab<P>::NFN; ++k) t.v[f][q][k] = basis(P, t.fn[f][k], x, -1);
      if (P == 1) {
        t.psi[f][q][0] = 1.0;
      } else {
        int m = 0;
        for (int j = 0; j < nloc(P - 1); ++j)
          if (alpha(P - 1, j).a[f] == 0) t.psi[f][q][m++] = basis(P - 1, j, x, -1);
      }
    }
    if (P == 1) {
      const double x0[3] = {0.25, 0.25, 0.25};
      for (int i = 0; i < NL; ++i)
        for (int d = 0; d < 3; ++d) t.dg[f][0][i][d] = basis(P, i, x0, d);
    } else {
      int m = 0;
      for (int j = 0; j < nloc(P - 1); ++j) {
        const tab::Alpha al = alpha(P - 1, j);
        if (al.a[f] != 0) continue;
        const double V[4][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
        double y[3] = {0, 0, 0};
        for (int kk = 0; kk < 4; ++kk)
          for (int d = 0; d < 3; ++d) y[d] += al.a[kk] * V[kk][d] / (P - 1);
        for (int i = 0; i < NL; ++i)
          for (int d = 0; d < 3; ++d) t.dg[f][m][i][d] = basis(P, i, y, d);
        ++m;
      }
    }
    // neighbour frame for f' = f: new vertices (kFV[f] ascending, f)
    const int sg[4] = {kFV[f][0], kFV[f][1], kFV[f][2], f};
    for (int j = 0; j < 4; ++j) t.sigma[f][j] = sg[j];
    for (int i = 0; i < NL; ++i) {
      const tab::Alpha an = alpha(P, i);
      int ao[4] = {0, 0, 0, 0};
      for (int j = 0; j < 4; ++j) ao[sg[j]] = an.a[j];
      t.perm[f][i] = -1;
      for (int o = 0; o < NL; ++o) {
        const tab::Alpha b = alpha(P, o);
        if (b.a[0] == ao[0] && b.a[1] == ao[1] && b.a[2] == ao[2] && b.a[3] == ao[3]) t.perm[f][i] = o;
      }
    }
  }
  return t;
}

// Every face lists its nodes in the same pattern (face vertices ascending,
// then edge nodes) and is parametrised from its lowest vertex, so the face
// node values and gradient-node functions at the face points are those of
// face 3: one set of face mass matrices serves all four faces.
template <int P>
constexpr bool faces_share_tables() {
  constexpr DGFaceTab<P> FT = make_face_tab<P>();
  for (int f = 0; f < 3; ++f)
    for (int q = 0; q < DGFaceTab<P>::NQF; ++q) {
      for (int k = 0; k < DGFaceTab<P>::NFN; ++k)
        if (FT.v[f][q][k] != FT.v[3][q][k]) return false;
      for (int j = 0; j < DGFaceTab<P>::NFG; ++j)
        if (FT.psi[f][q][j] != FT.psi[3][q][j]) return false;
    }
  return true;
}

template <int P>
constexpr DGFaceMass<P> make_face_mass() {
  constexpr DGFaceTab<P> FT = make_face_tab<P>();
  DGFaceMass<P> m{};
  for (int k = 0; k < DGFaceMass<P>::NFN; ++k) {
    for (int l = 0; l < DGFaceMass<P>::NFN; ++l) {
      double a = 0;
      for (int q = 0; q < DGFaceTab<P>::NQF; ++q) a += FT.w[q] * FT.v[3][q][k] * FT.v[3][q][l];
      m.vv[k][l] = a;
    }
    for (int g = 0; g < DGFaceMass<P>::NFG; ++g) {
      double a = 0;
      for (int q = 0; q < DGFaceTab<P>::NQF; ++q) a += FT.w[q] * FT.v[3][q][k] * FT.psi[3][q][g];
      m.vg[k][g] = a;
    }
  }
  return m;
}

// acc += c x for an entry c of a compile-time table (c folds to a constant
// after unrolling): zero entries vanish and +-1 entries add or subtract.
// x * 0.0 may not be dropped under IEEE rules (x could be Inf or NaN), so
// written as a plain product every zero of a table would cost an FMA (the
// P3 face-gradient table has 118 nonzeros of 360).
template <typename T>
SMF_HDI void tmac(T& acc, double c, T x) {
  if (c == 0.0) return;
  if (c == 1.0) acc += x;
  else if (c == -1.0) acc -= x;
  else acc += T(c) * x;
}

// outward reference normals of the faces (face f opposite vertex f), unnormalised
constexpr double kNhat[4][3] = {{1, 1, 1}, {-1, 0, 0}, {0, -1, 0}, {0, 0, -1}};

// a balanced tree of two-way selects on the bits of f: three v_cndmask per
// 32-bit half, no branches (f differs between the lanes of a wave)
template <typename T>
SMF_HD inline T select4(int f, T a0, T a1, T a2, T a3) {
  const bool odd = (f & 1) != 0;
  const T lo = odd ? a1 : a0, hi = odd ? a3 : a2;
  return (f & 2) ? hi : lo;
}

// One cell k: MODE 0 dst rows of k = (A src) rows; MODE 1 = diagonal of A.
// geo: SoA [11][ncells] = J^-T (row-major), |det J|, tau
template <int P, typename T, int NC, bool MASS, int MODE>
SMF_HDI void apply_cell(int64_t k, int64_t ncells, const T* __restrict__ geo, const int32_t* __restrict__ nbr,
                       const uint32_t* __restrict__ nface, T factor, T mass, T lap, const T* __restrict__ src,
                       T* __restrict__ dst, int pkind = 0) {
  constexpr DGCellTab<P, MASS> CT = make_cell_tab<P, MASS>();
  constexpr DGFaceTab<P> FT = make_face_tab<P>();
  constexpr int NL = nloc(P), NQ = DGCellTab<P, MASS>::NQ, NQF = DGFaceTab<P>::NQF;
  constexpr int NFN = DGFaceTab<P>::NFN, NFG = DGFaceTab<P>::NFG;
  using std::sqrt;

  T jit[9];
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
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      const T w = T(CT.w[q]);
      if (MODE == 0) {
        T gx = T(0), gy = T(0), gz = T(0), uv = T(0);
#pragma unroll
        for (int i = 0; i < NL; ++i) {
          tmac(gx, CT.g[q][i][0], u[i]);
          tmac(gy, CT.g[q][i][1], u[i]);
          tmac(gz, CT.g[q][i][2], u[i]);
          if (MASS) tmac(uv, CT.v[q][i], u[i]);
        }
        const T wl = w * lap;
        const T tx = wl * (M0 * gx + M1 * gy + M2 * gz);
        const T ty = wl * (M1 * gx + M3 * gy + M4 * gz);
        const T tz = wl * (M2 * gx + M4 * gy + M5 * gz);
        const T tv = MASS ? w * mass * adet * uv : T(0);
#pragma unroll
        for (int i = 0; i < NL; ++i) {
          tmac(y[i], CT.g[q][i][0], tx);
          tmac(y[i], CT.g[q][i][1], ty);
          tmac(y[i], CT.g[q][i][2], tz);
          if (MASS) tmac(y[i], CT.v[q][i], tv);
        }
      } else {
#pragma unroll
        for (int i = 0; i < NL; ++i) {
          const T gx = T(CT.g[q][i][0]), gy = T(CT.g[q][i][1]), gz = T(CT.g[q][i][2]);
          y[i] += w * lap * (gx * (M0 * gx + M1 * gy + M2 * gz) + gy * (M1 * gx + M3 * gy + M4 * gz) +
                             gz * (M2 * gx + M4 * gy + M5 * gz));
          if (MASS) y[i] += w * mass * adet * T(CT.v[q][i]) * T(CT.v[q][i]);
        }
      }
    }
    // ---- face integrals ----
#pragma unroll
    for (int f = 0; f < 4; ++f) {
#if defined(__HIP_DEVICE_COMPILE__)
      // keep each face's neighbour loads and temporaries inside its own
      // iteration (bounds register pressure)
      __builtin_amdgcn_sched_barrier(0);
#endif
      const T nj0 = jit[0] * T(kNhat[f][0]) + jit[1] * T(kNhat[f][1]) + jit[2] * T(kNhat[f][2]);
      const T nj1 = jit[3] * T(kNhat[f][0]) + jit[4] * T(kNhat[f][1]) + jit[5] * T(kNhat[f][2]);
      const T nj2 = jit[6] * T(kNhat[f][0]) + jit[7] * T(kNhat[f][1]) + jit[8] * T(kNhat[f][2]);
      const T len = sqrt(nj0 * nj0 + nj1 * nj1 + nj2 * nj2);
      const T il = T(1) / len;
      const T n0 = nj0 * il, n1 = nj1 * il, n2 = nj2 * il;
      const T area2 = adet * len * lap;
      // nJ = J^-1 n: d/dn of a function = (reference gradient) . nJ
      const T nJ[3] = {jit[0] * n0 + jit[3] * n1 + jit[6] * n2, jit[1] * n0 + jit[4] * n1 + jit[7] * n2,
                       jit[2] * n0 + jit[5] * n1 + jit[8] * n2};
      const int32_t kb = nbr[f * ncells + k];
      const bool interior = kb >= 0;
      const int64_t kbs = interior ? kb : k;
      T sigma;
      if (pkind == 1) {  // (p+1)(p+2) |F| / min |K|: factor * max(len, len |det J| / |det J+|)
        const T lb = len * adet / geo[9 * ncells + kbs];
        sigma = factor * (len > lb ? len : lb);
      } else {
        const T taub = geo[10 * ncells + kbs];
        sigma = factor * (interior ? (tau > taub ? tau : taub) : tau);
      }

      if (MODE == 0) {
        // own normal derivative at the P_{p-1} face nodes
        T gn[NFG];
#pragma unroll
        for (int j = 0; j < NFG; ++j) {
          T sx = T(0), sy = T(0), sz = T(0);
#pragma unroll
          for (int i = 0; i < NL; ++i) {
            tmac(sx, FT.dg[f][j][i][0], u[i]);
            tmac(sy, FT.dg[f][j][i][1], u[i]);
            tmac(sz, FT.dg[f][j][i][2], u[i]);
          }
          gn[j] = sx * nJ[0] + sy * nJ[1] + sz * nJ[2];
        }
        // neighbour, renumbered so that the shared face is its face 3
        T ub[NL], gb[NFG];
#pragma unroll
        for (int i = 0; i < NL; ++i) ub[i] = T(0);
#pragma unroll
        for (int j = 0; j < NFG; ++j) gb[j] = T(0);
        if (interior) {
          const int fb = (nf >> (8 * f)) & 3;
          const int64_t nbase = ((int64_t)kb * NC + c) * NL;
#pragma unroll
          for (int i = 0; i < NL; ++i)
            ub[i] = src[nbase + select4<int>(fb, FT.perm[0][i], FT.perm[1][i], FT.perm[2][i], FT.perm[3][i])];
          T jb[9];
#pragma unroll
          for (int t = 0; t < 9; ++t) jb[t] = geo[t * ncells + kb];
          // barycentric normal components beta_a = grad lambda_a . n of the neighbour
          const T b1 = jb[0] * n0 + jb[3] * n1 + jb[6] * n2;
          const T b2 = jb[1] * n0 + jb[4] * n1 + jb[7] * n2;
          const T b3 = jb[2] * n0 + jb[5] * n1 + jb[8] * n2;
          const T b0 = -(b1 + b2 + b3);
          const T bo[4] = {b0, b1, b2, b3};
          T nb[3];
#pragma unroll
          for (int d = 0; d < 3; ++d)
            nb[d] = select4<T>(fb, bo[FT.sigma[0][d + 1]], bo[FT.sigma[1][d + 1]], bo[FT.sigma[2][d + 1]],
                               bo[FT.sigma[3][d + 1]]);
#pragma unroll
          for (int j = 0; j < NFG; ++j) {
            T sx = T(0), sy = T(0), sz = T(0);
#pragma unroll
            for (int i = 0; i < NL; ++i) {
              tmac(sx, FT.dg[3][j][i][0], ub[i]);
              tmac(sy, FT.dg[3][j][i][1], ub[i]);
              tmac(sz, FT.dg[3][j][i][2], ub[i]);
            }
            gb[j] = sx * nb[0] + sy * nb[1] + sz * nb[2];
          }
        }
        // face terms in matrix form (as in apply_cell_lane): own face nodes
        // fn[f], the neighbour's fn[3] (same physical nodes, same order)
        static_assert(faces_share_tables<P>(), "face tables differ between faces");
        constexpr DGFaceMass<P> FM = make_face_mass<P>();
        T J[NFN], A[NFG];
#pragma unroll
        for (int k = 0; k < NFN; ++k)
          J[k] = interior ? u[FT.fn[f][k]] - ub[FT.fn[3][k]] : T(2) * u[FT.fn[f][k]];
#pragma unroll
        for (int j = 0; j < NFG; ++j) A[j] = interior ? T(0.5) * (gn[j] + gb[j]) : gn[j];
#pragma unroll
        for (int k = 0; k < NFN; ++k) {
          T sj = T(0), sa = T(0);
#pragma unroll
          for (int l = 0; l < NFN; ++l) tmac(sj, FM.vv[k][l], J[l]);
#pragma unroll
          for (int j = 0; j < NFG; ++j) tmac(sa, FM.vg[k][j], A[j]);
          y[FT.fn[f][k]] += area2 * (sigma * sj - sa);
        }
        T tgj[NFG];
#pragma unroll
        for (int j = 0; j < NFG; ++j) {
          T sj = T(0);
#pragma unroll
          for (int k = 0; k < NFN; ++k) tmac(sj, FM.vg[k][j], J[k]);
          tgj[j] = T(-0.5) * area2 * sj;
        }
        // gradient test: y_i += sum_j tgj[j] * grad phi_i(y_j) . nJ
#pragma unroll
        for (int j = 0; j < NFG; ++j) {
          const T ax = tgj[j] * nJ[0], ay = tgj[j] * nJ[1], az = tgj[j] * nJ[2];
#pragma unroll
          for (int i = 0; i < NL; ++i) {
            tmac(y[i], FT.dg[f][j][i][0], ax);
            tmac(y[i], FT.dg[f][j][i][1], ay);
            tmac(y[i], FT.dg[f][j][i][2], az);
          }
        }
      } else {
        // diagonal: u = phi_i on this cell only (u+ = 0 inside, mirror on the boundary)
        const T s = interior ? T(1) : T(2);
        T dnode[NFG][NL];
#pragma unroll
        for (int j = 0; j < NFG; ++j)
#pragma unroll
          for (int i = 0; i < NL; ++i)
            dnode[j][i] = T(FT.dg[f][j][i][0]) * nJ[0] + T(FT.dg[f][j][i][1]) * nJ[1] + T(FT.dg[f][j][i][2]) * nJ[2];
#pragma unroll
        for (int q = 0; q < NQF; ++q) {
          const T jxw = T(FT.w[q]) * area2;
#pragma unroll
          for (int i = 0; i < NL; ++i) {
            T vi = T(0);
#pragma unroll
            for (int kk = 0; kk < NFN; ++kk)
              if (FT.fn[f][kk] == i) vi = T(FT.v[f][q][kk]);
            T di = T(0);
#pragma unroll
            for (int j = 0; j < NFG; ++j) di += T(FT.psi[f][q][j]) * dnode[j][i];
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

// Share of cell-rule point q for one lane of the four-lane-per-cell form:
// the lane works in the frame where its face is face 3, and takes the points
// whose largest barycentric coordinate belongs to vertex 3 (ties split
// evenly).  The rules are symmetric under vertex permutations, so over the
// four lanes every point of the rule counts exactly once.
template <int P, bool MASS>
struct DGCellShare {
  double s[DGCellRule<P, MASS>::NQ];
};
template <int P, bool MASS>
constexpr DGCellShare<P, MASS> make_cell_share() {
  DGCellShare<P, MASS> r{};
  for (int q = 0; q < DGCellRule<P, MASS>::NQ; ++q) {
    const double* x = DGCellRule<P, MASS>::pts[q];
    const double l[4] = {1.0 - x[0] - x[1] - x[2], x[0], x[1], x[2]};
    double m = l[0];
    for (int a = 1; a < 4; ++a) m = l[a] > m ? l[a] : m;
    int cnt = 0;
    for (int a = 0; a < 4; ++a) cnt += (m - l[a] < 1e-12) ? 1 : 0;
    r.s[q] = (m - l[3] < 1e-12) ? 1.0 / cnt : 0.0;
  }
  return r;
}

// Slot of cell node i in lane g's row (lane g stores the row of its frame
// node j, cell node perm[g][j], at j): the inverse permutations.
template <int P>
struct LaneInv {
  int j[4][nloc(P)];
};
template <int P>
constexpr LaneInv<P> make_lane_inv() {
  constexpr DGFaceTab<P> FT = make_face_tab<P>();
  LaneInv<P> r{};
  for (int g = 0; g < 4; ++g)
    for (int j = 0; j < nloc(P); ++j) r.j[g][FT.perm[g][j]] = j;
  return r;
}
// slot of cell node f + 4 t in lane g's row for a run-time f in 0..3 and
// unrolled (compile-time) g, t: three selects between constants
template <int P>
SMF_HDI int lane_slot4(int g, int t, int f) {
  constexpr LaneInv<P> LI = make_lane_inv<P>();
  constexpr int NL = nloc(P);
  const int i0 = 4 * t;
  return select4<int>(f, LI.j[g][i0 < NL ? i0 : 0], LI.j[g][i0 + 1 < NL ? i0 + 1 : 0],
                      LI.j[g][i0 + 2 < NL ? i0 + 2 : 0], LI.j[g][i0 + 3 < NL ? i0 + 3 : 0]);
}

// Geometry of lane f of cell k (shared by all components of the cell):
// the face normal and its J^-1 n in the lane frame (nJ), the cell metric in
// the lane frame, the face's |F|-scaled Jacobian factor, and the neighbour
// (kn, its local face fb, its geometry).
// The neighbour's values (jb, taub, adetb) are first used in the face
// section, so their loads stay in flight across the cell term.
template <typename T>
struct LaneGeo {
  T nJ[3], M[6], area2, adet;
  T nb[3], sigma;  // the lane's face record
  int64_t kn;
  int fb;
  bool interior;
};

// Per-face record of the four-lane kernel, [cell][local face][4], built once
// at set-up from the geometry of both cells of the face: the neighbour's
// J^-1 n (its barycentric gradients in the frame of its face fb, dotted with
// this cell's outward unit normal n; zero on the boundary) and the face's
// penalty sigma (penalty factor included).  Lane f of cell k reads the 4
// values at (4 k + f) 4: one contiguous 32 B (FP64) read per lane instead of
// eleven gathers of the neighbour's geometry arrays.
constexpr int kFaceRec = 4;
// Neighbour word of the four-lane kernel, [cell][local face] (one coalesced
// 4 B read per lane): -1 on the boundary, else the neighbour's cell index in
// the low kNbrShift bits and its local face fb above them.
constexpr int kNbrShift = 29;
constexpr int32_t kNbrMask = (1 << kNbrShift) - 1;

// The lane-frame node permutations perm[f] packed into 32-bit words, B bits
// per entry (4 up to 16 nodes, 5 up to 32), so that a run-time f costs one
// select per word and one bit-field extract per node instead of a select per
// node.
template <int P>
struct PermWords {
  static constexpr int B = nloc(P) <= 16 ? 4 : 5, E = 32 / B, W = (nloc(P) + E - 1) / E;
  uint32_t w[4][W];
};
template <int P>
constexpr PermWords<P> make_perm_words() {
  constexpr DGFaceTab<P> FT = make_face_tab<P>();
  PermWords<P> r{};
  for (int f = 0; f < 4; ++f)
    for (int i = 0; i < nloc(P); ++i)
      r.w[f][i / PermWords<P>::E] |= (uint32_t)FT.perm[f][i] << (PermWords<P>::B * (i % PermWords<P>::E));
  return r;
}
// perm[f][i] for i = 0..NL-1 into out[]
template <int P>
SMF_HDI void lane_perm(int f, int* out) {
  constexpr PermWords<P> PW = make_perm_words<P>();
  constexpr int B = PermWords<P>::B, E = PermWords<P>::E, W = PermWords<P>::W;
  uint32_t w[W];
#pragma unroll
  for (int t = 0; t < W; ++t) w[t] = select4<uint32_t>(f, PW.w[0][t], PW.w[1][t], PW.w[2][t], PW.w[3][t]);
#pragma unroll
  for (int i = 0; i < nloc(P); ++i) out[i] = (int)((w[i / E] >> (B * (i % E))) & ((1u << B) - 1));
}

// Barycentric metric of a cell, B_ab = |det J| grad lambda_a . grad lambda_b
// (a, b = 0..3): symmetric, rows summing to zero, so the six off-diagonal
// entries determine it.  kBPair[a][b] is the slot of (a, b) in the cell's
// [7][ncells] record (b01 b02 b03 b12 b13 b23 |det J|).
constexpr int kBPair[4][4] = {{-1, 0, 1, 2}, {0, -1, 3, 4}, {1, 3, -1, 5}, {2, 4, 5, -1}};
constexpr int kBGeo = 7;
// vertex of frame role r of face f: 0 origin kFV[f][0], 1 and 2 the other
// face vertices kFV[f][1], kFV[f][2], 3 the opposite vertex f
constexpr int frame_vertex(int f, int r) { return r == 3 ? f : tab::kFV[f][r]; }
template <int RX, int RY, typename T>
SMF_HDI T bsel(int f, const T* b) {
  return select4<T>(f, b[kBPair[frame_vertex(0, RX)][frame_vertex(0, RY)]],
                    b[kBPair[frame_vertex(1, RX)][frame_vertex(1, RY)]],
                    b[kBPair[frame_vertex(2, RX)][frame_vertex(2, RY)]],
                    b[kBPair[frame_vertex(3, RX)][frame_vertex(3, RY)]]);
}

// Geometry of lane f of cell k (shared by all components of the cell).  The
// lane frame has origin vertex o = kFV[f][0] and reference directions
// (a, b, f) = (kFV[f][1], kFV[f][2], f).  Its metric is B restricted to
// {a, b, f} (diagonal from the zero row sums), and because face f is
// opposite vertex f, the outward unit normal is n = -grad lambda_f / |grad
// lambda_f|: J^-1 n in the frame is -B[., f] / s and the face Jacobian
// factor |det J| |grad lambda_f| is s, with s = sqrt(|det J| B_ff).  The
// neighbour's J^-1 n and the penalty come from the face record.
// NT: non-temporal loads (data read once per product, so that the caches
// keep the gathered cell values instead)
template <bool NT, typename T>
SMF_HDI T load_once(const T* p) {
  if constexpr (NT) return __builtin_nontemporal_load(p);
  else return *p;
}
template <int P, typename T, bool NT = false>
SMF_HDI LaneGeo<T> lane_geometry(int64_t k, int f, int64_t ncells, const T* __restrict__ bgeo,
                                 const T* __restrict__ fgeo, const int32_t* __restrict__ nbrf, T lap) {
  using std::sqrt;
  LaneGeo<T> g;
  T b[6];
#pragma unroll
  for (int t = 0; t < 6; ++t) b[t] = load_once<NT>(bgeo + t * ncells + k);
  const T adet = load_once<NT>(bgeo + 6 * ncells + k);
  g.adet = adet;
  const int32_t kb = load_once<NT>(nbrf + k * 4 + f);
  g.interior = kb >= 0;
  g.kn = g.interior ? (kb & kNbrMask) : k;
  g.fb = (kb >> kNbrShift) & 3;  // unused on the boundary
  const T* fr = fgeo + (k * 4 + f) * kFaceRec;
  g.nb[0] = load_once<NT>(fr);
  g.nb[1] = load_once<NT>(fr + 1);
  g.nb[2] = load_once<NT>(fr + 2);
  g.sigma = load_once<NT>(fr + 3);
  // B[x][y] for the frame roles x, y of lane f: one select over the four faces
  const T bab = bsel<1, 2>(f, b), baf = bsel<1, 3>(f, b), bbf = bsel<2, 3>(f, b);
  const T bao = bsel<1, 0>(f, b), bbo = bsel<2, 0>(f, b), bfo = bsel<3, 0>(f, b);
  const T bff = -((bfo + baf) + bbf);
  g.M[0] = -((bao + bab) + baf);
  g.M[1] = bab;
  g.M[2] = baf;
  g.M[3] = -((bbo + bab) + bbf);
  g.M[4] = bbf;
  g.M[5] = bff;
  const T sf = sqrt(adet * bff);
  const T is = T(-1) / sf;
  g.nJ[0] = baf * is;
  g.nJ[1] = bbf * is;
  g.nJ[2] = bff * is;
  g.area2 = sf * lap;
  return g;
}

// Lane f of cell k, component c, from the lane's geometry: the cell term's
// share (above) and the SIPG flux of face f, evaluated with the cell's nodes
// renumbered so that face f is face 3 (the neighbour's frame of apply_cell),
// so every table is the compile-time face-3 table.  yrow[j] receives the
// lane's contribution to the row of its frame node j (cell node perm[f][j];
// make_lane_inv inverts it) (MODE 0) or to that diagonal entry (MODE 1); the
// four lanes' rows add up to apply_cell's result.  Storing in the lane frame
// keeps the permutation out of the registers live across the lane's work.
// Reference mass matrix of the degree-p Lagrange basis (exact: the cell rule
// of degree 2p) and each node's share for the four lanes: lane f computes the
// mass rows of its frame nodes in proportion alpha_3 / p, the node's
// barycentric index on the frame vertex that is cell vertex f, so the four
// lanes' shares of every row sum to one.  The mass matrix of an affine cell is
// |det J| times the reference one, and the reference one is the same in every
// lane frame (its entries depend only on the nodes' barycentric indices).
template <int P>
struct DGMassTab {
  double m[nloc(P)][nloc(P)];
  double share[nloc(P)];
};
template <int P>
constexpr DGMassTab<P> make_mass_tab() {
  DGMassTab<P> t{};
  constexpr DGCellTab<P, true> CT = make_cell_tab<P, true>();
  for (int i = 0; i < nloc(P); ++i) {
    t.share[i] = (double)alpha(P, i).a[3] / P;
    for (int j = 0; j < nloc(P); ++j) {
      double a = 0;
      for (int q = 0; q < DGCellTab<P, true>::NQ; ++q) a += CT.w[q] * CT.v[q][i] * CT.v[q][j];
      t.m[i][j] = a;
    }
  }
  return t;
}

// ST: the values of the workgroup's cells (component c) are staged in LDS,
// sblk[(k - k0) NL + i] for k0 <= k < k0 + ncb; the own values and the
// neighbour's, when the neighbour is one of those cells, are read there.
template <int P, typename T, int NC, bool MASS, int MODE, bool ST = false>
SMF_HDI void lane_component(const LaneGeo<T>& g, int64_t k, int f, int c, T mass, T lap,
                            const T* __restrict__ src, T* __restrict__ yrow, const T* sblk = nullptr,
                            int64_t k0 = 0, int ncb = 0) {
  // the stiffness at the stiffness rule; a mass term through the reference
  // mass matrix (below)
  constexpr DGCellTab<P, false> CT = make_cell_tab<P, false>();
  constexpr DGCellShare<P, false> CS = make_cell_share<P, false>();
  constexpr DGFaceTab<P> FT = make_face_tab<P>();
  constexpr int NL = nloc(P), NQ = DGCellTab<P, false>::NQ, NQF = DGFaceTab<P>::NQF;
  constexpr int NFN = DGFaceTab<P>::NFN, NFG = DGFaceTab<P>::NFG;
  const T M0 = g.M[0], M1 = g.M[1], M2 = g.M[2], M3 = g.M[3], M4 = g.M[4], M5 = g.M[5];
  const T* nJ = g.nJ;
  const T adet = g.adet, area2 = g.area2;
  const bool interior = g.interior;
  int pf[NL];
  lane_perm<P>(f, pf);

  const int64_t base = (k * NC + c) * NL;
  T u[NL], y[NL];
#pragma unroll
  for (int i = 0; i < NL; ++i) {
    u[i] = MODE == 0 ? (ST ? sblk[(k - k0) * NL + pf[i]] : src[base + pf[i]]) : T(0);
    y[i] = T(0);
  }
  // neighbour values (a boundary face reads its own cell; those values are
  // not used): up to P2 issued before the cell term so that their latency
  // overlaps it; from P3 on after the cell term and the own normal
  // derivatives, where three live arrays of 20 values would cost a wave per
  // SIMD
  T ub[NL];
  auto load_ub = [&]() {
    const int64_t nbase = (g.kn * NC + c) * NL;
    int pb[NL];
    lane_perm<P>(g.fb, pb);
    if (ST && (uint64_t)(g.kn - k0) < (uint64_t)ncb) {
      const T* row = sblk + (g.kn - k0) * NL;
#pragma unroll
      for (int i = 0; i < NL; ++i) ub[i] = row[pb[i]];
    } else {
#pragma unroll
      for (int i = 0; i < NL; ++i) ub[i] = src[nbase + pb[i]];
    }
  };
  if (MODE == 0 && P <= 2) load_ub();
  // ---- the lane's share of the cell integral ----
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    if (CS.s[q] == 0.0) continue;
    const T w = T(CT.w[q] * CS.s[q]);
    if (MODE == 0) {
      T gx = T(0), gy = T(0), gz = T(0);
#pragma unroll
      for (int i = 0; i < NL; ++i) {
        tmac(gx, CT.g[q][i][0], u[i]);
        tmac(gy, CT.g[q][i][1], u[i]);
        tmac(gz, CT.g[q][i][2], u[i]);
      }
      const T wl = w * lap;
      const T tx = wl * (M0 * gx + M1 * gy + M2 * gz);
      const T ty = wl * (M1 * gx + M3 * gy + M4 * gz);
      const T tz = wl * (M2 * gx + M4 * gy + M5 * gz);
#pragma unroll
      for (int i = 0; i < NL; ++i) {
        tmac(y[i], CT.g[q][i][0], tx);
        tmac(y[i], CT.g[q][i][1], ty);
        tmac(y[i], CT.g[q][i][2], tz);
      }
    } else {
#pragma unroll
      for (int i = 0; i < NL; ++i) {
        const T gx = T(CT.g[q][i][0]), gy = T(CT.g[q][i][1]), gz = T(CT.g[q][i][2]);
        y[i] += w * lap * (gx * (M0 * gx + M1 * gy + M2 * gz) + gy * (M1 * gx + M3 * gy + M4 * gz) +
                           gz * (M2 * gx + M4 * gy + M5 * gz));
      }
    }
  }
  if (MASS) {
    // this lane's share of |det J| mass (M_ref u)
    constexpr DGMassTab<P> MT = make_mass_tab<P>();
    const T ma = mass * adet;
#pragma unroll
    for (int i = 0; i < NL; ++i) {
      if (MT.share[i] == 0.0) continue;
      T acc = T(0);
      if (MODE == 0) {
#pragma unroll
        for (int j = 0; j < NL; ++j) tmac(acc, MT.m[i][j], u[j]);
      } else {
        acc = T(MT.m[i][i]);
      }
      y[i] += (T(MT.share[i]) * ma) * acc;
    }
  }
  // ---- face f (= face 3 of the lane frame) ----
  const T sigma = g.sigma;
  if (MODE == 0) {
    T gn[NFG];
#pragma unroll
    for (int j = 0; j < NFG; ++j) {
      T sx = T(0), sy = T(0), sz = T(0);
#pragma unroll
      for (int i = 0; i < NL; ++i) {
        tmac(sx, FT.dg[3][j][i][0], u[i]);
        tmac(sy, FT.dg[3][j][i][1], u[i]);
        tmac(sz, FT.dg[3][j][i][2], u[i]);
      }
      gn[j] = sx * nJ[0] + sy * nJ[1] + sz * nJ[2];
    }
    // own values at the face nodes (the rest of u is no longer needed)
    T uf[NFN];
#pragma unroll
    for (int kk = 0; kk < NFN; ++kk) uf[kk] = u[FT.fn[3][kk]];
    if (P > 2) {
#ifdef __HIP_DEVICE_COMPILE__
      __builtin_amdgcn_sched_barrier(0);  // keep the loads after the cell term and gn
#endif
      load_ub();
    }
    T gb[NFG];
#pragma unroll
    for (int j = 0; j < NFG; ++j) gb[j] = T(0);
    if (interior) {
      // the neighbour's J^-1 n in the frame of its face fb (face record)
      const T* nb = g.nb;
#pragma unroll
      for (int j = 0; j < NFG; ++j) {
        T sx = T(0), sy = T(0), sz = T(0);
#pragma unroll
        for (int i = 0; i < NL; ++i) {
          tmac(sx, FT.dg[3][j][i][0], ub[i]);
          tmac(sy, FT.dg[3][j][i][1], ub[i]);
          tmac(sz, FT.dg[3][j][i][2], ub[i]);
        }
        gb[j] = sx * nb[0] + sy * nb[1] + sz * nb[2];
      }
    }
    // SIPG face terms in matrix form: jump J of the values at the face nodes
    // (mirror u+ = -u- on the boundary), mean A of the normal derivatives at
    // the gradient nodes (d_n u+ = d_n u- on the boundary); then
    //   y_face += |F| (sigma Mvv J - Mvg A),  tgj = -1/2 |F| Mvg^T J
    constexpr DGFaceMass<P> FM = make_face_mass<P>();
    T J[NFN], A[NFG];
#pragma unroll
    for (int kk = 0; kk < NFN; ++kk) J[kk] = interior ? uf[kk] - ub[FT.fn[3][kk]] : T(2) * uf[kk];
#pragma unroll
    for (int j = 0; j < NFG; ++j) A[j] = interior ? T(0.5) * (gn[j] + gb[j]) : gn[j];
#pragma unroll
    for (int kk = 0; kk < NFN; ++kk) {
      T sj = T(0), sa = T(0);
#pragma unroll
      for (int l = 0; l < NFN; ++l) tmac(sj, FM.vv[kk][l], J[l]);
#pragma unroll
      for (int j = 0; j < NFG; ++j) tmac(sa, FM.vg[kk][j], A[j]);
      y[FT.fn[3][kk]] += area2 * (sigma * sj - sa);
    }
    T tgj[NFG];
#pragma unroll
    for (int j = 0; j < NFG; ++j) {
      T sj = T(0);
#pragma unroll
      for (int kk = 0; kk < NFN; ++kk) tmac(sj, FM.vg[kk][j], J[kk]);
      tgj[j] = T(-0.5) * area2 * sj;
    }
#pragma unroll
    for (int j = 0; j < NFG; ++j) {
      const T ax = tgj[j] * nJ[0], ay = tgj[j] * nJ[1], az = tgj[j] * nJ[2];
#pragma unroll
      for (int i = 0; i < NL; ++i) {
        tmac(y[i], FT.dg[3][j][i][0], ax);
        tmac(y[i], FT.dg[3][j][i][1], ay);
        tmac(y[i], FT.dg[3][j][i][2], az);
      }
    }
  } else {
    const T s = interior ? T(1) : T(2);
    T dnode[NFG][NL];
#pragma unroll
    for (int j = 0; j < NFG; ++j)
#pragma unroll
      for (int i = 0; i < NL; ++i)
        dnode[j][i] = T(FT.dg[3][j][i][0]) * nJ[0] + T(FT.dg[3][j][i][1]) * nJ[1] + T(FT.dg[3][j][i][2]) * nJ[2];
#pragma unroll
    for (int q = 0; q < NQF; ++q) {
      const T jxw = T(FT.w[q]) * area2;
#pragma unroll
      for (int i = 0; i < NL; ++i) {
        T vi = T(0);
#pragma unroll
        for (int kk = 0; kk < NFN; ++kk)
          if (FT.fn[3][kk] == i) vi = T(FT.v[3][q][kk]);
        T di = T(0);
#pragma unroll
        for (int j = 0; j < NFG; ++j) di += T(FT.psi[3][q][j]) * dnode[j][i];
        const T jump = s * vi, avg = T(0.5) * s * di;
        y[i] += jxw * ((sigma * jump - avg) * vi - T(0.5) * jump * di);
      }
    }
  }
#pragma unroll
  for (int i = 0; i < NL; ++i) yrow[i] = y[i];
}

// Lane f of cell k, component c (geometry and component together)
template <int P, typename T, int NC, bool MASS, int MODE>
SMF_HDI void apply_cell_lane(int64_t k, int f, int c, int64_t ncells, const T* __restrict__ bgeo,
                             const T* __restrict__ fgeo, const int32_t* __restrict__ nbrf, T mass, T lap,
                             const T* __restrict__ src,
                             T* __restrict__ yrow) {
  const LaneGeo<T> g = lane_geometry<P, T>(k, f, ncells, bgeo, fgeo, nbrf, lap);
  lane_component<P, T, NC, MASS, MODE>(g, k, f, c, mass, lap, src, yrow);
}

}  // namespace dgt
}  // namespace smf
