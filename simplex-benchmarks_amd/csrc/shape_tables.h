// Compile-time reference-element tables for the HIP kernels.
//
// The Lagrange basis (closed barycentric product form, node order of fe.h)
// and the symmetric quadrature rules of quad_tables.h are evaluated by
// constexpr code, so each kernel sees its tables as literal constants: after
// unrolling, the compiler drops zero entries, turns +-1 into adds and keeps
// the few distinct values in scalar registers -- no table traffic at run time.
#pragma once
#include "quad_tables.h"

namespace smf {
namespace tab {

constexpr int nloc(int p) { return (p + 1) * (p + 2) * (p + 3) / 6; }

// lattice multi-index of local node i (same order as smf::Lagrange)
struct Alpha {
  int a[4];
};
constexpr int kEV[6][2] = {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {1, 3}, {2, 3}};
constexpr int kFV[4][3] = {{1, 2, 3}, {0, 2, 3}, {0, 1, 3}, {0, 1, 2}};

constexpr Alpha alpha(int p, int i) {
  Alpha r{{0, 0, 0, 0}};
  if (i < 4) {
    r.a[i] = p;
    return r;
  }
  int k = i - 4;
  if (k < 6 * (p - 1)) {
    const int e = k / (p - 1), s = k % (p - 1);
    r.a[kEV[e][0]] = p - 1 - s;
    r.a[kEV[e][1]] = 1 + s;
    return r;
  }
  k -= 6 * (p - 1);
  for (int t = 0; t < 3; ++t) r.a[kFV[k][t]] = 1;
  return r;
}

// value (d < 0) or d-th reference derivative of basis i at x
constexpr double basis(int p, int i, const double* x, int d) {
  const double l[4] = {1.0 - x[0] - x[1] - x[2], x[0], x[1], x[2]};
  const double dl[4][3] = {{-1, -1, -1}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
  const Alpha al = alpha(p, i);
  double fv[4] = {0, 0, 0, 0}, fd[4] = {0, 0, 0, 0};
  for (int k = 0; k < 4; ++k) {
    double v = 1.0, dv = 0.0;
    for (int j = 0; j < al.a[k]; ++j) {
      const double c = 1.0 / (j + 1);
      const double t = (p * l[k] - j) * c;
      dv = dv * t + v * p * c;
      v *= t;
    }
    fv[k] = v;
    fd[k] = dv;
  }
  if (d < 0) return fv[0] * fv[1] * fv[2] * fv[3];
  double g = 0.0;
  for (int k = 0; k < 4; ++k) {
    double o = 1.0;
    for (int m = 0; m < 4; ++m)
      if (m != k) o *= fv[m];
    g += fd[k] * o * dl[k][d];
  }
  return g;
}

// cell rule used for stiffness (exact for degree 2p-2)
template <int P>
struct CellRule;
template <>
struct CellRule<1> {
  static constexpr int NQ = 1;
  static constexpr const double (*pts)[3] = q_tet_d1_pts;
  static constexpr const double* w = q_tet_d1_w;
};
template <>
struct CellRule<2> {
  static constexpr int NQ = 4;
  static constexpr const double (*pts)[3] = q_tet_d2_pts;
  static constexpr const double* w = q_tet_d2_w;
};
template <>
struct CellRule<3> {
  static constexpr int NQ = 11;
  static constexpr const double (*pts)[3] = q_tet_d4_pts;
  static constexpr const double* w = q_tet_d4_w;
};

template <int P>
struct CellGrad {
  static constexpr int NL = nloc(P), NQ = CellRule<P>::NQ;
  double g[NQ][NL][3];
  double w[NQ];
};

template <int P>
constexpr CellGrad<P> make_cell_grad() {
  CellGrad<P> t{};
  for (int q = 0; q < CellRule<P>::NQ; ++q) {
    t.w[q] = CellRule<P>::w[q];
    for (int i = 0; i < nloc(P); ++i)
      for (int d = 0; d < 3; ++d) t.g[q][i][d] = basis(P, i, CellRule<P>::pts[q], d);
  }
  return t;
}

}  // namespace tab
}  // namespace smf
