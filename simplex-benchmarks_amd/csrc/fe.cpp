#include "fe.h"

#include <stdexcept>

#include "quad_tables.h"

namespace smf {

namespace {
template <int N, int D>
Quadrature make_rule(int degree, const double (&pts)[N][D], const double (&w)[N]) {
  Quadrature q;
  q.degree = degree;
  for (int i = 0; i < N; ++i) {
    std::array<double, 3> x{0, 0, 0};
    for (int d = 0; d < D; ++d) x[d] = pts[i][d];
    q.pts.push_back(x);
    q.w.push_back(w[i]);
  }
  return q;
}

const double kVerts[4][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
// gradient of barycentric l_k with respect to (x,y,z)
const double kGradL[4][3] = {{-1, -1, -1}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
}  // namespace

Quadrature tet_rule(int degree) {
  if (degree <= 1) return make_rule(1, q_tet_d1_pts, q_tet_d1_w);
  if (degree <= 2) return make_rule(2, q_tet_d2_pts, q_tet_d2_w);
  if (degree <= 4) return make_rule(4, q_tet_d4_pts, q_tet_d4_w);
  if (degree <= 5) return make_rule(5, q_tet_d5_pts, q_tet_d5_w);
  if (degree <= 6) return make_rule(6, q_tet_d6_pts, q_tet_d6_w);
  throw std::runtime_error("no tetrahedral rule of that degree");
}

Quadrature tri_rule(int degree) {
  if (degree <= 1) return make_rule(1, q_tri_d1_pts, q_tri_d1_w);
  if (degree <= 2) return make_rule(2, q_tri_d2_pts, q_tri_d2_w);
  if (degree <= 4) return make_rule(4, q_tri_d4_pts, q_tri_d4_w);
  if (degree <= 6) return make_rule(6, q_tri_d6_pts, q_tri_d6_w);
  throw std::runtime_error("no triangle rule of that degree");
}

Lagrange::Lagrange(int degree) : p(degree), n_loc(n_loc_of(degree)) {
  if (p < 1 || p > 3) throw std::runtime_error("degree must be 1..3");
  for (int v = 0; v < 4; ++v) {
    LocalNode n{0, v, 0, {0, 0, 0, 0}};
    n.alpha[v] = p;
    nodes.push_back(n);
  }
  for (int e = 0; e < 6; ++e)
    for (int s = 0; s < p - 1; ++s) {
      // s-th node from the lower vertex a towards b: alpha_a = p-1-s, alpha_b = 1+s
      LocalNode n{1, e, s, {0, 0, 0, 0}};
      n.alpha[kEdgeVerts[e][0]] = p - 1 - s;
      n.alpha[kEdgeVerts[e][1]] = 1 + s;
      nodes.push_back(n);
    }
  if (p == 3)
    for (int f = 0; f < 4; ++f) {
      LocalNode n{2, f, 0, {0, 0, 0, 0}};
      for (int k = 0; k < 3; ++k) n.alpha[kFaceVerts[f][k]] = 1;
      nodes.push_back(n);
    }
  if ((int)nodes.size() != n_loc) throw std::runtime_error("node count mismatch");
}

std::array<double, 3> Lagrange::node_point(int i) const {
  std::array<double, 3> x{0, 0, 0};
  for (int k = 0; k < 4; ++k)
    for (int d = 0; d < 3; ++d) x[d] += nodes[i].alpha[k] * kVerts[k][d] / p;
  return x;
}

// phi_alpha(l) = prod_k prod_{j<alpha_k} (p l_k - j) / (j + 1): the Lagrange
// polynomial of the lattice node alpha/p (vanishes on every other lattice node).
void Lagrange::eval(const double x[3], double* values, double* grads) const {
  const double l[4] = {1.0 - x[0] - x[1] - x[2], x[0], x[1], x[2]};
  for (int i = 0; i < n_loc; ++i) {
    // factor values and their derivatives with respect to l_k
    double fval[4], fder[4];
    for (int k = 0; k < 4; ++k) {
      double v = 1.0, dv = 0.0;
      for (int j = 0; j < nodes[i].alpha[k]; ++j) {
        const double c = 1.0 / (j + 1);
        const double t = (p * l[k] - j) * c;
        dv = dv * t + v * p * c;
        v *= t;
      }
      fval[k] = v;
      fder[k] = dv;
    }
    double val = fval[0] * fval[1] * fval[2] * fval[3];
    if (values) values[i] = val;
    if (grads) {
      double g[3] = {0, 0, 0};
      for (int k = 0; k < 4; ++k) {
        double others = 1.0;
        for (int m = 0; m < 4; ++m)
          if (m != k) others *= fval[m];
        const double dk = fder[k] * others;
        for (int d = 0; d < 3; ++d) g[d] += dk * kGradL[k][d];
      }
      for (int d = 0; d < 3; ++d) grads[i * 3 + d] = g[d];
    }
  }
}

std::array<double, 3> face_to_cell(int f, double s, double t) {
  const int a = kFaceVerts[f][0], b = kFaceVerts[f][1], c = kFaceVerts[f][2];
  std::array<double, 3> x;
  for (int d = 0; d < 3; ++d)
    x[d] = kVerts[a][d] + s * (kVerts[b][d] - kVerts[a][d]) + t * (kVerts[c][d] - kVerts[a][d]);
  return x;
}

}  // namespace smf
