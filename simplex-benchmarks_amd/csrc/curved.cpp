// Curved cells: a P2 geometry map per cell whose edge nodes lie on a
// manifold, and the quadrature-point geometry of the general D_e branch.
//
// Reference: MappingFE(FE_SimplexP(2)) with a SphericalManifold around
// (5, 0, 0) attached to every cell (CG/cells_batched/application.cc:1613,
// 1668-1677, 1763-1767, 1803-1806); the curvilinear quadrature-point
// operator g_q <- JxW_q jac_q^T (jac_q g_q) with jac = J^-T (:1361-1376).
//
// Geometry nodes of cell K: its four vertices as given, and per edge (a, b)
// the manifold's midpoint.  For the spherical manifold around c this is the
// geodesic midpoint: c + (r_a + r_b)/2 * (e_a + e_b)/|e_a + e_b| with
// r = |x - c|, e = (x - c)/r (the weight-1/2 point of the sphere-direction
// slerp with linearly interpolated radius).  The formula is symmetric in a
// and b, so the two cells of an edge agree on its node and the map is
// continuous.  Mode 2 uses the straight midpoint (x_a + x_b)/2.
#include <cmath>
#include <stdexcept>
#include <vector>

#include "fe.h"
#include "op.h"

namespace smf {

namespace {

void edge_node(const double* xa, const double* xb, int manifold, const double* c, double* out) {
  if (manifold == 2) {
    for (int d = 0; d < 3; ++d) out[d] = 0.5 * (xa[d] + xb[d]);
    return;
  }
  double va[3], vb[3], ra = 0, rb = 0;
  for (int d = 0; d < 3; ++d) {
    va[d] = xa[d] - c[d];
    vb[d] = xb[d] - c[d];
    ra += va[d] * va[d];
    rb += vb[d] * vb[d];
  }
  ra = std::sqrt(ra);
  rb = std::sqrt(rb);
  if (!(ra > 0) || !(rb > 0)) throw std::runtime_error("curved cells: a vertex at the manifold centre");
  double e[3], ne = 0;
  for (int d = 0; d < 3; ++d) {
    e[d] = va[d] / ra + vb[d] / rb;
    ne += e[d] * e[d];
  }
  ne = std::sqrt(ne);
  if (!(ne > 1e-12)) throw std::runtime_error("curved cells: an edge spans opposite directions from the centre");
  const double r = 0.5 * (ra + rb);
  for (int d = 0; d < 3; ++d) out[d] = c[d] + r * e[d] / ne;
}

}  // namespace

void p2_geometry_nodes(const double* const xv[4], int manifold, const double* center, double X[10][3]) {
  for (int v = 0; v < 4; ++v)
    for (int d = 0; d < 3; ++d) X[v][d] = xv[v][d];
  for (int e = 0; e < 6; ++e) edge_node(X[kEdgeVerts[e][0]], X[kEdgeVerts[e][1]], manifold, center, X[4 + e]);
}

void p2_jacobian_coeffs(const double X[10][3], double jc[27]) {
  // J(xi) = C + sum_f xi_f D_f.  With lambda = (1 - xi_1 - xi_2 - xi_3,
  // xi_1, xi_2, xi_3), dx/dlambda_a = 4 sum_b N_ab lambda_b - X_a with
  // N_aa = X_a and N_ab the edge node of (a, b); J column e (vertex e + 1)
  // is dx/dlambda_{e+1} - dx/dlambda_0, so
  //   C_e  = 4 X_0e - 3 X_0 - X_e,
  //   D_e column e = 4 (X_e + X_0 - 2 X_0e),
  //   D_f column e = 4 (X_ef - X_0e - X_0f + X_0)  (e != f).
  // Straight edges give C_e = X_e - X_0 and D = 0.
  const int e0[3] = {4, 5, 6};                                 // edge nodes (0,1) (0,2) (0,3)
  const int pair[3][3] = {{0, 1, 7}, {0, 2, 8}, {1, 2, 9}};  // (e, f, node of edge (e+1, f+1))
  for (int e = 0; e < 3; ++e)
    for (int d = 0; d < 3; ++d) {
      jc[3 * e + d] = 4 * X[e0[e]][d] - 3 * X[0][d] - X[e + 1][d];
      jc[9 + 3 * e + d] = 4 * (X[e + 1][d] + X[0][d] - 2 * X[e0[e]][d]);
    }
  for (int i = 0; i < 3; ++i)
    for (int d = 0; d < 3; ++d) {
      const int e = pair[i][0], f = pair[i][1], ef = pair[i][2];
      jc[18 + 3 * i + d] = 4 * (X[ef][d] - X[e0[e]][d] - X[e0[f]][d] + X[0][d]);
    }
}

void p2_jacobian_at(const double jc[27], const double* xi, double J[3][3]) {
  for (int d = 0; d < 3; ++d) {
    J[d][0] = jc[d] + xi[0] * jc[9 + d] + xi[1] * jc[18 + d] + xi[2] * jc[21 + d];
    J[d][1] = jc[3 + d] + xi[1] * jc[12 + d] + xi[0] * jc[18 + d] + xi[2] * jc[24 + d];
    J[d][2] = jc[6 + d] + xi[2] * jc[15 + d] + xi[0] * jc[21 + d] + xi[1] * jc[24 + d];
  }
}

void build_curved_geometry(const Mesh& mesh, const mf_options& opt, CGPlan& plan) {
  const int64_t nc = plan.n_cells;
  const int nl = plan.n_loc;
  const Lagrange geo(2), fe(plan.p);
  const Quadrature q = tet_rule(2 * plan.p);
  const int nq = q.size();
  // geometry-basis gradients at the rule's points and values at the FE nodes
  std::vector<double> gg((size_t)nq * 10 * 3), gv((size_t)nl * 10);
  for (int k = 0; k < nq; ++k) geo.eval(q.pts[k].data(), nullptr, &gg[(size_t)k * 30]);
  for (int i = 0; i < nl; ++i) {
    const auto xr = fe.node_point(i);
    geo.eval(xr.data(), &gv[(size_t)i * 10], nullptr);
  }
  plan.curved = true;
  plan.qmetric.assign((size_t)nq * 6 * nc, 0.0);
  plan.jgeo.assign((size_t)27 * nc, 0.0);
  const int ab[6][2] = {{0, 0}, {0, 1}, {0, 2}, {1, 1}, {1, 2}, {2, 2}};
  // each support point is written by the first cell of its DoF, so the
  // result does not depend on the thread schedule
  std::vector<int64_t> owner(plan.n_dofs, -1);
  for (int64_t k = 0; k < nc; ++k)
    for (int i = 0; i < nl; ++i) {
      const int64_t dof = plan.cell_dofs[k * nl + i];
      if (owner[dof] < 0) owner[dof] = k;
    }
  int bad = 0;
#pragma omp parallel for schedule(static)
  for (int64_t k = 0; k < nc; ++k) {
    const int64_t c = plan.cell_perm[k];
    double X[10][3];
    {
      const double* xv[4];
      for (int v = 0; v < 4; ++v) xv[v] = &mesh.x[3 * (int64_t)mesh.cells[4 * c + v]];
      p2_geometry_nodes(xv, opt.manifold, opt.manifold_center, X);
      double jc[27];
      p2_jacobian_coeffs(X, jc);
      for (int t = 0; t < 27; ++t) plan.jgeo[(size_t)t * nc + k] = jc[t];
    }
    for (int qq = 0; qq < nq; ++qq) {
      // J[d][e] = sum_i X_i[d] d(phi_i)/d(xi_e)
      double J[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
      for (int i = 0; i < 10; ++i)
        for (int d = 0; d < 3; ++d)
          for (int e = 0; e < 3; ++e) J[d][e] += X[i][d] * gg[((size_t)qq * 10 + i) * 3 + e];
      // set-up self-check of the kernels' affine form of J against the basis form
      {
        const double* xi = q.pts[qq].data();
        double dev = 0, mag = 0;
        for (int e = 0; e < 3; ++e)
          for (int d = 0; d < 3; ++d) {
            auto g = [&](int t) { return plan.jgeo[(size_t)t * nc + k]; };
            double v = g(3 * e + d) + xi[e] * g(9 + 3 * e + d);
            for (int i = 0; i < 3; ++i) {
              const int a = i == 2 ? 1 : 0, b = i == 0 ? 1 : 2;  // pairs (0,1) (0,2) (1,2)
              if (e == a) v += xi[b] * g(18 + 3 * i + d);
              if (e == b) v += xi[a] * g(18 + 3 * i + d);
            }
            dev = std::fmax(dev, std::fabs(v - J[d][e]));
            mag = std::fmax(mag, std::fabs(J[d][e]));
          }
        if (dev > 1e-12 * mag) {
#pragma omp atomic write
          bad = 2;
        }
      }
      const double c00 = J[1][1] * J[2][2] - J[1][2] * J[2][1], c01 = J[1][2] * J[2][0] - J[1][0] * J[2][2],
                   c02 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
      const double det = J[0][0] * c00 + J[0][1] * c01 + J[0][2] * c02;
      if (!(det > 0)) {
#pragma omp atomic write
        bad = 1;
        continue;
      }
      double Ji[3][3];  // J^-1 = adj(J) / det
      Ji[0][0] = c00 / det;
      Ji[1][0] = c01 / det;
      Ji[2][0] = c02 / det;
      Ji[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) / det;
      Ji[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) / det;
      Ji[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) / det;
      Ji[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) / det;
      Ji[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) / det;
      Ji[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) / det;
      for (int t = 0; t < 6; ++t) {
        double s = 0;
        for (int e = 0; e < 3; ++e) s += Ji[ab[t][0]][e] * Ji[ab[t][1]][e];
        plan.qmetric[((size_t)qq * 6 + t) * nc + k] = q.w[qq] * det * s;
      }
    }
    // support points of the cell's DoFs under the curved map
    for (int i = 0; i < nl; ++i) {
      const int64_t dof = plan.cell_dofs[k * nl + i];
      if (owner[dof] != k) continue;
      for (int d = 0; d < 3; ++d) {
        double s = 0;
        for (int j = 0; j < 10; ++j) s += gv[(size_t)i * 10 + j] * X[j][d];
        plan.dof_points[3 * dof + d] = s;
      }
    }
  }
  if (bad == 2) throw std::runtime_error("curved cells: affine Jacobian form disagrees with the P2 map");
  if (bad) throw std::runtime_error("curved cells: a quadrature point with non-positive Jacobian determinant");
}

}  // namespace smf
