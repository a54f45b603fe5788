// Discontinuous-Galerkin (SIPG) setup: cell order, canonical local vertex
// order, neighbours, geometry, penalty and reference tables; plus a host
// execution of the kernel's per-cell algorithm (self-check of the tables).
//
// Counterpart of Operator::reinit of the DG operator
// (DG/cells_batched/application.cc:1170-1277): MatrixFree face data and the
// per-cell penalty tau = surface/volume (1193-1245).  The reference groups
// faces into 4x4x6x6 (face number, orientation) classes so that each class
// shares one interpolation matrix (1737-1788); here every cell lists its
// vertices in ascending global order, which gives both cells of a face the
// same face parametrisation, so a single table per local face number serves
// every face.
#include <algorithm>
#include <cmath>
#include <stdexcept>

#include "dg_cell.h"
#include "op.h"

namespace smf {

namespace {
const double kNhat[4][3] = {{1, 1, 1}, {-1, 0, 0}, {0, -1, 0}, {0, 0, -1}};
}

int cg_quadrature_degree(int p);

void build_dg_plan(const Mesh& mesh, int p, int ncomp, const mf_options& opt, bool helmholtz,
                   DGPlan& plan) {
  const Lagrange fe(p);
  const int nl = fe.n_loc;
  const int64_t nc = mesh.nc;
  plan.p = p;
  plan.n_loc = nl;
  plan.ncomp = ncomp;
  plan.n_cells = nc;
  plan.n_dofs = nc * ncomp * nl;
  plan.penalty_kind = opt.penalty_kind;
  plan.penalty_factor = opt.penalty_kind == 1 ? 3.0 * (p + 1.0) * (p + 2.0) : (p + 1.0) * (p + 3.0) / 3.0;
  plan.mass = helmholtz ? opt.mass_coefficient : 0.0;
  plan.laplace = helmholtz ? opt.laplace_coefficient : 1.0;

  std::vector<int64_t> perm, patch_ptr;
  reorder_cells(mesh, opt.reorder != 0, 64, perm, patch_ptr);
  plan.cell_perm = perm;
  std::vector<int64_t> inv(nc);
  for (int64_t k = 0; k < nc; ++k) inv[perm[k]] = k;

  // canonical local vertex order
  plan.cell_verts.resize(4 * nc);
  for (int64_t k = 0; k < nc; ++k) {
    int32_t v[4];
    for (int t = 0; t < 4; ++t) v[t] = mesh.cells[4 * perm[k] + t];
    std::sort(v, v + 4);
    for (int t = 0; t < 4; ++t) plan.cell_verts[4 * k + t] = v[t];
  }
  // local face (canonical numbering) of an input (cell, input-local face)
  auto canon_face = [&](int64_t c, int lf) {
    const int32_t opp = mesh.cells[4 * c + lf];
    const int64_t k = inv[c];
    for (int t = 0; t < 4; ++t)
      if (plan.cell_verts[4 * k + t] == opp) return t;
    throw std::runtime_error("face bookkeeping");
  };

  // neighbours
  FaceInfo fi = build_faces(mesh);
  plan.nbr.assign(4 * nc, -1);
  plan.nface.assign(nc, 0);
  std::vector<uint8_t> is_bdry(4 * nc, 1);
  for (int64_t f = 0; f < fi.n_faces; ++f) {
    const int64_t c0 = fi.face_cell[f][0], c1 = fi.face_cell[f][1];
    const int f0 = canon_face(c0, fi.face_lf[f][0]);
    if (c1 < 0) continue;
    const int f1 = canon_face(c1, fi.face_lf[f][1]);
    const int64_t k0 = inv[c0], k1 = inv[c1];
    plan.nbr[(size_t)f0 * nc + k0] = (int32_t)k1;
    plan.nbr[(size_t)f1 * nc + k1] = (int32_t)k0;
    plan.nface[k0] |= (uint32_t)f1 << (8 * f0);
    plan.nface[k1] |= (uint32_t)f0 << (8 * f1);
    is_bdry[4 * k0 + f0] = 0;
    is_bdry[4 * k1 + f1] = 0;
  }
  plan.n_boundary_faces = fi.n_boundary;
  plan.n_interior_faces = fi.n_faces - fi.n_boundary;

  // geometry and penalty
  plan.geo.assign(11 * nc, 0.0);
  plan.dof_points.resize(3 * nc * nl);
  Mesh sorted_mesh;
  for (int64_t k = 0; k < nc; ++k) {
    double J[3][3], x0[3];
    const int32_t* v = &plan.cell_verts[4 * k];
    for (int d = 0; d < 3; ++d) {
      x0[d] = mesh.x[3 * v[0] + d];
      for (int t = 0; t < 3; ++t) J[d][t] = mesh.x[3 * v[t + 1] + d] - x0[d];
    }
    const double det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
                       J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                       J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
    if (!(std::fabs(det) > 0)) throw std::runtime_error("degenerate cell");
    // J^-T = cof(J) / det
    double jit[3][3];
    jit[0][0] = (J[1][1] * J[2][2] - J[1][2] * J[2][1]) / det;
    jit[0][1] = -(J[1][0] * J[2][2] - J[1][2] * J[2][0]) / det;
    jit[0][2] = (J[1][0] * J[2][1] - J[1][1] * J[2][0]) / det;
    jit[1][0] = -(J[0][1] * J[2][2] - J[0][2] * J[2][1]) / det;
    jit[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) / det;
    jit[1][2] = -(J[0][0] * J[2][1] - J[0][1] * J[2][0]) / det;
    jit[2][0] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) / det;
    jit[2][1] = -(J[0][0] * J[1][2] - J[0][2] * J[1][0]) / det;
    jit[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) / det;
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) plan.geo[(size_t)(3 * a + b) * nc + k] = jit[a][b];
    const double adet = std::fabs(det);
    plan.geo[(size_t)9 * nc + k] = adet;
    double surf = 0;
    for (int f = 0; f < 4; ++f) {
      double nn = 0;
      for (int a = 0; a < 3; ++a) {
        double s = 0;
        for (int b = 0; b < 3; ++b) s += jit[a][b] * kNhat[f][b];
        nn += s * s;
      }
      const double area = 0.5 * adet * std::sqrt(nn);
      surf += (is_bdry[4 * k + f] ? 1.0 : 0.5) * area;
    }
    plan.geo[(size_t)10 * nc + k] = surf / (adet / 6.0);
    for (int i = 0; i < nl; ++i) {
      const auto xr = fe.node_point(i);
      for (int a = 0; a < 3; ++a) {
        double s = x0[a];
        for (int b = 0; b < 3; ++b) s += J[a][b] * xr[b];
        plan.dof_points[3 * (k * nl + i) + a] = s;
      }
    }
  }

  // tables
  const Quadrature qc = tet_rule(helmholtz ? 2 * p : cg_quadrature_degree(p));
  const Quadrature qf = tri_rule(2 * p);
  plan.n_q = qc.size();
  plan.n_qf = qf.size();
  std::vector<double> val(nl), grad(3 * nl);
  plan.cell_tab.assign((size_t)plan.n_q * nl * 4 + plan.n_q, 0.0);
  for (int q = 0; q < plan.n_q; ++q) {
    fe.eval(qc.pts[q].data(), val.data(), grad.data());
    for (int i = 0; i < nl; ++i) {
      for (int d = 0; d < 3; ++d) plan.cell_tab[((size_t)q * nl + i) * 3 + d] = grad[3 * i + d];
      plan.cell_tab[(size_t)plan.n_q * nl * 3 + plan.n_q + (size_t)q * nl + i] = val[i];
    }
    plan.cell_tab[(size_t)plan.n_q * nl * 3 + q] = qc.w[q];
  }
  // face tables: per face f: values [q][i] then gradients [q][i][d]; then w[q]
  const int nqf = plan.n_qf;
  const size_t fstride = (size_t)nqf * nl * 4;
  plan.face_tab.assign(4 * fstride + nqf, 0.0);
  for (int f = 0; f < 4; ++f)
    for (int q = 0; q < nqf; ++q) {
      const auto x = face_to_cell(f, qf.pts[q][0], qf.pts[q][1]);
      fe.eval(x.data(), val.data(), grad.data());
      for (int i = 0; i < nl; ++i) {
        plan.face_tab[f * fstride + (size_t)q * nl + i] = val[i];
        for (int d = 0; d < 3; ++d)
          plan.face_tab[f * fstride + (size_t)nqf * nl + ((size_t)q * nl + i) * 3 + d] = grad[3 * i + d];
      }
    }
  for (int q = 0; q < nqf; ++q) plan.face_tab[4 * fstride + q] = qf.w[q];
}

namespace {
// The kernel's per-cell algorithm (dg_cell.h), executed on the host.
template <int P, typename T, int NC, bool MASS>
void host_dg_run(const Operator& op, int mode, T* dst, const T* src) {
  const DGPlan& pl = op.dg;
  std::vector<T> geo(pl.geo.begin(), pl.geo.end());
  if (op.opt.dg_kernel == 3) {
    // four lanes per cell (one face each), summed in lane order
    constexpr int NL = dgt::nloc(P);
    T yr[4][NL];
    for (int64_t k = 0; k < pl.n_cells; ++k)
      for (int c = 0; c < NC; ++c) {
        for (int f = 0; f < 4; ++f) {
          if (mode == 0)
            dgt::apply_cell_lane<P, T, NC, MASS, 0>(k, f, c, pl.n_cells, geo.data(), pl.nbr.data(), pl.nface.data(),
                                                    (T)pl.penalty_factor, (T)pl.mass, (T)pl.laplace, src, yr[f],
                                                    pl.penalty_kind);
          else
            dgt::apply_cell_lane<P, T, NC, MASS, 1>(k, f, c, pl.n_cells, geo.data(), pl.nbr.data(), pl.nface.data(),
                                                    (T)pl.penalty_factor, (T)pl.mass, (T)pl.laplace, src, yr[f],
                                                    pl.penalty_kind);
        }
        constexpr dgt::LaneInv<P> LI = dgt::make_lane_inv<P>();
        for (int i = 0; i < NL; ++i)
          dst[(k * NC + c) * NL + i] =
              ((yr[0][LI.j[0][i]] + yr[1][LI.j[1][i]]) + yr[2][LI.j[2][i]]) + yr[3][LI.j[3][i]];
      }
    return;
  }
  for (int64_t k = 0; k < pl.n_cells; ++k) {
    if (mode == 0)
      dgt::apply_cell<P, T, NC, MASS, 0>(k, pl.n_cells, geo.data(), pl.nbr.data(), pl.nface.data(),
                                         (T)pl.penalty_factor, (T)pl.mass, (T)pl.laplace, src, dst, pl.penalty_kind);
    else
      dgt::apply_cell<P, T, NC, MASS, 1>(k, pl.n_cells, geo.data(), pl.nbr.data(), pl.nface.data(),
                                         (T)pl.penalty_factor, (T)pl.mass, (T)pl.laplace, src, dst, pl.penalty_kind);
  }
}

template <int P, typename T>
void host_dg_p(const Operator& op, int mode, T* dst, const T* src) {
  if (op.kind == MF_DG_SIPG) host_dg_run<P, T, 1, false>(op, mode, dst, src);
  else if (op.kind == MF_DG_VECTOR) host_dg_run<P, T, 3, false>(op, mode, dst, src);
  else host_dg_run<P, T, 3, true>(op, mode, dst, src);
}

template <typename T>
void host_dg_t(const Operator& op, int mode, T* dst, const T* src) {
  if (op.dg.p == 1) host_dg_p<1, T>(op, mode, dst, src);
  else if (op.dg.p == 2) host_dg_p<2, T>(op, mode, dst, src);
  else host_dg_p<3, T>(op, mode, dst, src);
}
}  // namespace

void host_apply_dg(const Operator& op, int mode, void* dst, const void* src) {
  if (op.precision == MF_F64)
    host_dg_t<double>(op, mode, static_cast<double*>(dst), static_cast<const double*>(src));
  else
    host_dg_t<float>(op, mode, static_cast<float*>(dst), static_cast<const float*>(src));
}

}  // namespace smf
