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
#include <vector>

#include "dg_cell.h"
#include "op.h"

namespace smf {

namespace {
const double kNhat[4][3] = {{1, 1, 1}, {-1, 0, 0}, {0, -1, 0}, {0, 0, -1}};

// Curved cells (the reference's MappingFE(FE_SimplexP(2)) + manifold,
// CG/cells_batched/application.cc:1613, 1668-1677, with the DG operator of
// DG/cells_batched/application.cc:1614-1666 (general D_e) and 2191-2238
// (curvilinear face flux)): each cell's P2 geometry map in the ascending
// vertex order (its edge nodes are symmetric in the two ends, so both cells
// of a face map it identically), tau_K = (sum_bdry |F| + 1/2 sum_int |F|)/|K|
// with |K| = sum_q w_q |det J(xi_q)| and |F| = sum_q w_q |cof J(xi_q) N_f| by
// the operator's rules (N_f the reference face normal scaled by the
// reference face's area over 1/2), and the support points under the map.
void build_dg_curved(const Mesh& mesh, const mf_options& opt, const Quadrature& qc, const Quadrature& qf,
                     const std::vector<uint8_t>& is_bdry, DGPlan& plan) {
  const int64_t nc = plan.n_cells;
  const int nl = plan.n_loc;
  const Lagrange geo2(2), fe(plan.p);
  plan.curved = true;
  plan.jgeo.assign((size_t)27 * nc, 0.0);
  std::vector<double> gv((size_t)nl * 10);
  for (int i = 0; i < nl; ++i) {
    const auto xr = fe.node_point(i);
    geo2.eval(xr.data(), &gv[(size_t)i * 10], nullptr);
  }
  int bad = 0;
#pragma omp parallel for schedule(static)
  for (int64_t k = 0; k < nc; ++k) {
    const double* xv[4];
    for (int v = 0; v < 4; ++v) xv[v] = &mesh.x[3 * (int64_t)plan.cell_verts[4 * k + v]];
    double X[10][3], jc[27];
    p2_geometry_nodes(xv, opt.manifold, opt.manifold_center, X);
    p2_jacobian_coeffs(X, jc);
    for (int t = 0; t < 27; ++t) plan.jgeo[(size_t)t * nc + k] = jc[t];
    auto cof = [&](const double* xi, double C[3][3], double* det) {
      double J[3][3];
      p2_jacobian_at(jc, xi, J);
      // cof(J)[d][e] = cofactor of J[d][e]
      C[0][0] = J[1][1] * J[2][2] - J[1][2] * J[2][1];
      C[0][1] = J[1][2] * J[2][0] - J[1][0] * J[2][2];
      C[0][2] = J[1][0] * J[2][1] - J[1][1] * J[2][0];
      C[1][0] = J[0][2] * J[2][1] - J[0][1] * J[2][2];
      C[1][1] = J[0][0] * J[2][2] - J[0][2] * J[2][0];
      C[1][2] = J[0][1] * J[2][0] - J[0][0] * J[2][1];
      C[2][0] = J[0][1] * J[1][2] - J[0][2] * J[1][1];
      C[2][1] = J[0][2] * J[1][0] - J[0][0] * J[1][2];
      C[2][2] = J[0][0] * J[1][1] - J[0][1] * J[1][0];
      *det = J[0][0] * C[0][0] + J[0][1] * C[0][1] + J[0][2] * C[0][2];
    };
    // ascending vertex order may reverse the orientation: det J keeps one
    // sign over a valid cell, |det J| weights the integrals
    double vol = 0, sgn = 0;
    for (int q = 0; q < qc.size(); ++q) {
      double C[3][3], det;
      cof(qc.pts[q].data(), C, &det);
      if (q == 0) sgn = det > 0 ? 1.0 : -1.0;
      if (!(det * sgn > 0)) {
#pragma omp atomic write
        bad = 1;
      }
      vol += qc.w[q] * std::fabs(det);
    }
    double surf = 0;
    for (int f = 0; f < 4; ++f) {
      double area = 0;
      for (int q = 0; q < qf.size(); ++q) {
        const auto xi = face_to_cell(f, qf.pts[q][0], qf.pts[q][1]);
        double C[3][3], det;
        cof(xi.data(), C, &det);
        double a2 = 0;
        for (int d = 0; d < 3; ++d) {
          const double a = C[d][0] * kNhat[f][0] + C[d][1] * kNhat[f][1] + C[d][2] * kNhat[f][2];
          a2 += a * a;
        }
        area += qf.w[q] * std::sqrt(a2);
      }
      surf += (is_bdry[4 * k + f] ? 1.0 : 0.5) * area;
    }
    plan.geo[(size_t)10 * nc + k] = surf / vol;
    for (int i = 0; i < nl; ++i)
      for (int d = 0; d < 3; ++d) {
        double sxy = 0;
        for (int j = 0; j < 10; ++j) sxy += gv[(size_t)i * 10 + j] * X[j][d];
        plan.dof_points[3 * (k * nl + i) + d] = sxy;
      }
  }
  if (bad) throw std::runtime_error("curved cells: the Jacobian determinant changes sign inside a cell");
}
// Data of the four-lane kernel: the neighbour words (dgt::kNbrShift), the
// cell's barycentric metric (dgt::kBGeo
// values per cell, [7][cell]) and face records (dgt::kFaceRec values per cell and
// local face, [cell][face][4]): the neighbour's barycentric gradients of the
// reference directions of the frame of its face fb (frame vertices kFV[fb]
// ascending, then fb; direction d is frame vertex d + 1), dotted with
// this cell's outward unit normal, and the face penalty sigma with the
// penalty factor: kind 0 factor max(tau-, tau+) (tau on the boundary), kind 1
// factor max(len, len |det J-| / |det J+|), len = |grad lambda_f|.  Formed in
// FP64 from the cell geometry (geo rows: J^-T row-major, |det J|, tau).
void build_dg_lane_data(DGPlan& plan) {
  const int64_t nc = plan.n_cells;
  plan.fgeo.assign((size_t)nc * 4 * dgt::kFaceRec, 0.0);
  auto grad = [&](int64_t k, int a, double* g) {  // grad lambda_a of cell k
    for (int d = 0; d < 3; ++d) {
      if (a == 0) {
        g[d] = 0;
        for (int b = 0; b < 3; ++b) g[d] -= plan.geo[(size_t)(3 * d + b) * nc + k];
      } else {
        g[d] = plan.geo[(size_t)(3 * d + a - 1) * nc + k];
      }
    }
  };
  // neighbour words: cell index and local face fb packed, -1 on the boundary
  if (nc > (int64_t)dgt::kNbrMask) throw std::runtime_error("DG: too many cells for the neighbour words");
  plan.nbrf.assign((size_t)nc * 4, -1);
  for (int64_t k = 0; k < nc; ++k)
    for (int f = 0; f < 4; ++f) {
      const int32_t kb = plan.nbr[(size_t)f * nc + k];
      if (kb >= 0)
        plan.nbrf[(size_t)k * 4 + f] = kb | (int32_t)(((plan.nface[k] >> (8 * f)) & 3) << dgt::kNbrShift);
    }
  // barycentric metric B_ab = |det J| grad lambda_a . grad lambda_b, a < b, and |det J|
  plan.bgeo.assign((size_t)dgt::kBGeo * nc, 0.0);
  for (int64_t k = 0; k < nc; ++k) {
    const double adet = plan.geo[(size_t)9 * nc + k];
    double gl[4][3];
    for (int a = 0; a < 4; ++a) grad(k, a, gl[a]);
    for (int a = 0; a < 4; ++a)
      for (int b = a + 1; b < 4; ++b)
        plan.bgeo[(size_t)dgt::kBPair[a][b] * nc + k] =
            adet * (gl[a][0] * gl[b][0] + gl[a][1] * gl[b][1] + gl[a][2] * gl[b][2]);
    plan.bgeo[(size_t)6 * nc + k] = adet;
  }
  for (int64_t k = 0; k < nc; ++k) {
    const double adet = plan.geo[(size_t)9 * nc + k], tau = plan.geo[(size_t)10 * nc + k];
    for (int f = 0; f < 4; ++f) {
      double* r = &plan.fgeo[((size_t)k * 4 + f) * dgt::kFaceRec];
      double gf[3];
      grad(k, f, gf);
      const double len = std::sqrt(gf[0] * gf[0] + gf[1] * gf[1] + gf[2] * gf[2]);
      const double n[3] = {-gf[0] / len, -gf[1] / len, -gf[2] / len};
      const int32_t kb = plan.nbr[(size_t)f * nc + k];
      if (kb >= 0) {
        const int fb = (int)((plan.nface[k] >> (8 * f)) & 3);
        // reference direction d of the frame is its vertex d + 1: (kFV[fb][0], kFV[fb][1], kFV[fb][2], fb)
        const int frame[3] = {tab::kFV[fb][1], tab::kFV[fb][2], fb};
        for (int d = 0; d < 3; ++d) {
          double gb[3];
          grad(kb, frame[d], gb);
          r[d] = gb[0] * n[0] + gb[1] * n[1] + gb[2] * n[2];
        }
      }
      const double adetb = kb >= 0 ? plan.geo[(size_t)9 * nc + kb] : adet;
      const double taub = kb >= 0 ? plan.geo[(size_t)10 * nc + kb] : tau;
      r[3] = plan.penalty_kind == 1 ? plan.penalty_factor * std::max(len, len * adet / adetb)
                                    : plan.penalty_factor * std::max(tau, taub);
    }
  }
}

}  // namespace

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

  // tables (curved cells: the cell rule of degree 2p, as the integrands are
  // no longer polynomials)
  const bool curved = opt.manifold != 0;
  const Quadrature qc = tet_rule(helmholtz || curved ? 2 * p : cg_quadrature_degree(p));
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
  // the rule points (read by the curved-cell kernel): cell points [q][3],
  // face points in cell coordinates [f][q][3]
  for (int q = 0; q < plan.n_q; ++q)
    for (int d = 0; d < 3; ++d) plan.cell_tab.push_back(qc.pts[q][d]);
  for (int f = 0; f < 4; ++f)
    for (int q = 0; q < nqf; ++q) {
      const auto x = face_to_cell(f, qf.pts[q][0], qf.pts[q][1]);
      for (int d = 0; d < 3; ++d) plan.face_tab.push_back(x[d]);
    }
  if (curved) build_dg_curved(mesh, opt, qc, qf, is_bdry, plan);
  else build_dg_lane_data(plan);
}

namespace {
// The kernel's per-cell algorithm (dg_cell.h), executed on the host.
template <int P, typename T, int NC, bool MASS>
void host_dg_run(const Operator& op, int mode, T* dst, const T* src) {
  const DGPlan& pl = op.dg;
  std::vector<T> geo(pl.geo.begin(), pl.geo.end());
  if (op.opt.dg_kernel == 3 || op.opt.dg_kernel == 4) {
    const std::vector<T> fgeo(pl.fgeo.begin(), pl.fgeo.end()), bgeo(pl.bgeo.begin(), pl.bgeo.end());
    const int32_t* nbrf = pl.nbrf.data();
    // four lanes per cell (one face each), summed in lane order
    constexpr int NL = dgt::nloc(P);
    T yr[4][NL];
    for (int64_t k = 0; k < pl.n_cells; ++k)
      for (int c = 0; c < NC; ++c) {
        for (int f = 0; f < 4; ++f) {
          if (mode == 0)
            dgt::apply_cell_lane<P, T, NC, MASS, 0>(k, f, c, pl.n_cells, bgeo.data(), fgeo.data(), nbrf,
                                                    (T)pl.mass, (T)pl.laplace, src, yr[f]);
          else
            dgt::apply_cell_lane<P, T, NC, MASS, 1>(k, f, c, pl.n_cells, bgeo.data(), fgeo.data(), nbrf,
                                                    (T)pl.mass, (T)pl.laplace, src, yr[f]);
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
  if (op.dg.curved) throw std::runtime_error("host schedule: affine DG cells only");
  if (op.precision == MF_F64)
    host_dg_t<double>(op, mode, static_cast<double*>(dst), static_cast<const double*>(src));
  else
    host_dg_t<float>(op, mode, static_cast<float*>(dst), static_cast<const float*>(src));
}

}  // namespace smf
