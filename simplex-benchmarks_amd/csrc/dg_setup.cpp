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
  plan.penalty_factor = (p + 1.0) * (p + 3.0) / 3.0;
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
// The kernel's per-cell algorithm on the host (double or float).
template <typename T>
void host_dg_t(const Operator& op, int mode, T* dst, const T* src) {
  const DGPlan& pl = op.dg;
  const int nl = pl.n_loc, nq = pl.n_q, nqf = pl.n_qf, NC = pl.ncomp;
  const int64_t nc = pl.n_cells;
  std::vector<T> ct(pl.cell_tab.begin(), pl.cell_tab.end()), ft(pl.face_tab.begin(), pl.face_tab.end());
  const T* CG = ct.data();
  const T* CW = ct.data() + (size_t)nq * nl * 3;
  const T* CV = CW + nq;
  const size_t fstride = (size_t)nqf * nl * 4;
  const T* FW = ft.data() + 4 * fstride;
  const T factor = (T)pl.penalty_factor, mass = (T)pl.mass, lap = (T)pl.laplace;
  auto geo = [&](int t, int64_t k) { return (T)pl.geo[(size_t)t * nc + k]; };
  for (int64_t k = 0; k < nc; ++k) {
    T jit[9];
    for (int t = 0; t < 9; ++t) jit[t] = geo(t, k);
    const T adet = geo(9, k), tau = geo(10, k);
    T M[6];
    {
      const int ab[6][2] = {{0, 0}, {0, 1}, {0, 2}, {1, 1}, {1, 2}, {2, 2}};
      for (int t = 0; t < 6; ++t) {
        T s = 0;
        for (int e = 0; e < 3; ++e) s += jit[3 * e + ab[t][0]] * jit[3 * e + ab[t][1]];
        M[t] = adet * s;
      }
    }
    for (int c = 0; c < NC; ++c) {
      const int64_t base = (k * NC + c) * nl;
      T u[20], y[20];
      for (int i = 0; i < nl; ++i) {
        u[i] = mode == 0 ? src[base + i] : T(0);
        y[i] = T(0);
      }
      // cell term
      for (int q = 0; q < nq; ++q) {
        const T* G = CG + (size_t)q * nl * 3;
        const T* V = CV + (size_t)q * nl;
        if (mode == 0) {
          T g[3] = {0, 0, 0}, uv = 0;
          for (int i = 0; i < nl; ++i) {
            for (int d = 0; d < 3; ++d) g[d] += G[3 * i + d] * u[i];
            uv += V[i] * u[i];
          }
          const T w = CW[q];
          const T tx = w * lap * (M[0] * g[0] + M[1] * g[1] + M[2] * g[2]);
          const T ty = w * lap * (M[1] * g[0] + M[3] * g[1] + M[4] * g[2]);
          const T tz = w * lap * (M[2] * g[0] + M[4] * g[1] + M[5] * g[2]);
          const T tv = w * mass * adet * uv;
          for (int i = 0; i < nl; ++i) y[i] += G[3 * i] * tx + G[3 * i + 1] * ty + G[3 * i + 2] * tz + V[i] * tv;
        } else {
          for (int i = 0; i < nl; ++i) {
            const T* g = &G[3 * i];
            y[i] += CW[q] * (lap * (g[0] * (M[0] * g[0] + M[1] * g[1] + M[2] * g[2]) +
                                    g[1] * (M[1] * g[0] + M[3] * g[1] + M[4] * g[2]) +
                                    g[2] * (M[2] * g[0] + M[4] * g[1] + M[5] * g[2])) +
                             mass * adet * V[i] * V[i]);
          }
        }
      }
      // faces
      for (int f = 0; f < 4; ++f) {
        const T* FV = ft.data() + f * fstride;
        const T* FG = FV + (size_t)nqf * nl;
        T nj[3], nn = 0;
        for (int a = 0; a < 3; ++a) {
          nj[a] = jit[3 * a] * (T)kNhat[f][0] + jit[3 * a + 1] * (T)kNhat[f][1] + jit[3 * a + 2] * (T)kNhat[f][2];
          nn += nj[a] * nj[a];
        }
        const T len = std::sqrt(nn);
        T n[3] = {nj[0] / len, nj[1] / len, nj[2] / len};
        const T area2 = adet * len;  // JxW_q = w_q * area2
        // nJ = J^-1 n = jit^T n
        T nJ[3];
        for (int a = 0; a < 3; ++a) nJ[a] = jit[a] * n[0] + jit[3 + a] * n[1] + jit[6 + a] * n[2];
        const int32_t kb = pl.nbr[(size_t)f * nc + k];
        T nJb[3] = {0, 0, 0}, ub[20];
        T sigma = factor * tau;
        const T* FVb = nullptr;
        const T* FGb = nullptr;
        if (kb >= 0) {
          const int fb = (pl.nface[k] >> (8 * f)) & 3;
          FVb = ft.data() + fb * fstride;
          FGb = FVb + (size_t)nqf * nl;
          T jb[9];
          for (int t = 0; t < 9; ++t) jb[t] = geo(t, kb);
          for (int a = 0; a < 3; ++a) nJb[a] = jb[a] * n[0] + jb[3 + a] * n[1] + jb[6 + a] * n[2];
          sigma = factor * std::max(tau, geo(10, kb));
          for (int i = 0; i < nl; ++i) ub[i] = mode == 0 ? src[(kb * NC + c) * nl + i] : T(0);
        }
        for (int q = 0; q < nqf; ++q) {
          const T* V = FV + (size_t)q * nl;
          const T* G = FG + (size_t)q * nl * 3;
          const T jxw = FW[q] * area2 * lap;
          if (mode == 0) {
            T uv = 0, dn = 0;
            for (int i = 0; i < nl; ++i) {
              uv += V[i] * u[i];
              dn += (G[3 * i] * nJ[0] + G[3 * i + 1] * nJ[1] + G[3 * i + 2] * nJ[2]) * u[i];
            }
            T up, dnp;
            if (kb >= 0) {
              up = 0;
              dnp = 0;
              const T* Vb = FVb + (size_t)q * nl;
              const T* Gb = FGb + (size_t)q * nl * 3;
              for (int i = 0; i < nl; ++i) {
                up += Vb[i] * ub[i];
                dnp += (Gb[3 * i] * nJb[0] + Gb[3 * i + 1] * nJb[1] + Gb[3 * i + 2] * nJb[2]) * ub[i];
              }
            } else {
              up = -uv;
              dnp = dn;
            }
            const T jump = uv - up, avg = T(0.5) * (dn + dnp);
            const T tv = (sigma * jump - avg) * jxw, tg = T(-0.5) * jump * jxw;
            for (int i = 0; i < nl; ++i)
              y[i] += V[i] * tv + (G[3 * i] * nJ[0] + G[3 * i + 1] * nJ[1] + G[3 * i + 2] * nJ[2]) * tg;
          } else {
            const T s = kb >= 0 ? T(1) : T(2);
            for (int i = 0; i < nl; ++i) {
              const T vi = V[i];
              const T di = G[3 * i] * nJ[0] + G[3 * i + 1] * nJ[1] + G[3 * i + 2] * nJ[2];
              const T jump = s * vi, avg = T(0.5) * s * di;
              y[i] += jxw * ((sigma * jump - avg) * vi - T(0.5) * jump * di);
            }
          }
        }
      }
      for (int i = 0; i < nl; ++i) dst[(k * NC + c) * nl + i] = y[i];
    }
  }
}
}  // namespace

void host_apply_dg(const Operator& op, int mode, void* dst, const void* src) {
  if (op.precision == MF_F64)
    host_dg_t<double>(op, mode, static_cast<double*>(dst), static_cast<const double*>(src));
  else
    host_dg_t<float>(op, mode, static_cast<float*>(dst), static_cast<const float*>(src));
}

}  // namespace smf
