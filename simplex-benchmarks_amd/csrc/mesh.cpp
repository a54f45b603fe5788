#include "mesh.h"

#include <algorithm>
#include <cmath>
#include <stdexcept>

#include "fe.h"

namespace smf {

Mesh kuhn_cube(int n, double perturb, uint64_t seed) {
  if (n < 1) throw std::runtime_error("kuhn_cube: n must be >= 1");
  Mesh m;
  const int64_t n1 = n + 1;
  m.nv = n1 * n1 * n1;
  m.nc = 6ll * n * n * n;
  m.kuhn_n = n;
  m.x.resize(3 * m.nv);
  m.lattice.resize(3 * m.nv);
  const double h = 1.0 / n;
  SplitMix64 rng(seed);
  for (int64_t k = 0; k < n1; ++k)
    for (int64_t j = 0; j < n1; ++j)
      for (int64_t i = 0; i < n1; ++i) {
        const int64_t v = i + n1 * (j + n1 * k);
        const int64_t ijk[3] = {i, j, k};
        const bool interior = i > 0 && i < n && j > 0 && j < n && k > 0 && k < n;
        for (int d = 0; d < 3; ++d) {
          m.lattice[3 * v + d] = (int32_t)ijk[d];
          double xd = ijk[d] * h;
          if (interior && perturb != 0.0) xd += perturb * h * rng.uniform_pm1();
          m.x[3 * v + d] = xd;
        }
      }
  static const int kPerm[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
  static const bool kOdd[6] = {false, true, true, false, false, true};
  m.cells.resize(4 * m.nc);
  for (int64_t k = 0; k < n; ++k)
    for (int64_t j = 0; j < n; ++j)
      for (int64_t i = 0; i < n; ++i) {
        const int64_t hex = i + n * (j + n * k);
        for (int s = 0; s < 6; ++s) {
          int64_t p[3] = {i, j, k};
          int32_t vid[4];
          vid[0] = (int32_t)(p[0] + n1 * (p[1] + n1 * p[2]));
          for (int t = 0; t < 3; ++t) {
            p[kPerm[s][t]] += 1;
            vid[t + 1] = (int32_t)(p[0] + n1 * (p[1] + n1 * p[2]));
          }
          if (kOdd[s]) std::swap(vid[2], vid[3]);
          for (int t = 0; t < 4; ++t) m.cells[4 * (6 * hex + s) + t] = vid[t];
        }
      }
  return m;
}

FaceInfo build_faces(const Mesh& m) {
  FaceInfo fi;
  const int64_t nc = m.nc, nv = m.nv;
  // bucket the 4*nc cell faces by their smallest vertex
  std::vector<int64_t> cnt(nv + 1, 0);
  auto sorted_face = [&](int64_t c, int f, int32_t out[3]) {
    for (int k = 0; k < 3; ++k) out[k] = m.cells[4 * c + kFaceVerts[f][k]];
    std::sort(out, out + 3);
  };
  for (int64_t c = 0; c < nc; ++c)
    for (int f = 0; f < 4; ++f) {
      int32_t t[3];
      sorted_face(c, f, t);
      cnt[t[0] + 1]++;
    }
  for (int64_t v = 0; v < nv; ++v) cnt[v + 1] += cnt[v];
  struct Entry {
    int32_t b, c;
    int64_t cf;
  };
  std::vector<Entry> ent(4 * nc);
  {
    std::vector<int64_t> pos(cnt.begin(), cnt.end() - 1);
    for (int64_t c = 0; c < nc; ++c)
      for (int f = 0; f < 4; ++f) {
        int32_t t[3];
        sorted_face(c, f, t);
        ent[pos[t[0]]++] = Entry{t[1], t[2], 4 * c + f};
      }
  }
  fi.cell_face.assign(4 * nc, -1);
  for (int64_t v = 0; v < nv; ++v) {
    auto b = ent.begin() + cnt[v], e = ent.begin() + cnt[v + 1];
    std::sort(b, e, [](const Entry& x, const Entry& y) {
      return x.b != y.b ? x.b < y.b : (x.c != y.c ? x.c < y.c : x.cf < y.cf);
    });
    for (auto it = b; it != e;) {
      auto jt = it + 1;
      while (jt != e && jt->b == it->b && jt->c == it->c) ++jt;
      const int64_t mult = jt - it;
      if (mult > 2) throw std::runtime_error("non-manifold mesh: face shared by >2 cells");
      const int64_t id = fi.n_faces++;
      std::array<int64_t, 2> fc{it->cf / 4, -1};
      std::array<int8_t, 2> fl{(int8_t)(it->cf % 4), -1};
      if (mult == 2) {
        fc[1] = (it + 1)->cf / 4;
        fl[1] = (int8_t)((it + 1)->cf % 4);
      } else {
        fi.n_boundary++;
      }
      fi.face_cell.push_back(fc);
      fi.face_lf.push_back(fl);
      for (auto kt = it; kt != jt; ++kt) fi.cell_face[kt->cf] = id;
      it = jt;
    }
  }
  return fi;
}

EdgeInfo build_edges(const Mesh& m) {
  EdgeInfo ei;
  const int64_t nc = m.nc, nv = m.nv;
  std::vector<int64_t> cnt(nv + 1, 0);
  for (int64_t c = 0; c < nc; ++c)
    for (int e = 0; e < 6; ++e) {
      int32_t a = m.cells[4 * c + kEdgeVerts[e][0]], b = m.cells[4 * c + kEdgeVerts[e][1]];
      cnt[std::min(a, b) + 1]++;
    }
  for (int64_t v = 0; v < nv; ++v) cnt[v + 1] += cnt[v];
  std::vector<std::pair<int32_t, int64_t>> ent(6 * nc);
  {
    std::vector<int64_t> pos(cnt.begin(), cnt.end() - 1);
    for (int64_t c = 0; c < nc; ++c)
      for (int e = 0; e < 6; ++e) {
        int32_t a = m.cells[4 * c + kEdgeVerts[e][0]], b = m.cells[4 * c + kEdgeVerts[e][1]];
        ent[pos[std::min(a, b)]++] = {std::max(a, b), 6 * c + e};
      }
  }
  ei.cell_edge.assign(6 * nc, -1);
  for (int64_t v = 0; v < nv; ++v) {
    auto b = ent.begin() + cnt[v], e = ent.begin() + cnt[v + 1];
    std::sort(b, e);
    for (auto it = b; it != e;) {
      auto jt = it + 1;
      while (jt != e && jt->first == it->first) ++jt;
      const int64_t id = ei.n_edges++;
      ei.verts.push_back({(int32_t)v, it->first});
      for (auto kt = it; kt != jt; ++kt) ei.cell_edge[kt->second] = id;
      it = jt;
    }
  }
  return ei;
}

namespace {
void rcb(const std::vector<double>& cen, int64_t* idx, int64_t n, int max_patch,
         std::vector<int64_t>& leaves, int64_t offset) {
  if (n <= max_patch) {
    leaves.push_back(offset + n);
    return;
  }
  double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
  for (int64_t i = 0; i < n; ++i)
    for (int d = 0; d < 3; ++d) {
      lo[d] = std::min(lo[d], cen[3 * idx[i] + d]);
      hi[d] = std::max(hi[d], cen[3 * idx[i] + d]);
    }
  int ax = 0;
  for (int d = 1; d < 3; ++d)
    if (hi[d] - lo[d] > hi[ax] - lo[ax]) ax = d;
  // split into halves whose sizes are as equal as possible (ties broken by id
  // so the result does not depend on the std::nth_element implementation)
  const int64_t half = n / 2;
  std::nth_element(idx, idx + half, idx + n, [&](int64_t a, int64_t b) {
    const double ka = cen[3 * a + ax], kb = cen[3 * b + ax];
    return ka != kb ? ka < kb : a < b;
  });
  rcb(cen, idx, half, max_patch, leaves, offset);
  rcb(cen, idx + half, n - half, max_patch, leaves, offset + half);
}
}  // namespace

void reorder_cells(const Mesh& m, bool reorder, int max_patch, std::vector<int64_t>& perm,
                   std::vector<int64_t>& patch_ptr) {
  const int64_t nc = m.nc;
  perm.resize(nc);
  for (int64_t c = 0; c < nc; ++c) perm[c] = c;
  patch_ptr.assign(1, 0);
  if (nc == 0) return;
  if (!reorder) {
    for (int64_t s = max_patch; s < nc; s += max_patch) patch_ptr.push_back(s);
    patch_ptr.push_back(nc);
    return;
  }
  std::vector<double> cen(3 * nc);
  for (int64_t c = 0; c < nc; ++c)
    for (int d = 0; d < 3; ++d) {
      double s = 0;
      for (int k = 0; k < 4; ++k) s += m.x[3 * m.cells[4 * c + k] + d];
      cen[3 * c + d] = 0.25 * s;
    }
  std::vector<int64_t> leaves;
  rcb(cen, perm.data(), nc, max_patch, leaves, 0);
  for (int64_t e : leaves) patch_ptr.push_back(e);
}

CellGeom cell_geometry(const Mesh& m, int64_t c) {
  CellGeom g;
  const int32_t* v = &m.cells[4 * c];
  for (int d = 0; d < 3; ++d)
    for (int k = 0; k < 3; ++k) g.J[d][k] = m.x[3 * v[k + 1] + d] - m.x[3 * v[0] + d];
  const auto& J = g.J;
  g.det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
          J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
          J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
  const double id = 1.0 / g.det;
  g.Jinv[0][0] = (J[1][1] * J[2][2] - J[1][2] * J[2][1]) * id;
  g.Jinv[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) * id;
  g.Jinv[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) * id;
  g.Jinv[1][0] = (J[1][2] * J[2][0] - J[1][0] * J[2][2]) * id;
  g.Jinv[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * id;
  g.Jinv[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * id;
  g.Jinv[2][0] = (J[1][0] * J[2][1] - J[1][1] * J[2][0]) * id;
  g.Jinv[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * id;
  g.Jinv[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * id;
  return g;
}

}  // namespace smf
