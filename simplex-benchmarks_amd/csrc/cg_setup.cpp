// Continuous-Galerkin setup: DoF numbering, Dirichlet constraints, patch and
// cell colourings, patch-local index tables and per-cell geometry.
//
// Counterpart of Operator::reinit (CG/cells_batched/application.cc:1056-1128),
// which builds manual_dof_indices (global index per lane, invalid for
// constrained DoFs) on top of deal.II's DoFHandler + MatrixFree.  Here the
// table is patch-local: each patch of cells (from the hierarchical reordering)
// stages its DoFs in LDS, so a cell's index is a 16-bit offset into its
// patch's DoF list and constrained DoFs carry the invalid marker 0xFFFF, as
// in the reference's masked gather (application.cc:1287-1292).
#include <algorithm>
#include <cstring>
#include <stdexcept>

#include "op.h"

namespace smf {

namespace {

// local edge index of the pair of local vertices (a,b)
int local_edge(int a, int b) {
  if (a > b) std::swap(a, b);
  for (int e = 0; e < 6; ++e)
    if (kEdgeVerts[e][0] == a && kEdgeVerts[e][1] == b) return e;
  return -1;
}

}  // namespace

void build_cg_plan(const Mesh& mesh, int p, const mf_options& opt, CGPlan& plan) {
  const Lagrange fe(p);
  const int nl = fe.n_loc;
  plan.p = p;
  plan.n_loc = nl;
  plan.n_cells = mesh.nc;
  const int64_t nc = mesh.nc, nv = mesh.nv;

  // cells per wave by degree (P1 1024, P2 512, P3 128); a patch is one
  // workgroup of up to 4 waves.  The default patch is 4 waves' worth of
  // cells, halved (down to 64 cells) until there are >= 1024 patches, so
  // that small operators (multigrid levels) still fill the 256 CUs.  (FP64
  // P2 with per-wave accumulation: 4 x 512 0.312 ms, 2 x 512 0.322 ms at
  // 128^3.)  The 3-component operator, whose LDS is 3x larger, takes patches
  // of the scalar wave size split over two waves (P2 FP64 64^3: 0.181 ->
  // 0.167 ms against one wave of 512)
  const int per_wave = opt.cells_per_wave > 0
                           ? opt.cells_per_wave
                           : (p == 1 ? 1024 : (p == 2 ? 512 : 128)) / (plan.ncomp == 3 ? 2 : 1);
  int patch = opt.patch_cells;
  int waves_default = 0;
  if (patch <= 0) {
    waves_default = plan.ncomp == 3 ? 2 : 4;
    patch = waves_default * per_wave;
    while (patch > 64 && mesh.nc / patch < 1024) patch /= 2;
  }
  if (patch > 4096) throw std::runtime_error("patch_cells must be <= 4096");
  plan.waves = std::min(4, std::max(1, (patch + per_wave - 1) / per_wave));
  // a halved default patch keeps up to two waves of at least 256 cells each
  // (48^3 multigrid levels: P2 FP32 26.4 -> 25.6 us, P1 15.5 -> 15.0 us)
  if (waves_default > 0 && opt.cells_per_wave <= 0 && patch < waves_default * per_wave)
    plan.waves = std::max(1, std::min(std::min(waves_default, 2), patch / 256));
  plan.max_patch_cells = 0;

  // ---- entities and raw DoFs (entity-based numbering) ----
  FaceInfo fi = build_faces(mesh);
  EdgeInfo ei;
  if (p >= 2) ei = build_edges(mesh);
  const int64_t n_raw = nv + ei.n_edges * (p - 1) + (p == 3 ? fi.n_faces : 0);
  std::vector<uint8_t> raw_con(n_raw, 0);
  plan.nitsche = opt.boundary == 1;
  for (int64_t f = 0; f < fi.n_faces && !plan.nitsche; ++f) {
    if (fi.face_cell[f][1] >= 0) continue;
    const int64_t c = fi.face_cell[f][0];
    const int lf = fi.face_lf[f][0];
    for (int k = 0; k < 3; ++k) raw_con[mesh.cells[4 * c + kFaceVerts[lf][k]]] = 1;
    if (p >= 2)
      for (int k = 0; k < 3; ++k) {
        const int e = local_edge(kFaceVerts[lf][k], kFaceVerts[lf][(k + 1) % 3]);
        const int64_t ge = ei.cell_edge[6 * c + e];
        for (int s = 0; s < p - 1; ++s) raw_con[nv + ge * (p - 1) + s] = 1;
      }
    if (p == 3) raw_con[nv + ei.n_edges * 2 + f] = 1;
  }
  // raw DoF of every (input cell, local node)
  auto raw_of = [&](int64_t c, int i) -> int64_t {
    const LocalNode& n = fe.nodes[i];
    if (n.kind == 0) return mesh.cells[4 * c + n.entity];
    if (n.kind == 1) {
      const int32_t ga = mesh.cells[4 * c + kEdgeVerts[n.entity][0]];
      const int32_t gb = mesh.cells[4 * c + kEdgeVerts[n.entity][1]];
      const int s = ga < gb ? n.sub : p - 2 - n.sub;
      return nv + ei.cell_edge[6 * c + n.entity] * (p - 1) + s;
    }
    return nv + ei.n_edges * (p - 1) + fi.cell_face[4 * c + n.entity];
  };

  // ---- hierarchical patches ----
  plan.cell_dofs.clear();
  std::vector<int64_t> perm, patch_ptr;
  reorder_cells(mesh, opt.reorder != 0, patch, perm, patch_ptr);
  const int64_t np = (int64_t)patch_ptr.size() - 1;
  plan.n_patches = np;

  // ---- intra-patch cell colouring (cells sharing a vertex conflict), cells
  //      of a patch sorted by colour ----
  plan.ccol.assign(nc, 0);
  {
    std::vector<uint64_t> vmask(2 * nv, 0);
    std::vector<int64_t> newperm(nc);
    for (int64_t q = 0; q < np; ++q) {
      const int64_t b = patch_ptr[q], e = patch_ptr[q + 1];
      std::vector<int> col(e - b);
      int ncol = 0;
      for (int64_t k = b; k < e; ++k) {
        const int32_t* v = &mesh.cells[4 * perm[k]];
        uint64_t u0 = 0, u1 = 0;
        for (int t = 0; t < 4; ++t) {
          u0 |= vmask[2 * v[t]];
          u1 |= vmask[2 * v[t] + 1];
        }
        int c;
        if (~u0) c = __builtin_ctzll(~u0);
        else if (~u1) c = 64 + __builtin_ctzll(~u1);
        else throw std::runtime_error("cell colouring needs more than 128 colours");
        if (c > 254) throw std::runtime_error("cell colouring needs more than 255 colours");
        col[k - b] = c;
        ncol = std::max(ncol, c + 1);
        for (int t = 0; t < 4; ++t) vmask[2 * v[t] + (c >> 6)] |= 1ull << (c & 63);
      }
      for (int64_t k = b; k < e; ++k)
        for (int t = 0; t < 4; ++t) {
          const int32_t v = mesh.cells[4 * perm[k] + t];
          vmask[2 * v] = vmask[2 * v + 1] = 0;
        }
      // stable counting sort by colour
      std::vector<int> cnt(ncol + 1, 0);
      for (int c : col) cnt[c + 1]++;
      for (int c = 0; c < ncol; ++c) cnt[c + 1] += cnt[c];
      std::vector<int> pos(cnt.begin(), cnt.end() - 1);
      for (int64_t k = b; k < e; ++k) {
        const int64_t dstk = b + pos[col[k - b]]++;
        newperm[dstk] = perm[k];
        plan.ccol[dstk] = (uint8_t)col[k - b];
      }
      plan.max_icolors = std::max(plan.max_icolors, ncol);
      plan.max_patch_cells = std::max<int>(plan.max_patch_cells, (int)(e - b));
    }
    perm.swap(newperm);
  }
  plan.cell_perm = perm;

  // ---- interior vs interface DoFs: a free DoF touched by one patch only is
  //      interior to it; touched by several it is an interface DoF ----
  std::vector<int32_t> first_patch(n_raw, -1);
  std::vector<uint8_t> multi(n_raw, 0);
  for (int64_t q = 0; q < np; ++q)
    for (int64_t k = patch_ptr[q]; k < patch_ptr[q + 1]; ++k)
      for (int i = 0; i < nl; ++i) {
        const int64_t r = raw_of(perm[k], i);
        if (raw_con[r]) continue;
        if (first_patch[r] < 0) first_patch[r] = (int32_t)q;
        else if (first_patch[r] != q) multi[r] = 1;
      }

  // ---- numbering: interior DoFs patch by patch (contiguous per patch), then
  //      interface DoFs, then constrained DoFs, each in first-touch order ----
  std::vector<int64_t> number(n_raw, -1);
  int64_t next = 0;
  plan.own_start.assign(1, 0);
  for (int64_t q = 0; q < np; ++q) {
    for (int64_t k = patch_ptr[q]; k < patch_ptr[q + 1]; ++k)
      for (int i = 0; i < nl; ++i) {
        const int64_t r = raw_of(perm[k], i);
        if (!raw_con[r] && !multi[r] && number[r] < 0) number[r] = next++;
      }
    plan.own_start.push_back((int32_t)next);
  }
  plan.n_int = next;
  for (int64_t k = 0; k < nc; ++k)
    for (int i = 0; i < nl; ++i) {
      const int64_t r = raw_of(perm[k], i);
      if (!raw_con[r] && number[r] < 0) number[r] = next++;
    }
  plan.n_free = next;
  for (int64_t k = 0; k < nc; ++k)
    for (int i = 0; i < nl; ++i) {
      const int64_t r = raw_of(perm[k], i);
      if (number[r] < 0) number[r] = next++;
    }
  plan.n_dofs = next;  // entities no cell touches (unused vertices) get no DoF
  if (plan.n_dofs > INT32_MAX) throw std::runtime_error("too many DoFs for 32-bit indices");

  // ---- global cell table, support points ----
  plan.cell_dofs.resize(nc * nl);
  plan.dof_points.assign(3 * plan.n_dofs, 0.0);
  for (int64_t k = 0; k < nc; ++k) {
    const int64_t c = perm[k];
    const CellGeom g = cell_geometry(mesh, c);
    const double* x0 = &mesh.x[3 * mesh.cells[4 * c]];
    for (int i = 0; i < nl; ++i) {
      const int64_t d = number[raw_of(c, i)];
      plan.cell_dofs[k * nl + i] = d;
      const auto xr = fe.node_point(i);
      for (int a = 0; a < 3; ++a) {
        double s = x0[a];
        for (int b = 0; b < 3; ++b) s += g.J[a][b] * xr[b];
        plan.dof_points[3 * d + a] = s;
      }
    }
  }

  // ---- patch-local tables: [interior DoFs of the patch | interface DoFs it
  //      touches]; cell indices are 16-bit offsets into that list ----
  plan.cell_start.resize(np + 1);
  for (int64_t q = 0; q <= np; ++q) plan.cell_start[q] = (int32_t)patch_ptr[q];
  plan.ghost_start.assign(1, 0);
  plan.lidx.assign((size_t)nl * nc, 0xFFFF);
  {
    std::vector<int32_t> stamp(plan.n_dofs, -1), loc(plan.n_dofs, 0);
    for (int64_t q = 0; q < np; ++q) {
      const int32_t o0 = plan.own_start[q], o1 = plan.own_start[q + 1];
      int32_t nloc_p = o1 - o0;
      for (int64_t k = patch_ptr[q]; k < patch_ptr[q + 1]; ++k)
        for (int i = 0; i < nl; ++i) {
          const int64_t d = plan.cell_dofs[k * nl + i];
          if (d >= plan.n_free) continue;  // constrained: 0xFFFF
          int32_t li;
          if (d < plan.n_int) {
            if (d < o0 || d >= o1) throw std::runtime_error("interior DoF outside its patch range");
            li = (int32_t)(d - o0);
          } else if (stamp[d] == q) {
            li = loc[d];
          } else {
            stamp[d] = (int32_t)q;
            li = loc[d] = nloc_p++;
            plan.ghost_idx.push_back((int32_t)d);
          }
          if (li >= 0xFFFF) throw std::runtime_error("patch has too many local DoFs");
          plan.lidx[(size_t)i * nc + k] = (uint16_t)li;
        }
      plan.ghost_start.push_back((int32_t)plan.ghost_idx.size());
      plan.max_local = std::max(plan.max_local, nloc_p);
    }
  }
  plan.n_ghost_total = (int64_t)plan.ghost_idx.size();

  // ---- interface fix-up: interface DoF n_int + t is the sum of its patch
  //      partials, slots k in [fix_ptr[t], fix_ptr[t+1]) in patch order;
  //      ghost entry g writes its partial to slot pslot[g] ----
  {
    const int64_t ng = plan.n_ghost_total, nif = plan.n_free - plan.n_int;
    std::vector<int32_t> cnt(nif + 1, 0);
    for (int64_t k = 0; k < ng; ++k) cnt[plan.ghost_idx[k] - plan.n_int + 1]++;
    for (int64_t t = 0; t < nif; ++t) {
      if (cnt[t + 1] < 2) throw std::runtime_error("interface DoF with fewer than two patches");
      cnt[t + 1] += cnt[t];
    }
    plan.fix_ptr.assign(cnt.begin(), cnt.end());
    plan.fix_src.resize(ng);
    std::vector<int32_t> pos(cnt.begin(), cnt.end() - 1);
    for (int64_t k = 0; k < ng; ++k) plan.fix_src[pos[plan.ghost_idx[k] - plan.n_int]++] = (int32_t)k;
    plan.pslot.resize(ng);
    for (int64_t k = 0; k < ng; ++k) plan.pslot[plan.fix_src[k]] = (int32_t)k;
    plan.gfix.resize(ng);
    for (int64_t g = 0; g < ng; ++g) {
      const int64_t t = plan.ghost_idx[g] - plan.n_int;
      const int32_t pos = plan.pslot[g] - plan.fix_ptr[t], len = plan.fix_ptr[t + 1] - plan.fix_ptr[t];
      if (len > 0xFFFF) throw std::runtime_error("interface DoF shared by too many patches");
      plan.gfix[g] = (pos << 16) | len;
    }
  }

  // ---- geometry ----
  plan.metric.resize(6 * nc);
  plan.jit.resize(9 * nc);
  plan.detj.resize(nc);
  for (int64_t k = 0; k < nc; ++k) {
    const CellGeom g = cell_geometry(mesh, perm[k]);
    if (!(g.det > 0)) throw std::runtime_error("cell with non-positive Jacobian determinant");
    // M = |det J| J^-1 J^-T ; (J^-1 J^-T)_{ab} = sum_e Jinv[a][e] Jinv[b][e]
    const int ab[6][2] = {{0, 0}, {0, 1}, {0, 2}, {1, 1}, {1, 2}, {2, 2}};
    for (int t = 0; t < 6; ++t) {
      double s = 0;
      for (int e = 0; e < 3; ++e) s += g.Jinv[ab[t][0]][e] * g.Jinv[ab[t][1]][e];
      plan.metric[(size_t)t * nc + k] = g.det * s;
    }
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) plan.jit[(size_t)(3 * a + b) * nc + k] = g.Jinv[b][a];
    plan.detj[k] = g.det;
  }
  if (plan.nitsche) build_nitsche_boundary(mesh, plan, opt.penalty_kind);
  if (opt.manifold != 0) build_curved_geometry(mesh, opt, plan);

  // ---- 12-bit packed local indices (fewer and smaller loads per cell) ----
  plan.lidx12.clear();
  plan.lidx12_words = 0;
  // the cell's colour takes the 8 bits after the indices (P2: bits 120-127,
  // P3: 240-247), so the kernels need no separate colour load
  if (p >= 2 && plan.max_local < 0xFFF) {
    const int nw = (nl * 12 + 31) / 32;
    const int cb = nl * 12, cw = cb >> 5, cs = cb & 31;
    if (cs > 24) throw std::runtime_error("packed local indices: no room for the colour");
    plan.lidx12.assign((size_t)nw * nc, 0u);
    for (int64_t k = 0; k < nc; ++k) {
      for (int i = 0; i < nl; ++i) {
        const uint16_t li = plan.lidx[(size_t)i * nc + k];
        const uint64_t v = li == 0xFFFF ? 0xFFFu : li;
        const int b = 12 * i, w = b >> 5, s = b & 31;
        plan.lidx12[(size_t)w * nc + k] |= (uint32_t)(v << s);
        if (s > 20) plan.lidx12[(size_t)(w + 1) * nc + k] |= (uint32_t)(v >> (32 - s));
      }
      plan.lidx12[(size_t)cw * nc + k] |= (uint32_t)plan.ccol[k] << cs;
    }
    // set-up self-check: the packed form decodes to the unpacked one
    for (int64_t k = 0; k < nc; ++k)
      for (int i = 0; i < nl; ++i) {
        const int b = 12 * i, w = b >> 5, s = b & 31;
        uint32_t v = plan.lidx12[(size_t)w * nc + k] >> s;
        if (s > 20) v |= plan.lidx12[(size_t)(w + 1) * nc + k] << (32 - s);
        v &= 0xFFFu;
        const uint16_t li = plan.lidx[(size_t)i * nc + k];
        if (v != (li == 0xFFFF ? 0xFFFu : li)) throw std::runtime_error("packed local indices do not decode");
      }
    for (int64_t k = 0; k < nc; ++k)
      if (((plan.lidx12[(size_t)cw * nc + k] >> cs) & 0xFFu) != plan.ccol[k])
        throw std::runtime_error("packed cell colour does not decode");
    plan.lidx12_words = nw;
  }
}

namespace {
const double kNref[4][3] = {{1, 1, 1}, {-1, 0, 0}, {0, -1, 0}, {0, 0, -1}};  // outward, |.| = 2 area / ref area
}

std::vector<double> cell_tau(const Mesh& mesh) {
  const FaceInfo fi = build_faces(mesh);
  std::vector<double> surf(mesh.nc, 0.0);
  for (int64_t f = 0; f < fi.n_faces; ++f) {
    const bool bdry = fi.face_cell[f][1] < 0;
    for (int side = 0; side < (bdry ? 1 : 2); ++side) {
      const int64_t c = fi.face_cell[f][side];
      const int lf = fi.face_lf[f][side];
      const CellGeom g = cell_geometry(mesh, c);
      double nn = 0;
      for (int a = 0; a < 3; ++a) {
        double v = 0;
        for (int b = 0; b < 3; ++b) v += g.Jinv[b][a] * kNref[lf][b];
        nn += v * v;
      }
      surf[c] += (bdry ? 1.0 : 0.5) * 0.5 * std::fabs(g.det) * std::sqrt(nn);
    }
  }
  std::vector<double> tau(mesh.nc);
  for (int64_t c = 0; c < mesh.nc; ++c) tau[c] = surf[c] / (std::fabs(cell_geometry(mesh, c).det) / 6.0);
  return tau;
}

// B_ij = sum over boundary faces F of the cell of
//   int_F 2 sigma phi_j phi_i - (n.grad phi_j) phi_i - phi_j (n.grad phi_i),
// sigma = (p+1)(p+3)/3 tau_K (or (p+1)(p+2)|F|/|K|): the boundary part of the SIPG form for
// continuous functions (DG/cells_batched/application.cc:2688-2701 with the
// mirror principle; as the CG level of application_phc.cc:1798-1832).
void build_nitsche_boundary(const Mesh& mesh, CGPlan& plan, int penalty_kind) {
  const int p = plan.p, nl = plan.n_loc;
  const Lagrange fe(p);
  const Quadrature qf = tri_rule(2 * p);
  double wsum = 0;
  for (double w : qf.w) wsum += w;
  const std::vector<double> tau = cell_tau(mesh);
  const double factor = (p + 1.0) * (p + 3.0) / 3.0;
  std::vector<int64_t> inv(plan.n_cells);
  for (int64_t k = 0; k < plan.n_cells; ++k) inv[plan.cell_perm[k]] = k;
  const FaceInfo fi = build_faces(mesh);
  struct Trip {
    int32_t r, c;
    double v;
  };
  std::vector<Trip> trips;
  std::vector<double> val(nl), grad(3 * nl), dn(nl);
  plan.bcell.assign(plan.n_cells, -1);
  plan.bmat.clear();
  for (int64_t f = 0; f < fi.n_faces; ++f) {
    if (fi.face_cell[f][1] >= 0) continue;
    const int64_t c = fi.face_cell[f][0];
    const int lf = fi.face_lf[f][0];
    const CellGeom g = cell_geometry(mesh, c);
    double n[3], nn = 0;
    for (int a = 0; a < 3; ++a) {
      n[a] = 0;
      for (int b = 0; b < 3; ++b) n[a] += g.Jinv[b][a] * kNref[lf][b];
      nn += n[a] * n[a];
    }
    const double area = 0.5 * std::fabs(g.det) * std::sqrt(nn);
    for (int a = 0; a < 3; ++a) n[a] /= std::sqrt(nn);
    // penalty_kind 1: (p+1)(p+2) |F| / |K|
    const double sigma = penalty_kind == 1 ? (p + 1.0) * (p + 2.0) * area / (std::fabs(g.det) / 6.0) : factor * tau[c];
    const int64_t k = inv[c];
    std::vector<double> B(nl * nl, 0.0);
    for (int q = 0; q < qf.size(); ++q) {
      const auto x = face_to_cell(lf, qf.pts[q][0], qf.pts[q][1]);
      fe.eval(x.data(), val.data(), grad.data());
      for (int i = 0; i < nl; ++i) {
        double s = 0;
        for (int a = 0; a < 3; ++a) {
          double ga = 0;  // physical gradient component a = sum_b Jinv[b][a] dphi/dxhat_b
          for (int b = 0; b < 3; ++b) ga += g.Jinv[b][a] * grad[3 * i + b];
          s += n[a] * ga;
        }
        dn[i] = s;
      }
      const double ds = qf.w[q] * area / wsum;
      for (int i = 0; i < nl; ++i)
        for (int j = 0; j < nl; ++j)
          B[i * nl + j] += ds * (2.0 * sigma * val[j] * val[i] - dn[j] * val[i] - val[j] * dn[i]);
    }
    for (int i = 0; i < nl; ++i)
      for (int j = 0; j < nl; ++j)
        trips.push_back({(int32_t)plan.cell_dofs[k * nl + i], (int32_t)plan.cell_dofs[k * nl + j], B[i * nl + j]});
    if (plan.bcell[k] < 0) {
      plan.bcell[k] = (int32_t)(plan.bmat.size() / ((size_t)nl * nl));
      plan.bmat.resize(plan.bmat.size() + (size_t)nl * nl, 0.0);
    }
    double* Bk = &plan.bmat[(size_t)plan.bcell[k] * nl * nl];
    for (int e = 0; e < nl * nl; ++e) Bk[e] += B[e];
  }
  // stable: contributions to one entry are summed in boundary-face order
  std::stable_sort(trips.begin(), trips.end(),
                   [](const Trip& a, const Trip& b) { return a.r < b.r || (a.r == b.r && a.c < b.c); });
  plan.bnd_rows.clear();
  plan.bnd_ptr.assign(1, 0);
  plan.bnd_col.clear();
  plan.bnd_val.clear();
  for (size_t t = 0; t < trips.size(); ++t) {
    if (!plan.bnd_col.empty() && plan.bnd_rows.back() == trips[t].r && plan.bnd_col.back() == trips[t].c) {
      plan.bnd_val.back() += trips[t].v;
      continue;
    }
    if (plan.bnd_rows.empty() || plan.bnd_rows.back() != trips[t].r) {
      if (!plan.bnd_rows.empty()) plan.bnd_ptr.push_back((int32_t)plan.bnd_col.size());
      plan.bnd_rows.push_back(trips[t].r);
    }
    plan.bnd_col.push_back(trips[t].c);
    plan.bnd_val.push_back(trips[t].v);
  }
  if (!plan.bnd_rows.empty()) plan.bnd_ptr.push_back((int32_t)plan.bnd_col.size());
}

}  // namespace smf
