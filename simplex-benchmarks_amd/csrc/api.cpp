// C ABI of the operator library (declared in include/simplexmf.h).
#include <hip/hip_runtime_api.h>

#include <cmath>
#include <cstring>
#include <stdexcept>
#include <string>

#include "amg.h"
#include "api_handles.h"
#include "mg.h"
#include "op.h"

namespace smf {

int cg_quadrature_degree(int p);
void launch_invert_diagonal(int precision, int64_t n, void* d, hipStream_t s);

thread_local std::string g_last_error;
void set_error(const std::string& msg) { g_last_error = msg; }

template <typename T>
static std::vector<T> convert(const std::vector<double>& v) {
  return std::vector<T>(v.begin(), v.end());
}

void make_cg_shape(CGPlan& pl) {
  // reference shape gradients at the quadrature points [q][i][d], then w[q]
  const Lagrange fe(pl.p);
  // curved cells: a rule of degree 2p for the non-polynomial integrand
  const Quadrature q = tet_rule(pl.curved ? 2 * pl.p : cg_quadrature_degree(pl.p));
  pl.n_q = q.size();
  pl.quad_pts.resize(3 * pl.n_q);
  for (int k = 0; k < pl.n_q; ++k)
    for (int d = 0; d < 3; ++d) pl.quad_pts[3 * k + d] = q.pts[k][d];
  pl.shape.assign(pl.n_loc * pl.n_q * 3 + pl.n_q, 0.0);
  std::vector<double> grads(pl.n_loc * 3);
  for (int k = 0; k < pl.n_q; ++k) {
    fe.eval(q.pts[k].data(), nullptr, grads.data());
    for (int i = 0; i < pl.n_loc; ++i)
      for (int d = 0; d < 3; ++d) pl.shape[(k * pl.n_loc + i) * 3 + d] = grads[i * 3 + d];
    pl.shape[pl.n_loc * pl.n_q * 3 + k] = q.w[k];
  }
  // curved cells: the kernels also read the points, after the weights
  if (pl.curved) pl.shape.insert(pl.shape.end(), pl.quad_pts.begin(), pl.quad_pts.end());
}

void upload_cg(Operator& op) {
  CGPlan& pl = op.cg;
  op.d_cell_start.upload(pl.cell_start);
  op.d_own_start.upload(pl.own_start);
  op.d_ghost_start.upload(pl.ghost_start);
  op.d_ghost_idx.upload(pl.ghost_idx);
  op.d_ccol.upload(pl.ccol);
  op.d_fix_ptr.upload(pl.fix_ptr);
  op.d_pslot.upload(pl.pslot);
  op.d_gfix.upload(pl.gfix);
  op.d_lidx.upload(pl.lidx);
  op.d_lidx12.upload(pl.lidx12);
  op.d_partial.alloc(std::max<int64_t>(pl.n_ghost_total, 1) * op.ncomp * op.wbytes());
  // curved cells: the kernels read the 27 Jacobian coefficients per cell
  const std::vector<double>& metric = pl.curved ? pl.jgeo : pl.metric;
  if (op.precision == MF_F64) {
    op.d_metric.upload(metric);
    op.d_shape.upload(pl.shape);
  } else {
    op.d_metric.upload(convert<float>(metric));
    op.d_shape.upload(convert<float>(pl.shape));
  }
  if (pl.nitsche) {
    op.d_bnd_rows.upload(pl.bnd_rows);
    op.d_bnd_ptr.upload(pl.bnd_ptr);
    op.d_bnd_col.upload(pl.bnd_col);
    if (op.precision == MF_F64) op.d_bnd_val.upload(pl.bnd_val);
    else op.d_bnd_val.upload(convert<float>(pl.bnd_val));
    op.d_bcell.upload(pl.bcell);
    if (!pl.bmat.empty()) {
      if (op.precision == MF_F64) op.d_bmat.upload(pl.bmat);
      else op.d_bmat.upload(convert<float>(pl.bmat));
    }
    // dense boundary row sums for the fused epilogues: rows without boundary
    // entries stay zero
    op.d_bnd_dense.alloc(std::max<int64_t>(pl.n_dofs, 1) * op.wbytes());
    SMF_HIP(hipMemset(op.d_bnd_dense.p, 0, op.d_bnd_dense.bytes));
  }
  op.on_device = true;
}

void upload_dg(Operator& op) {
  DGPlan& pl = op.dg;
  op.d_dg_nbr.upload(pl.nbr);
  op.d_dg_nface.upload(pl.nface);
  if (!pl.nbrf.empty()) op.d_dg_nbrf.upload(pl.nbrf);
  if (op.precision == MF_F64) {
    op.d_dg_geo.upload(pl.geo);
    if (!pl.fgeo.empty()) op.d_dg_fgeo.upload(pl.fgeo);
    if (!pl.bgeo.empty()) op.d_dg_bgeo.upload(pl.bgeo);
    op.d_dg_cell_tab.upload(pl.cell_tab);
    op.d_dg_face_tab.upload(pl.face_tab);
    if (pl.curved) op.d_dg_jgeo.upload(pl.jgeo);
  } else {
    op.d_dg_geo.upload(convert<float>(pl.geo));
    if (!pl.fgeo.empty()) op.d_dg_fgeo.upload(convert<float>(pl.fgeo));
    if (!pl.bgeo.empty()) op.d_dg_bgeo.upload(convert<float>(pl.bgeo));
    op.d_dg_cell_tab.upload(convert<float>(pl.cell_tab));
    op.d_dg_face_tab.upload(convert<float>(pl.face_tab));
    if (pl.curved) op.d_dg_jgeo.upload(convert<float>(pl.jgeo));
  }
  op.on_device = true;
}

namespace {
// Same schedule as cg_wave_kernel + cg_fixup_kernel, executed serially on the host.
template <typename T>
void host_apply_cg_t(const Operator& op, int mode, T* dst, const T* src) {
  const CGPlan& pl = op.cg;
  const int nl = pl.n_loc, nq = pl.n_q;
  std::vector<T> G(pl.shape.begin(), pl.shape.end());
  const T* W = G.data() + nl * nq * 3;
  const int NCc = op.ncomp;
  for (int cc = 0; cc < NCc; ++cc) {
  const int64_t off = cc * pl.n_dofs;
  T* dstc = dst + off;
  const T* srcc = src ? src + off : nullptr;
  std::vector<T> su, sy, partial(pl.n_ghost_total, T(0));
  for (int64_t patch = 0; patch < pl.n_patches; ++patch) {
    const int c0 = pl.cell_start[patch], nc = pl.cell_start[patch + 1] - c0;
    const int o0 = pl.own_start[patch], no = pl.own_start[patch + 1] - o0;
    const int g0 = pl.ghost_start[patch], ng = pl.ghost_start[patch + 1] - g0;
    su.assign(no + ng, T(0));
    sy.assign(no + ng, T(0));
    if (mode == 0)
      for (int j = 0; j < no + ng; ++j) su[j] = j < no ? srcc[o0 + j] : srcc[pl.ghost_idx[g0 + j - no]];
    for (int lc = 0; lc < nc; ++lc) {
      const int64_t cell = c0 + lc;
      T m[6];
      if (!pl.curved)
        for (int t = 0; t < 6; ++t) m[t] = (T)pl.metric[(size_t)t * pl.n_cells + cell];
      T u[20], y[20];
      uint16_t li[20];
      for (int i = 0; i < nl; ++i) {
        li[i] = pl.lidx[(size_t)i * pl.n_cells + cell];
        u[i] = (mode == 0 && li[i] != 0xFFFF) ? su[li[i]] : T(0);
        y[i] = T(0);
      }
      for (int q = 0; q < nq; ++q) {
        const T* Gq = &G[(size_t)q * nl * 3];
        if (pl.curved)  // w_q |det J_q| J_q^-1 J_q^-T, weight included
          for (int t = 0; t < 6; ++t) m[t] = (T)pl.qmetric[((size_t)q * 6 + t) * pl.n_cells + cell];
        const T wq = pl.curved ? T(1) : W[q];
        if (mode == 0) {
          T g[3] = {0, 0, 0};
          for (int i = 0; i < nl; ++i)
            for (int d = 0; d < 3; ++d) g[d] += Gq[i * 3 + d] * u[i];
          const T tx = wq * (m[0] * g[0] + m[1] * g[1] + m[2] * g[2]);
          const T ty = wq * (m[1] * g[0] + m[3] * g[1] + m[4] * g[2]);
          const T tz = wq * (m[2] * g[0] + m[4] * g[1] + m[5] * g[2]);
          for (int i = 0; i < nl; ++i) y[i] += Gq[i * 3] * tx + Gq[i * 3 + 1] * ty + Gq[i * 3 + 2] * tz;
        } else {
          for (int i = 0; i < nl; ++i) {
            const T* g = &Gq[i * 3];
            y[i] += wq * (g[0] * (m[0] * g[0] + m[1] * g[1] + m[2] * g[2]) +
                            g[1] * (m[1] * g[0] + m[3] * g[1] + m[4] * g[2]) +
                            g[2] * (m[2] * g[0] + m[4] * g[1] + m[5] * g[2]));
          }
        }
      }
      for (int i = 0; i < nl; ++i)
        if (li[i] != 0xFFFF) sy[li[i]] += y[i];
    }
    for (int j = 0; j < no + ng; ++j) {
      if (j < no) dstc[o0 + j] = sy[j];
      else partial[pl.pslot[g0 + j - no]] = sy[j];
    }
  }
  for (int64_t t = 0; t < pl.n_free - pl.n_int; ++t) {
    T s = partial[pl.fix_ptr[t]];
    for (int32_t k = pl.fix_ptr[t] + 1; k < pl.fix_ptr[t + 1]; ++k) s += partial[k];
    dstc[pl.n_int + t] = s;
  }
  for (int64_t i = pl.n_free; i < pl.n_dofs; ++i) dstc[i] = mode == 0 ? srcc[i] : T(1);
  }
  if (pl.nitsche)
    for (size_t r = 0; r < pl.bnd_rows.size(); ++r) {
      const int32_t row = pl.bnd_rows[r];
      T s = T(0);
      for (int32_t k = pl.bnd_ptr[r]; k < pl.bnd_ptr[r + 1]; ++k) {
        if (mode == 0) s += T(pl.bnd_val[k]) * src[pl.bnd_col[k]];
        else if (pl.bnd_col[k] == row) s += T(pl.bnd_val[k]);
      }
      dst[row] += s;
    }
}
}  // namespace

void host_apply_cg(const Operator& op, int mode, void* dst, const void* src) {
  if (op.precision == MF_F64)
    host_apply_cg_t<double>(op, mode, static_cast<double*>(dst), static_cast<const double*>(src));
  else
    host_apply_cg_t<float>(op, mode, static_cast<float*>(dst), static_cast<const float*>(src));
}

}  // namespace smf

using namespace smf;


extern "C" {

int mf_debug_host_apply(mf_op_t h, int32_t mode, void* dst, const void* src) {
  MF_TRY({
    if (!h || !dst || (mode == 0 && !src)) throw std::runtime_error("null argument");
    if (mode != 0 && mode != 1) throw std::runtime_error("mode must be 0 or 1");
    if (h->op.is_dg()) host_apply_dg(h->op, mode, dst, src);
    else host_apply_cg(h->op, mode, dst, src);
  });
}

const char* mf_last_error(void) { return g_last_error.c_str(); }

void mf_default_options(mf_options* opt) {
  std::memset(opt, 0, sizeof(*opt));
  opt->reorder = 1;
  opt->patch_cells = 0;
  opt->mass_coefficient = 1.0;
  opt->laplace_coefficient = 1.0;
  opt->penalty_kind = 0;
  opt->host_only = 0;
  opt->accumulate = 0;
  opt->dg_kernel = 0;
  opt->boundary = 0;
  opt->manifold = 0;
  opt->manifold_center[0] = 5.0;
  opt->manifold_center[1] = 0.0;
  opt->manifold_center[2] = 0.0;
}

int mf_mesh_kuhn_cube(int32_t n, double perturb, uint64_t seed, double* vertices, int32_t* cells) {
  MF_TRY({
    if (!vertices || !cells) throw std::runtime_error("null output array");
    Mesh m = kuhn_cube(n, perturb, seed);
    std::memcpy(vertices, m.x.data(), m.x.size() * sizeof(double));
    std::memcpy(cells, m.cells.data(), m.cells.size() * sizeof(int32_t));
  });
}

int mf_get_quadrature(mf_op_t h, double* points, double* weights) {
  MF_TRY({
    if (!h || !points || !weights) throw std::runtime_error("null argument");
    if (h->op.is_dg()) {
      // cell_tab: gradients [q][i][d], weights [q], values [q][i], points [q][3]
      const DGPlan& pl = h->op.dg;
      const size_t w0 = (size_t)pl.n_q * pl.n_loc * 3;
      const size_t p0 = w0 + pl.n_q + (size_t)pl.n_q * pl.n_loc;
      for (int q = 0; q < pl.n_q; ++q) {
        weights[q] = pl.cell_tab[w0 + q];
        for (int d = 0; d < 3; ++d) points[3 * q + d] = pl.cell_tab[p0 + 3 * q + d];
      }
      return 0;
    }
    const CGPlan& pl = h->op.cg;
    std::memcpy(points, pl.quad_pts.data(), pl.quad_pts.size() * sizeof(double));
    for (int q = 0; q < pl.n_q; ++q) weights[q] = pl.shape[(size_t)pl.n_loc * pl.n_q * 3 + q];
  });
}

int mf_get_face_quadrature(mf_op_t h, int32_t* n_points, double* points, double* weights) {
  MF_TRY({
    if (!h || !n_points) throw std::runtime_error("null argument");
    if (!h->op.is_dg()) throw std::runtime_error("face quadrature query: DG operators only");
    const DGPlan& pl = h->op.dg;
    *n_points = pl.n_qf;
    if (!points || !weights) return 0;
    // face_tab: per face values and gradients (stride n_qf n_loc 4), weights [q], points [f][q][3]
    const size_t w0 = (size_t)4 * pl.n_qf * pl.n_loc * 4;
    const size_t p0 = w0 + pl.n_qf;
    for (int q = 0; q < pl.n_qf; ++q) {
      weights[q] = pl.face_tab[w0 + q];
      // (s, t) of face 3, whose vertices (0, 1, 2) put v_0 + s (v_1 - v_0) + t (v_2 - v_0) at (s, t, 0)
      points[2 * q] = pl.face_tab[p0 + (3 * pl.n_qf + q) * 3];
      points[2 * q + 1] = pl.face_tab[p0 + (3 * pl.n_qf + q) * 3 + 1];
    }
  });
}

int mf_mesh_read_vtk(const char* path, int64_t* n_vertices, int64_t* n_cells, int64_t* n_flipped,
                     double* vertices, int32_t* cells) {
  MF_TRY({
    if (!path || !n_vertices || !n_cells) throw std::runtime_error("null argument");
    if ((vertices == nullptr) != (cells == nullptr)) throw std::runtime_error("give both arrays or neither");
    int64_t flipped = 0;
    Mesh m = read_vtk(path, &flipped);
    if (vertices) {
      if (*n_vertices != m.nv || *n_cells != m.nc)
        throw std::runtime_error("VTK: array sizes differ from the file's (" + std::to_string(m.nv) + " points, " +
                                 std::to_string(m.nc) + " tetrahedra)");
      std::memcpy(vertices, m.x.data(), m.x.size() * sizeof(double));
      std::memcpy(cells, m.cells.data(), m.cells.size() * sizeof(int32_t));
    }
    *n_vertices = m.nv;
    *n_cells = m.nc;
    if (n_flipped) *n_flipped = flipped;
  });
}

int mf_create(const mf_mesh_desc* desc, int32_t degree, int32_t kind, int32_t precision,
              const mf_options* opt, mf_op_t* out) {
  MF_TRY({
    if (!desc || !out) throw std::runtime_error("null argument");
    *out = nullptr;
    if (degree < 1 || degree > 3) throw std::runtime_error("degree must be 1, 2 or 3");
    if (opt && (opt->accumulate < 0 || opt->accumulate > 3)) throw std::runtime_error("accumulate must be 0 to 3");
    if (opt && (opt->boundary < 0 || opt->boundary > 1)) throw std::runtime_error("boundary must be 0 or 1");
    if (opt && (opt->dg_kernel < 0 || opt->dg_kernel > 4)) throw std::runtime_error("dg_kernel must be 0..4");
    if (opt && (opt->penalty_kind < 0 || opt->penalty_kind > 1)) throw std::runtime_error("penalty_kind must be 0 or 1");
    if (opt && (opt->manifold < 0 || opt->manifold > 2)) throw std::runtime_error("manifold must be 0, 1 or 2");
    if (opt && opt->manifold != 0 && opt->boundary != 0)
      throw std::runtime_error("curved cells (manifold != 0): CG operators with the Dirichlet boundary, DG operators");
    if (opt && opt->manifold != 0 && kind != MF_CG_POISSON && kind != MF_CG_VECTOR && opt->penalty_kind != 0)
      throw std::runtime_error("curved DG cells take the reference penalty (penalty_kind 0)");
    if (opt && opt->boundary == 1 && kind != MF_CG_POISSON)
      throw std::runtime_error("Nitsche boundary (boundary = 1) is available for the scalar CG operator");
    if (precision != MF_F64 && precision != MF_F32) throw std::runtime_error("precision must be MF_F64 or MF_F32");
    if (kind != MF_CG_POISSON && kind != MF_CG_VECTOR && kind != MF_DG_SIPG && kind != MF_DG_VECTOR &&
        kind != MF_DG_HELMHOLTZ)
      throw std::runtime_error("operator kind not available in this build");
    if (desc->n_vertices <= 0 || desc->n_cells < 0 || !desc->vertices || (!desc->cells && desc->n_cells))
      throw std::runtime_error("invalid mesh description");
    auto h = std::make_unique<mf_operator>();
    Operator& op = h->op;
    op.kind = kind;
    op.precision = precision;
    op.degree = degree;
    if (opt) op.opt = *opt; else mf_default_options(&op.opt);
    op.mesh.nv = desc->n_vertices;
    op.mesh.nc = desc->n_cells;
    op.mesh.x.assign(desc->vertices, desc->vertices + 3 * desc->n_vertices);
    op.mesh.cells.assign(desc->cells, desc->cells + 4 * desc->n_cells);
    for (int32_t v : op.mesh.cells)
      if (v < 0 || v >= desc->n_vertices) throw std::runtime_error("cell references a missing vertex");
    op.mesh.kuhn_n = desc->kuhn_n;
    if (desc->kuhn_n > 0) {
      const int64_t n = desc->kuhn_n;
      const int64_t n1 = n + 1;
      if (desc->n_vertices != n1 * n1 * n1 || desc->n_cells != 6 * n * n * n)
        throw std::runtime_error("kuhn_n does not match the mesh sizes");
      op.mesh.lattice.resize(3 * op.mesh.nv);
      for (int64_t v = 0; v < op.mesh.nv; ++v) {
        op.mesh.lattice[3 * v] = (int32_t)(v % n1);
        op.mesh.lattice[3 * v + 1] = (int32_t)((v / n1) % n1);
        op.mesh.lattice[3 * v + 2] = (int32_t)(v / (n1 * n1));
      }
    }
    if (op.is_dg()) {
      op.ncomp = kind == MF_DG_SIPG ? 1 : 3;
      build_dg_plan(op.mesh, degree, op.ncomp, op.opt, kind == MF_DG_HELMHOLTZ, op.dg);
      if (!op.opt.host_only) upload_dg(op);
    } else {
      op.ncomp = kind == MF_CG_VECTOR ? 3 : 1;
      op.cg.ncomp = op.ncomp;
      op.cg.value_bytes = op.precision == MF_F64 ? 8 : 4;
      build_cg_plan(op.mesh, degree, op.opt, op.cg);
      make_cg_shape(op.cg);
      if (!op.opt.host_only) upload_cg(op);
    }
    if (!op.opt.host_only) SMF_HIP(hipDeviceSynchronize());
    *out = h.release();
  });
}

int mf_destroy(mf_op_t op) {
  MF_TRY({ delete op; });
}

int mf_vmult(mf_op_t h, void* dst, const void* src, void* stream) {
  MF_TRY({
    if (!h) throw std::runtime_error("null argument");
    if (dst && dst == src) throw std::runtime_error("dst and src must not alias");
    require_device(h->op);
    if (h->op.m() == 0) return 0;  // empty mesh: nothing to do (vectors may be null)
    if (!dst || !src) throw std::runtime_error("null argument");
    if (h->op.is_dg()) launch_dg(h->op, 0, dst, src, static_cast<hipStream_t>(stream));
    else launch_cg(h->op, 0, dst, src, static_cast<hipStream_t>(stream));
  });
}

int mf_tvmult(mf_op_t h, void* dst, const void* src, void* stream) {
  return mf_vmult(h, dst, src, stream);
}

int mf_compute_diagonal(mf_op_t h, void* diag, void* stream) {
  MF_TRY({
    if (!h) throw std::runtime_error("null argument");
    require_device(h->op);
    if (h->op.m() == 0) return 0;
    if (!diag) throw std::runtime_error("null argument");
    if (h->op.is_dg()) launch_dg(h->op, 1, diag, nullptr, static_cast<hipStream_t>(stream));
    else launch_cg(h->op, 1, diag, nullptr, static_cast<hipStream_t>(stream));
  });
}

int mf_compute_inverse_diagonal(mf_op_t h, void* inv, void* stream) {
  MF_TRY({
    if (!h) throw std::runtime_error("null argument");
    require_device(h->op);
    if (h->op.m() == 0) return 0;
    if (!inv) throw std::runtime_error("null argument");
    const hipStream_t s = static_cast<hipStream_t>(stream);
    if (h->op.is_dg()) launch_dg(h->op, 1, inv, nullptr, s);
    else launch_cg(h->op, 1, inv, nullptr, s);
    launch_invert_diagonal(h->op.precision, h->op.m(), inv, s);
  });
}

int64_t mf_m(mf_op_t h) { return h ? h->op.m() : -1; }

int mf_get_info(mf_op_t h, mf_info* info) {
  MF_TRY({
    if (!h || !info) throw std::runtime_error("null argument");
    const Operator& op = h->op;
    std::memset(info, 0, sizeof(*info));
    const double w = (double)op.wbytes();
    info->n_dofs = op.m();
    info->degree = op.degree;
    info->n_components = op.ncomp;
    info->precision = op.precision;
    if (op.is_dg()) {
      const DGPlan& pl = op.dg;
      info->n_cells = pl.n_cells;
      info->dofs_per_cell = pl.n_loc;
      info->n_q_points = pl.n_q;
      info->n_interior_faces = pl.n_interior_faces;
      info->n_boundary_faces = pl.n_boundary_faces;
      // reference data model: src + dst; per cell J^-T, det J, tau + start index;
      // per interior face nJ-, nJ+, face det J + 2 cell ids; per boundary face nJ, det J + id
      info->algorithmic_bytes = 2.0 * w * pl.n_dofs + pl.n_cells * (11.0 * w + 4) +
                                pl.n_interior_faces * (7.0 * w + 8) + pl.n_boundary_faces * (4.0 * w + 4);
      info->flops = pl.ncomp * pl.n_cells *
                    ((12.0 * pl.n_loc * pl.n_q + 18.0 * pl.n_q) + 4.0 * pl.n_qf * 24.0 * pl.n_loc);
    } else {
      const CGPlan& pl = op.cg;
      info->n_cells = pl.n_cells;
      info->n_constrained = pl.n_dofs - pl.n_free;
      info->dofs_per_cell = pl.n_loc;
      info->n_q_points = pl.n_q;
      info->n_patches = pl.n_patches;
      info->max_cell_colors = pl.max_icolors;
      info->max_local_dofs = pl.max_local;
      // reference data model: src + dst, 32-bit index per cell DoF, J^-T + det J per cell
      // (curved cells: J^-T + JxW per quadrature point)
      info->algorithmic_bytes = 2.0 * w * pl.n_dofs * op.ncomp + 4.0 * pl.n_loc * pl.n_cells +
                                10.0 * w * pl.n_cells * (pl.curved ? pl.n_q : 1);
      info->flops = op.ncomp * (12.0 * pl.n_loc * pl.n_q + 18.0 * pl.n_q) * pl.n_cells;
    }
  });
}

int mf_get_cell_order(mf_op_t h, int64_t* perm) {
  MF_TRY({
    if (!h || !perm) throw std::runtime_error("null argument");
    const auto& v = h->op.is_dg() ? h->op.dg.cell_perm : h->op.cg.cell_perm;
    std::memcpy(perm, v.data(), v.size() * sizeof(int64_t));
  });
}

int mf_get_cell_vertices(mf_op_t h, int32_t* verts) {
  MF_TRY({
    if (!h || !verts) throw std::runtime_error("null argument");
    const Operator& op = h->op;
    if (op.is_dg()) {
      std::memcpy(verts, op.dg.cell_verts.data(), op.dg.cell_verts.size() * sizeof(int32_t));
    } else {
      for (int64_t k = 0; k < op.cg.n_cells; ++k)
        for (int t = 0; t < 4; ++t) verts[4 * k + t] = op.mesh.cells[4 * op.cg.cell_perm[k] + t];
    }
  });
}

int mf_get_cell_dofs(mf_op_t h, int64_t* dofs) {
  MF_TRY({
    if (!h || !dofs) throw std::runtime_error("null argument");
    const Operator& op = h->op;
    if (op.is_dg()) {
      const int64_t n = op.dg.n_dofs;
      for (int64_t r = 0; r < n; ++r) dofs[r] = r;
    } else {
      std::memcpy(dofs, op.cg.cell_dofs.data(), op.cg.cell_dofs.size() * sizeof(int64_t));
    }
  });
}

int mf_get_dof_points(mf_op_t h, double* xyz) {
  MF_TRY({
    if (!h || !xyz) throw std::runtime_error("null argument");
    const auto& v = h->op.is_dg() ? h->op.dg.dof_points : h->op.cg.dof_points;
    std::memcpy(xyz, v.data(), v.size() * sizeof(double));
  });
}

int mf_get_constrained(mf_op_t h, uint8_t* mask) {
  MF_TRY({
    if (!h || !mask) throw std::runtime_error("null argument");
    const Operator& op = h->op;
    if (op.is_dg()) {
      std::memset(mask, 0, op.dg.n_cells * op.dg.n_loc);
    } else {
      const CGPlan& pl = op.cg;
      for (int64_t i = 0; i < pl.n_dofs; ++i) mask[i] = i >= pl.n_free ? 1 : 0;
    }
  });
}


void mf_mg_default_options(mf_mg_options* opt) {
  MGOptions d;
  opt->smoother_degree = d.smoother_degree;
  opt->smoothing_range = d.smoothing_range;
  opt->eig_cg_iterations = d.eig_cg_iterations;
  opt->level_precision = d.level_precision;
  opt->max_coarse_dofs = d.max_coarse_dofs;
  opt->fuse_steps = d.fuse_steps;
  opt->coarse_solver = d.coarse_solver;
  opt->amg_threshold = d.amg_threshold;
}

int mf_mg_create(mf_op_t fine, const mf_mg_options* opt, mf_mg_t* out) {
  MF_TRY({
    if (!fine || !out) throw std::runtime_error("null argument");
    *out = nullptr;
    require_device(fine->op);
    MGOptions o;
    if (opt) {
      o.smoother_degree = opt->smoother_degree;
      o.smoothing_range = opt->smoothing_range;
      o.eig_cg_iterations = opt->eig_cg_iterations;
      o.level_precision = opt->level_precision;
      o.max_coarse_dofs = opt->max_coarse_dofs;
      o.fuse_steps = opt->fuse_steps;
      o.coarse_solver = opt->coarse_solver;
      if (o.coarse_solver < 0 || o.coarse_solver > 2) throw std::runtime_error("coarse_solver must be 0, 1 or 2");
      o.amg_threshold = opt->amg_threshold;
      if (!(o.amg_threshold >= 0.0 && o.amg_threshold < 1.0)) throw std::runtime_error("amg_threshold must be in [0, 1)");
    }
    if (o.smoother_degree < 1 || o.smoothing_range <= 1 || o.eig_cg_iterations < 1)
      throw std::runtime_error("invalid multigrid options");
    if (o.level_precision != MF_F64 && o.level_precision != MF_F32) throw std::runtime_error("invalid level precision");
    auto h = std::make_unique<mf_multigrid>();
    h->mg = build_multigrid(fine->op, o);
    *out = h.release();
  });
}

int mf_mg_vmult(mf_mg_t mg, void* dst, const void* src, void* stream) {
  MF_TRY({
    if (!mg || !dst || !src) throw std::runtime_error("null argument");
    if (dst == src) throw std::runtime_error("dst and src must not alias");
    mg_vmult(*mg->mg, dst, src, static_cast<hipStream_t>(stream));
  });
}

int mf_mg_get_info(mf_mg_t mg, mf_mg_info* info) {
  MF_TRY({
    if (!mg || !info) throw std::runtime_error("null argument");
    std::memset(info, 0, sizeof(*info));
    const auto& L = mg->mg->levels;
    info->n_levels = (int32_t)std::min<size_t>(L.size(), MF_MAX_LEVELS);
    for (int l = 0; l < info->n_levels; ++l) {
      info->degree[l] = L[l].degree;
      info->kuhn_n[l] = L[l].kuhn_n;
      info->n_dofs[l] = L[l].op->m();
      info->lambda_max[l] = L[l].lambda_max;
      info->level_kind[l] = L[l].op->kind;
      info->nitsche[l] = !L[l].op->is_dg() && L[l].op->cg.nitsche;
    }
    info->vcycles = mg->mg->vcycles;
    if (mg->mg->amg) {
      const auto& lv = mg->mg->amg->lv;
      info->amg_levels = (int32_t)std::min<size_t>(lv.size(), MF_MAX_LEVELS);
      for (int l = 0; l < info->amg_levels; ++l) info->amg_rows[l] = lv[l].A.nrows;
    }
  });
}

int mf_mg_transfer(mf_mg_t mg, int32_t level, int32_t direction, void* dst, const void* src, void* stream) {
  MF_TRY({
    if (!mg || !dst || !src) throw std::runtime_error("null argument");
    auto& L = mg->mg->levels;
    if (level < 1 || level >= (int)L.size()) throw std::runtime_error("level out of range");
    if (direction != 0 && direction != 1) throw std::runtime_error("direction must be 0 or 1");
    const hipStream_t s = static_cast<hipStream_t>(stream);
    if (mg->mg->opt.level_precision == MF_F64)
      level_transfer<double>(L[level], direction, static_cast<double*>(dst), static_cast<const double*>(src), false,
                             s);
    else
      level_transfer<float>(L[level], direction, static_cast<float*>(dst), static_cast<const float*>(src), false, s);
  });
}

int mf_mg_level_dofs(mf_mg_t mg, int32_t level, double* xyz, uint8_t* constrained) {
  MF_TRY({
    if (!mg) throw std::runtime_error("null argument");
    auto& L = mg->mg->levels;
    if (level < 0 || level >= (int)L.size()) throw std::runtime_error("level out of range");
    const Operator& op = *L[level].op;
    if (op.is_dg()) {
      if (xyz) std::memcpy(xyz, op.dg.dof_points.data(), op.dg.dof_points.size() * sizeof(double));
      if (constrained) std::memset(constrained, 0, op.dg.n_dofs);
    } else {
      const CGPlan& pl = op.cg;
      if (xyz) std::memcpy(xyz, pl.dof_points.data(), pl.dof_points.size() * sizeof(double));
      if (constrained)
        for (int64_t i = 0; i < pl.n_dofs; ++i) constrained[i] = i >= pl.n_free ? 1 : 0;
    }
  });
}

int mf_mg_destroy(mf_mg_t mg) {
  MF_TRY({ delete mg; });
}

int mf_solve_cg(mf_op_t A, mf_mg_t mg, void* x, const void* b, double rel_tol, int32_t max_iter, int32_t flexible,
                mf_solve_stats* stats, void* stream) {
  MF_TRY({
    if (!A || !x || !b) throw std::runtime_error("null argument");
    require_device(A->op);
    if (A->op.kind != MF_CG_POISSON && A->op.kind != MF_DG_SIPG)
      throw std::runtime_error("solver needs a scalar CG Poisson or DG SIPG operator");
    if (mg && mg->mg->n_fine != A->op.m()) throw std::runtime_error("multigrid does not match the operator");
    SolveStats st = solve_cg(A->op, mg ? mg->mg.get() : nullptr, static_cast<double*>(x),
                             static_cast<const double*>(b), rel_tol, max_iter, flexible != 0,
                             static_cast<hipStream_t>(stream));
    if (stats) {
      stats->iterations = st.iterations;
      stats->converged = st.converged;
      stats->initial_residual = st.r0;
      stats->final_residual = st.rn;
      stats->rhs_norm = st.b_norm;
    }
  });
}

int mf_assemble_rhs(mf_op_t op, int32_t kind, uint64_t seed, int32_t n_features, double corr_length, double* b) {
  MF_TRY({
    if (!op || !b) throw std::runtime_error("null argument");
    require_device(op->op);
    if (kind != 0 && kind != 1) throw std::runtime_error("rhs kind must be 0 or 1");
    if (kind == 1 && (n_features < 1 || corr_length <= 0)) throw std::runtime_error("invalid random field parameters");
    std::vector<double> v;
    assemble_rhs(op->op, kind, seed, n_features, corr_length, v);
    std::memcpy(b, v.data(), v.size() * sizeof(double));
  });
}

int mf_exact_solution(mf_op_t op, double* u) {
  MF_TRY({
    if (!op || !u) throw std::runtime_error("null argument");
    if (op->op.kind != MF_CG_POISSON && op->op.kind != MF_DG_SIPG) throw std::runtime_error("scalar CG and DG operators only");
    std::vector<double> v;
    exact_solution(op->op, v);
    std::memcpy(u, v.data(), v.size() * sizeof(double));
  });
}

}  // extern "C"
