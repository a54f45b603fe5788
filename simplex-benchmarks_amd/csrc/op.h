// Internal operator state shared by the setup code (C++) and the HIP launchers.
#pragma once
#include <hip/hip_runtime_api.h>

#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "../../include/simplexmf.h"
#include "fe.h"
#include "mesh.h"

namespace smf {

// Thread-local error message returned by mf_last_error().
void set_error(const std::string& msg);

#define SMF_HIP(call)                                                          \
  do {                                                                         \
    hipError_t e_ = (call);                                                    \
    if (e_ != hipSuccess)                                                      \
      throw std::runtime_error(std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

// Device allocation owned by an operator (freed in its destructor).
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p(o.p), bytes(o.bytes) {
    o.p = nullptr;
    o.bytes = 0;
  }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) {
      release();
      p = o.p;
      bytes = o.bytes;
      o.p = nullptr;
      o.bytes = 0;
    }
    return *this;
  }
  ~DevBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  template <typename T>
  T* as() const { return static_cast<T*>(p); }
  template <typename T>
  void upload(const std::vector<T>& v) {
    release();
    bytes = v.size() * sizeof(T);
    if (bytes == 0) return;
    SMF_HIP(hipMalloc(&p, bytes));
    SMF_HIP(hipMemcpy(p, v.data(), bytes, hipMemcpyHostToDevice));
  }
  void alloc(size_t b) {
    release();
    bytes = b;
    if (b) SMF_HIP(hipMalloc(&p, b));
  }
};

// Host-side plan of a continuous-Galerkin operator: DoF numbering, patches,
// colourings, local index tables and geometry (all in internal cell order).
struct CGPlan {
  int p = 0, n_loc = 0, n_q = 0, ncomp = 1;
  int value_bytes = 8;                // 8 (FP64) or 4 (FP32): sets the default patch shape
  // scalar DoFs (per component): [0,n_int) patch-interior, [n_int,n_free)
  // patch-interface, [n_free,n_dofs) Dirichlet-constrained
  int64_t n_cells = 0, n_dofs = 0, n_free = 0, n_int = 0;
  std::vector<int64_t> cell_perm;   // internal cell -> input cell
  std::vector<int64_t> cell_dofs;   // n_cells*n_loc global DoF per local node
  std::vector<double> dof_points;   // 3*n_dofs support points
  // patches: cells [cell_start[p], cell_start[p+1]); owned DoFs
  // [own_start[p], own_start[p+1]); ghost DoFs ghost_idx[ghost_start[p]..)
  int64_t n_patches = 0;
  int max_icolors = 0, max_local = 0, max_patch_cells = 0;
  int waves = 1;                     // wavefronts per patch workgroup (1..4)
  std::vector<int32_t> cell_start, own_start, ghost_start, ghost_idx;
  std::vector<uint8_t> ccol;         // per cell: colour within its patch (cells sorted by colour)
  // interface fix-up: DoF n_int+t = sum of the partials k in [fix_ptr[t],
  // fix_ptr[t+1]) (patch order); ghost entry g (patch order) writes its
  // partial to slot pslot[g], fix_src is the inverse map
  std::vector<int32_t> fix_ptr, fix_src, pslot;
  // per ghost entry: (position of its patch in the DoF's partial list) << 16 | list length
  std::vector<int32_t> gfix;
  std::vector<uint16_t> lidx;        // n_loc x n_cells patch-local index (0xFFFF = constrained)
  // the same packed to 12 bits (P >= 2, every patch below 4095 local DoFs;
  // 0xFFF = constrained): lidx12[w][n_cells], index i in bits 12 i .. 12 i + 11
  // of the little-endian word sequence, then the cell colour (ccol) in the
  // next 8 bits; lidx12_words = 0 when not packed
  std::vector<uint32_t> lidx12;
  int lidx12_words = 0;
  std::vector<double> metric;        // 6 x n_cells: |det J| J^-1 J^-T (xx xy xz yy yz zz)
  std::vector<double> jit;           // 9 x n_cells: J^-T row-major
  std::vector<double> detj;          // n_cells
  int64_t n_ghost_total = 0;
  std::vector<double> shape;         // G (n_loc x n_q x 3) then w_q G
  std::vector<double> quad_pts;      // 3 x n_q: the rule of shape (unit tetrahedron)
  // Nitsche boundary (weak Dirichlet): sparse rows of the boundary-face terms,
  // row bnd_rows[r] has entries [bnd_ptr[r], bnd_ptr[r+1]) (columns ascending)
  bool nitsche = false;
  std::vector<int32_t> bnd_rows, bnd_ptr, bnd_col;
  std::vector<double> bnd_val;
  // the same boundary terms per cell (Chebyshev runs add them inside the
  // cell kernel): bcell[k] = index of cell k's dense n_loc x n_loc matrix
  // in bmat (row-major, y_i += sum_j B_ij u_j), -1 without boundary faces
  std::vector<int32_t> bcell;
  std::vector<double> bmat;
  // curved cells (P2 geometry map): qmetric[(q*6 + t)*n_cells + k] =
  // w_q |det J_q| (J_q^-1 J_q^-T)_t at the points of the degree-2p rule,
  // which shape then holds; metric keeps the straight-cell values
  bool curved = false;
  std::vector<double> qmetric;
  // the same geometry in the kernels' form: J(xi) = C + sum_f xi_f D_f is
  // affine in xi for a P2 map; jgeo[t][n_cells], t < 27: C columns e = 0..2
  // (t = 3e + d), then E_e = D_e column e (9 + 3e + d), then F_ef = D_f
  // column e = D_e column f for (e, f) = (0,1), (0,2), (1,2) (18 + 3i + d)
  std::vector<double> jgeo;
};

// Host-side plan of a discontinuous-Galerkin (SIPG) operator.  Cell local
// vertices are ordered by ascending global id, so the two cells of a face
// parametrise it identically (no orientation tables).  DoFs are
// cell-contiguous: cell k, component c, node i -> (k*ncomp + c)*n_loc + i.
struct DGPlan {
  int p = 0, n_loc = 0, n_q = 0, n_qf = 0, ncomp = 1;
  int64_t n_cells = 0, n_dofs = 0;  // n_dofs counts all components
  int64_t n_interior_faces = 0, n_boundary_faces = 0;
  std::vector<int64_t> cell_perm;   // internal cell -> input cell
  std::vector<int32_t> cell_verts;  // 4*n_cells, ascending global ids (internal order)
  std::vector<double> geo;          // 11 x n_cells: J^-T (row-major 9), |det J|, tau
  std::vector<double> fgeo;         // [n_cells][4][4] face records of the four-lane kernel (dg_cell.h)
  std::vector<double> bgeo;         // [7][n_cells] barycentric metric of the four-lane kernel (dg_cell.h)
  std::vector<int32_t> nbrf;        // [n_cells][4] neighbour words of the four-lane kernel (dg_cell.h)
  std::vector<int32_t> nbr;         // 4 x n_cells: neighbour across local face f, -1 on the boundary
  std::vector<uint32_t> nface;      // per cell: neighbour's local face for f in bits 8f..8f+1
  std::vector<double> dof_points;   // 3 * n_cells * n_loc
  // tables: cell gradients [q][i][d] + w[q], cell values [q][i]; per face f:
  // values [f][q][i], gradients [f][q][i][d], weights [q]
  std::vector<double> cell_tab, face_tab;
  double penalty_factor = 0, mass = 0, laplace = 1;
  // 0: sigma = (p+1)(p+3)/3 max(tau-, tau+) (reference); 1: sigma = (p+1)(p+2)
  // max(|F|/|K-|, |F|/|K+|) = penalty_factor max(len, len |det J-|/|det J+|)
  // with len = |J^-T nhat_f| (|F|/|K| = 3 len), penalty_factor = 3 (p+1)(p+2)
  int penalty_kind = 0;
  // curved cells (opt.manifold != 0): the P2 geometry map of each cell as the
  // 27 coefficients of J(xi) (27 x n_cells); tau (geo row 10) from the
  // curved volumes and face areas by the operator's rules; the tables carry
  // the rule points: cell_tab ends with the cell points [q][3], face_tab with
  // the face points in cell coordinates [f][q][3]
  bool curved = false;
  std::vector<double> jgeo;
};

#ifdef __HIPCC__
// Workgroups are dispatched round-robin over the 8 XCDs of an MI355X, each
// with its own L2.  This bijection gives XCD x the contiguous logical range
// [x q + min(x, r), ...) of the nb blocks (q = nb / 8, r = nb % 8), so
// neighbouring cells/patches, which the hierarchical order keeps close in
// index, share one L2.
__device__ inline int64_t xcd_block(int64_t b, int64_t nb) {
  constexpr int kXcd = 8;
  const int64_t q = nb / kXcd, r = nb % kXcd, x = b % kXcd, i = b / kXcd;
  return x * q + (x < r ? x : r) + i;
}
#endif

struct Operator {
  int kind = MF_CG_POISSON, precision = MF_F64, degree = 1, ncomp = 1;
  mf_options opt{};
  Mesh mesh;            // kept for transfers and RHS assembly
  CGPlan cg;
  DGPlan dg;
  bool is_dg() const { return kind == MF_DG_SIPG || kind == MF_DG_VECTOR || kind == MF_DG_HELMHOLTZ; }
  // device data
  DevBuf d_cell_start, d_own_start, d_ghost_start, d_ghost_idx, d_ccol, d_fix_ptr,
      d_pslot, d_gfix, d_lidx, d_lidx12, d_metric, d_shape, d_partial;
  DevBuf d_bnd_rows, d_bnd_ptr, d_bnd_col, d_bnd_val, d_bnd_dense, d_bcell, d_bmat;
  DevBuf d_dg_geo, d_dg_fgeo, d_dg_bgeo, d_dg_nbrf, d_dg_nbr, d_dg_nface, d_dg_cell_tab, d_dg_face_tab, d_dg_jgeo;
  bool on_device = false;
  int64_t m() const { return is_dg() ? dg.n_dofs : cg.n_dofs * ncomp; }
  size_t wbytes() const { return precision == MF_F64 ? 8 : 4; }
};

// What the operator kernels do with each row of A x (scalar operators,
// MODE 0): kind 0 store it; kind 1 one Chebyshev step (d = c1 d + c2 Dinv
// (b - Ax), or d = c2 Dinv (b - Ax) when first; then x += d in place, x being
// the vmult's input, or xout = x + d when xout is set); kind 2 the residual
// r = b - A x.  bnd: Nitsche boundary row sums B x (dense), added to A x.
// In place is safe for the CG kernels (a row's epilogue runs after every read
// of that DoF); the DG kernels read neighbour cells' values while other cells
// finish, so they write xout.
template <typename T>
struct RowEpilogue {
  int kind = 0, first = 0;
  T c1 = T(0), c2 = T(0);
  const T* b = nullptr;
  const T* dinv = nullptr;
  T* d = nullptr;
  T* x = nullptr;
  T* xout = nullptr;
  T* r = nullptr;
  const T* bnd = nullptr;
};
// A x with the epilogue (kind 1 or 2) applied row by row inside the kernels
template <typename T>
void launch_cg_epilogue(const Operator& op, const T* x, RowEpilogue<T> ep, hipStream_t s);

// A run of Chebyshev steps on a scalar CG level without a fix-up launch per
// step.  Step k's kernel leaves its interface rows pending: their partial
// sums go to part[k % 2], and the next step's kernel applies the pending
// update to every interface DoF it stages (each patch computes the same
// value in the same order; the first patch of the DoF's partial list also
// stores it to xi[k % 2] / di[k % 2] for the step after).  Dirichlet rows
// take no part in the products of free rows, so their updates wait for the
// closing kernel, which also completes the last step's interface rows.
template <typename T>
struct ChebyStep {
  T c1, c2;
  int first;
};
constexpr int kMaxChebySteps = 16;
template <typename T>
struct ChebyRun {
  const T* b = nullptr;
  const T* dinv = nullptr;
  T* x = nullptr;  // iterate (interior rows updated in place by each step)
  T* d = nullptr;  // direction
  T* part[2] = {nullptr, nullptr};  // partial sums, fix-up order (n_ghost_total each)
  T* xi[2] = {nullptr, nullptr};    // interface iterate / direction of pending steps (n_iface each)
  T* di[2] = {nullptr, nullptr};
  int n = 0;                         // steps issued so far
  ChebyStep<T> steps[kMaxChebySteps];
};
// whether the operator's kernels take part in a ChebyRun (scalar CG; a
// Nitsche boundary is applied per cell inside the kernel during a run)
bool cg_has_cheby_run(const Operator& op);
// one Chebyshev step (c1, c2, first) of the run on x
template <typename T>
void launch_cg_cheby_step(const Operator& op, ChebyRun<T>& run, T c1, T c2, int first, hipStream_t s);
// completes the run: pending interface rows of the last step, Dirichlet rows
// of every step
template <typename T>
void launch_cg_cheby_close(const Operator& op, ChebyRun<T>& run, hipStream_t s);
// the same for a scalar DG SIPG operator (four-lane kernel; ep.xout required
// for kind 1); dg_has_epilogue says whether the operator's kernel has it
bool dg_has_epilogue(const Operator& op);
template <typename T>
void launch_dg_epilogue(const Operator& op, const T* x, RowEpilogue<T> ep, hipStream_t s);

void build_cg_plan(const Mesh& mesh, int p, const mf_options& opt, CGPlan& plan);
// curved-cell geometry (opt.manifold != 0): qmetric and the mapped support points
void build_curved_geometry(const Mesh& mesh, const mf_options& opt, CGPlan& plan);
// P2 geometry nodes of a cell: its vertices xv[0..3], then per edge (kEdgeVerts
// order) the manifold's midpoint (manifold 1 spherical around center, 2 straight)
void p2_geometry_nodes(const double* const xv[4], int manifold, const double* center, double X[10][3]);
// the 27 coefficients of J(xi) = C + sum_f xi_f D_f of the P2 map with nodes X,
// in the kernels' layout: C (3 columns), D_e column e, the shared columns of D_e, D_f
void p2_jacobian_coeffs(const double X[10][3], double jc[27]);
// J(xi) from those coefficients
void p2_jacobian_at(const double jc[27], const double* xi, double J[3][3]);
// boundary-face matrix of the Nitsche CG operator (plan.nitsche)
void build_nitsche_boundary(const Mesh& mesh, CGPlan& plan, int penalty_kind);
// SIPG penalty tau_K = (sum_bdry |F| + 1/2 sum_int |F|) / |K| per input cell
std::vector<double> cell_tau(const Mesh& mesh);
// any operator: dispatch to the CG or DG launcher
void launch_op(const Operator& op, int mode, void* dst, const void* src, hipStream_t s);
void make_cg_shape(CGPlan& plan);
// host execution of the kernels' schedule (self-check of the tables)
void host_apply_cg(const Operator& op, int mode, void* dst, const void* src);
void upload_cg(Operator& op);
// mode 0: dst = A src; mode 1: dst = diag(A)
void launch_cg(const Operator& op, int mode, void* dst, const void* src, hipStream_t s);

void build_dg_plan(const Mesh& mesh, int p, int ncomp, const mf_options& opt, bool helmholtz,
                   DGPlan& plan);
void upload_dg(Operator& op);
void launch_dg(const Operator& op, int mode, void* dst, const void* src, hipStream_t s);
void host_apply_dg(const Operator& op, int mode, void* dst, const void* src);

}  // namespace smf
