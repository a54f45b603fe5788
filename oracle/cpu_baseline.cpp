// CPU baseline: a vectorised, multi-core C++ restatement of the reference's
// optimised DG SIPG vmult (TEST / MEASUREMENT INFRASTRUCTURE ONLY: bench.py
// times it as the CPU baseline and tests/ check it against the oracle; the
// product never links or loads it).
//
// It follows the reference's manual path, DG/cells_batched/application.cc:
//   * cells (do_cell_integral_manual, :1493-1735): contiguous DoFs per cell,
//     a batch of four vector lanes-groups shares the shape-gradient matrix
//     (batch_size = 4, :1508), reference gradients at the quadrature points
//     by a GEMM, the affine D_e g <- w_q |K| J^-1 J^-T g (:1614-1666),
//     the transposed GEMM, contiguous write (:1707-1731);
//   * interior faces (do_face_integral_manual + compute_batched_face_integrals,
//     :1738-2431): faces grouped by the (interior face number, exterior face
//     number, orientations) class so a batch shares its face shape matrices;
//     values and reference gradients of both sides at the face points, jump
//     [u] = u- - u+, {d_n u} = (grad u- . nJ- + grad u+ . nJ+)/2 with
//     nJ = J^-1 n, sigma = (p+1)(p+3)/3 max(tau-, tau+) (:2104-2123), test
//     values +-(sigma [u] - {d_n u}) detJ_f w_q and test reference gradients
//     (-[u]/2 detJ_f w_q) nJ^+- (:2131-2187);
//   * boundary faces (compute_batched_boundary_integrals, :2433-2845): mirror
//     u+ = -u-, d_n u+ = d_n u-: test value (2 sigma u - d_n u) detJ_f w_q,
//     test gradient (-u detJ_f w_q) nJ (:2688-2701), sigma = factor tau_K.
// Each cell lists its vertices in ascending global order, so the two sides of
// a face parametrise it identically and the 576 orientation classes of the
// reference reduce to the 16 (face-, face+) classes.  A vector lane is one
// cell or one face (VectorizedArray<double> of 8 lanes: AVX-512, or two AVX2
// halves).  The reference runs one MPI rank per core; here OpenMP threads
// take cell batches, then face batches colour by colour (no two faces of one
// colour share a cell, so the scatter-adds do not race).
//
// Quadrature: P2 uses the symmetric 4-point tetrahedron rule of degree 2
// (stiffness, degree 2p-2) and the 6-point triangle rule of degree 4 (face
// terms, degree 2p); P1 the centroid and the 3-point rule of degree 2; P3
// collapsed Gauss-Legendre rules of degree 4 (tetrahedron) and 6 (triangle).
// On affine cells every rule of that degree gives the same operator.
#include <omp.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <map>
#include <vector>

#include "oracle.h"

namespace {

constexpr int L = 8;      // lanes per vector (cells or faces)
constexpr int BATCH = 4;  // vectors sharing one shape matrix in the GEMMs
constexpr int MAXL = 20;  // dofs per cell (P3)
constexpr int MAXQ = 64;  // cell points
constexpr int MAXQF = 16; // face points
// aligned(64) explicitly: without AVX-512 enabled GCC gives a 64-byte vector
// type only 32-byte alignment, while the AVX-512 kernels load it with aligned
// 64-byte moves; the setup code (compiled without AVX-512) allocates it
typedef double vd __attribute__((vector_size(8 * L), aligned(64)));
static_assert(alignof(vd) == 64, "vector lanes need 64-byte alignment");

// array of vd in 64-byte-aligned memory (std::vector<vd> would drop the
// alignment attribute of the element type)
struct VdArray {
  vd* p = nullptr;
  size_t n = 0;
  VdArray() = default;
  VdArray(const VdArray&) = delete;
  VdArray& operator=(const VdArray&) = delete;
  ~VdArray() { std::free(p); }
  void assign(size_t count) {
    std::free(p);
    n = count;
    p = static_cast<vd*>(std::aligned_alloc(64, std::max<size_t>(count, 1) * sizeof(vd)));
    std::memset(static_cast<void*>(p), 0, std::max<size_t>(count, 1) * sizeof(vd));
  }
  void grow(size_t count) {  // keeps the contents
    vd* q = static_cast<vd*>(std::aligned_alloc(64, std::max<size_t>(count, 1) * sizeof(vd)));
    std::memset(static_cast<void*>(q), 0, std::max<size_t>(count, 1) * sizeof(vd));
    if (p) std::memcpy(static_cast<void*>(q), static_cast<const void*>(p), n * sizeof(vd));
    std::free(p);
    p = q;
    n = count;
  }
  vd& operator[](size_t i) { return p[i]; }
  const vd& operator[](size_t i) const { return p[i]; }
  const vd* data() const { return p; }
  size_t size() const { return n; }
};

struct Rule {
  std::vector<std::array<double, 4>> bc;  // barycentric coordinates
  std::vector<double> w;                  // weights summing to 1
};

// Gauss-Legendre on [0, 1] by Newton iteration on P_n
void gauss01(int n, std::vector<double>& x, std::vector<double>& w) {
  x.resize(n);
  w.resize(n);
  for (int i = 0; i < n; ++i) {
    double z = std::cos(M_PI * (i + 0.75) / (n + 0.5)), dp = 1.0;
    for (int it = 0; it < 100; ++it) {
      double p0 = 1.0, p1 = z;
      for (int k = 2; k <= n; ++k) {
        const double p2 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
        p0 = p1;
        p1 = p2;
      }
      dp = n * (z * p1 - p0) / (z * z - 1.0);
      const double dz = p1 / dp;
      z -= dz;
      if (std::fabs(dz) < 1e-16) break;
    }
    double p0 = 1.0, p1 = z;
    for (int k = 2; k <= n; ++k) {
      const double p2 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
      p0 = p1;
      p1 = p2;
    }
    dp = n * (z * p1 - p0) / (z * z - 1.0);
    x[i] = 0.5 * (1.0 - z);
    w[i] = 1.0 / ((1.0 - z * z) * dp * dp);  // 2/((1-z^2) P'^2) on [-1,1], halved
  }
}

Rule tet_rule(int p) {
  Rule r;
  if (p == 1) {
    r.bc.push_back({0.25, 0.25, 0.25, 0.25});
    r.w.push_back(1.0);
  } else if (p == 2) {
    const double a = (5.0 + 3.0 * std::sqrt(5.0)) / 20.0, b = (5.0 - std::sqrt(5.0)) / 20.0;
    for (int k = 0; k < 4; ++k) {
      std::array<double, 4> c = {b, b, b, b};
      c[k] = a;
      r.bc.push_back(c);
      r.w.push_back(0.25);
    }
  } else {
    // collapsed Gauss-Legendre, 4 points per direction (degree >= 4)
    std::vector<double> x, w;
    gauss01(4, x, w);
    for (int i = 0; i < 4; ++i)
      for (int j = 0; j < 4; ++j)
        for (int k = 0; k < 4; ++k) {
          const double u = x[i], v = x[j] * (1 - u), t = x[k] * (1 - u) * (1 - x[j]);
          r.bc.push_back({1 - u - v - t, u, v, t});
          r.w.push_back(6.0 * w[i] * w[j] * w[k] * (1 - u) * (1 - u) * (1 - x[j]));
        }
  }
  return r;
}

Rule tri_rule(int p) {
  Rule r;
  auto sym3 = [&](double a, double wt) {  // (a, a, 1-2a) and its rotations
    r.bc.push_back({a, a, 1 - 2 * a, 0});
    r.bc.push_back({a, 1 - 2 * a, a, 0});
    r.bc.push_back({1 - 2 * a, a, a, 0});
    for (int k = 0; k < 3; ++k) r.w.push_back(wt);
  };
  if (p == 1) {
    sym3(1.0 / 6.0, 1.0 / 3.0);
  } else if (p == 2) {
    // Dunavant (1985), degree 4
    sym3(0.445948490915965, 0.223381589678011);
    sym3(0.091576213509771, 0.109951743655322);
  } else {
    std::vector<double> x, w;
    gauss01(4, x, w);
    for (int i = 0; i < 4; ++i)
      for (int j = 0; j < 4; ++j) {
        const double u = x[i], v = x[j] * (1 - u);
        r.bc.push_back({1 - u - v, u, v, 0});
        r.w.push_back(2.0 * w[i] * w[j] * (1 - u));
      }
  }
  return r;
}

const int kFaceVerts[4][3] = {{1, 2, 3}, {0, 2, 3}, {0, 1, 3}, {0, 1, 2}};  // face f: opposite vertex f

struct FaceGroup {
  int fm, fp;           // face numbers (fp = -1: boundary)
  int nvec;             // vectors in the group (1..BATCH)
  int64_t first;        // first face slot (BATCH*L slots per group, padded)
};

struct Plan {
  int p = 0, nl = 0, nq = 0, nqf = 0, nthreads = 1;
  bool avx512 = false;  // kernels: AVX-512 when the host has it, else AVX2
  int64_t ncells = 0, ncgroups = 0;
  double factor = 0;
  // tables
  std::vector<double> cG;        // [nq][nl][3]
  std::vector<double> cw;        // [nq]
  std::vector<double> fV, fG;    // [4][nqf][nl], [4][nqf][nl][3]
  std::vector<double> fw;        // [nqf]
  // cells: metric |K| J^-1 J^-T (6) per cell, SoA per vector
  VdArray metric;                // [ncells/L padded][6]
  // faces
  std::vector<FaceGroup> groups;
  std::vector<int64_t> color_ptr;  // group ranges per colour
  std::vector<int64_t> fcell_m, fcell_p;  // per slot (-1: empty lane)
  VdArray fgeo;                    // per vector: nJm(3), nJp(3), area, sigma
};

void inv3(const double J[3][3], double Ji[3][3], double* det) {
  const double d = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) - J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                   J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
  const double id = 1.0 / d;
  Ji[0][0] = (J[1][1] * J[2][2] - J[1][2] * J[2][1]) * id;
  Ji[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) * id;
  Ji[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) * id;
  Ji[1][0] = (J[1][2] * J[2][0] - J[1][0] * J[2][2]) * id;
  Ji[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * id;
  Ji[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * id;
  Ji[2][0] = (J[1][0] * J[2][1] - J[1][1] * J[2][0]) * id;
  Ji[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * id;
  Ji[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * id;
  *det = d;
}

const double kRefVert[4][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};

Plan* build(const orc_mesh* m, int p, int nthreads) {
  auto* P = new Plan;
  P->p = p;
  P->nl = (p + 1) * (p + 2) * (p + 3) / 6;
  P->nthreads = nthreads > 0 ? nthreads : omp_get_max_threads();
  P->factor = (p + 1.0) * (p + 3.0) / 3.0;
  const int nl = P->nl;
  const int64_t nc = m->n_cells;
  P->ncells = nc;
  // cell tables
  const Rule tr = tet_rule(p);
  P->nq = (int)tr.w.size();
  P->cG.resize(P->nq * nl * 3);
  P->cw = tr.w;
  std::vector<double> val(MAXL), grd(3 * MAXL);
  for (int q = 0; q < P->nq; ++q) {
    const double x[3] = {tr.bc[q][1], tr.bc[q][2], tr.bc[q][3]};
    orc_basis(p, x, val.data(), grd.data());
    for (int i = 0; i < nl; ++i)
      for (int d = 0; d < 3; ++d) P->cG[(q * nl + i) * 3 + d] = grd[3 * i + d];
  }
  // face tables: point q of face f at sum_a bc_a * (reference vertex kFaceVerts[f][a])
  const Rule fr = tri_rule(p);
  P->nqf = (int)fr.w.size();
  P->fw = fr.w;
  P->fV.resize(4 * P->nqf * nl);
  P->fG.resize(4 * P->nqf * nl * 3);
  for (int f = 0; f < 4; ++f)
    for (int q = 0; q < P->nqf; ++q) {
      double x[3] = {0, 0, 0};
      for (int a = 0; a < 3; ++a)
        for (int d = 0; d < 3; ++d) x[d] += fr.bc[q][a] * kRefVert[kFaceVerts[f][a]][d];
      orc_basis(p, x, val.data(), grd.data());
      for (int i = 0; i < nl; ++i) {
        P->fV[(f * P->nqf + q) * nl + i] = val[i];
        for (int d = 0; d < 3; ++d) P->fG[((f * P->nqf + q) * nl + i) * 3 + d] = grd[3 * i + d];
      }
    }
  // geometry
  std::vector<double> Ji(9 * nc), det(nc), tau(nc);
  orc_dg_tau(m, tau.data());
  const int64_t nvc = (nc + L - 1) / L;
  P->metric.assign(nvc * 6);
  for (int64_t k = 0; k < nc; ++k) {
    const int32_t* v = m->cells + 4 * k;
    double J[3][3], I[3][3], dt;
    for (int d = 0; d < 3; ++d)
      for (int c = 0; c < 3; ++c) J[d][c] = m->vertices[3 * v[c + 1] + d] - m->vertices[3 * v[0] + d];
    inv3(J, I, &dt);
    for (int a = 0; a < 9; ++a) Ji[9 * k + a] = I[a / 3][a % 3];
    det[k] = std::fabs(dt);
    // |K| J^-1 J^-T (|K| = |det J|/6; the rule's weights sum to 1), entries 00 01 02 11 12 22
    const int ij[6][2] = {{0, 0}, {0, 1}, {0, 2}, {1, 1}, {1, 2}, {2, 2}};
    for (int e = 0; e < 6; ++e) {
      double s = 0;
      for (int c = 0; c < 3; ++c) s += I[ij[e][0]][c] * I[ij[e][1]][c];
      P->metric[(k / L) * 6 + e][k % L] = det[k] / 6.0 * s;
    }
  }
  P->ncgroups = (nvc + BATCH - 1) / BATCH;
  // faces: sorted vertex triples
  struct FE {
    std::array<int32_t, 3> v;
    int64_t cell;
    int lf;
  };
  std::vector<FE> fe(4 * nc);
  for (int64_t k = 0; k < nc; ++k)
    for (int f = 0; f < 4; ++f) {
      FE& e = fe[4 * k + f];
      for (int a = 0; a < 3; ++a) e.v[a] = m->cells[4 * k + kFaceVerts[f][a]];
      std::sort(e.v.begin(), e.v.end());
      e.cell = k;
      e.lf = f;
    }
  std::sort(fe.begin(), fe.end(), [](const FE& a, const FE& b) {
    return a.v != b.v ? a.v < b.v : a.cell < b.cell;
  });
  struct Face {
    int64_t cm, cp;
    int fm, fp;
    double area, n[3];
  };
  std::vector<Face> faces;
  for (size_t k = 0; k < fe.size();) {
    size_t j = k + 1;
    while (j < fe.size() && fe[j].v == fe[k].v) ++j;
    Face F;
    F.cm = fe[k].cell;
    F.fm = fe[k].lf;
    F.cp = j - k == 2 ? fe[k + 1].cell : -1;
    F.fp = j - k == 2 ? fe[k + 1].lf : -1;
    const double* X0 = m->vertices + 3 * fe[k].v[0];
    const double* X1 = m->vertices + 3 * fe[k].v[1];
    const double* X2 = m->vertices + 3 * fe[k].v[2];
    double a[3], b[3];
    for (int d = 0; d < 3; ++d) {
      a[d] = X1[d] - X0[d];
      b[d] = X2[d] - X0[d];
    }
    double n[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
    const double ln = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    F.area = 0.5 * ln;
    const double* Xo = m->vertices + 3 * m->cells[4 * F.cm + F.fm];  // opposite vertex of cm
    double s = 0;
    for (int d = 0; d < 3; ++d) s += n[d] * (X0[d] - Xo[d]);
    for (int d = 0; d < 3; ++d) F.n[d] = (s < 0 ? -n[d] : n[d]) / ln;
    faces.push_back(F);
    k = j;
  }
  // greedy face colouring: no two faces of a colour share a cell
  std::vector<uint8_t> used(nc, 0);  // bit c: colour c taken at the cell
  std::vector<int> color(faces.size());
  int ncol = 0;
  for (size_t f = 0; f < faces.size(); ++f) {
    const uint8_t u = used[faces[f].cm] | (faces[f].cp >= 0 ? used[faces[f].cp] : 0);
    int c = 0;
    while (u & (1u << c)) ++c;
    color[f] = c;
    ncol = std::max(ncol, c + 1);
    used[faces[f].cm] |= (uint8_t)(1u << c);
    if (faces[f].cp >= 0) used[faces[f].cp] |= (uint8_t)(1u << c);
  }
  // groups: per colour, per class (fm, fp), chunks of BATCH vectors of L faces
  P->color_ptr.push_back(0);
  for (int c = 0; c < ncol; ++c) {
    std::map<std::pair<int, int>, std::vector<size_t>> cls;
    for (size_t f = 0; f < faces.size(); ++f)
      if (color[f] == c) cls[{faces[f].fm, faces[f].fp}].push_back(f);
    for (auto& kv : cls) {
      const std::vector<size_t>& list = kv.second;
      for (size_t s = 0; s < list.size(); s += BATCH * L) {
        const size_t n = std::min<size_t>(BATCH * L, list.size() - s);
        FaceGroup g;
        g.fm = kv.first.first;
        g.fp = kv.first.second;
        g.nvec = (int)((n + L - 1) / L);
        g.first = (int64_t)P->fcell_m.size();
        for (int slot = 0; slot < BATCH * L; ++slot) {
          const bool on = (size_t)slot < n;
          const Face* F = on ? &faces[list[s + slot]] : nullptr;
          P->fcell_m.push_back(on ? F->cm : -1);
          P->fcell_p.push_back(on ? F->cp : -1);
          if (slot % L == 0) {
            const size_t need = (P->fcell_m.size() + L - 1) / L * 8;
            if (need > P->fgeo.size()) P->fgeo.grow(std::max(need, 2 * P->fgeo.size()));
          }
          vd* G = &P->fgeo[((P->fcell_m.size() - 1) / L) * 8];
          const int l = slot % L;
          if (!on) continue;
          for (int d = 0; d < 3; ++d) {  // nJ = J^-1 n
            double sm = 0, sp = 0;
            for (int e = 0; e < 3; ++e) {
              sm += Ji[9 * F->cm + 3 * d + e] * F->n[e];
              if (F->cp >= 0) sp += Ji[9 * F->cp + 3 * d + e] * F->n[e];
            }
            G[d][l] = sm;
            G[3 + d][l] = sp;
          }
          G[6][l] = F->area;
          G[7][l] = P->factor * (F->cp >= 0 ? std::max(tau[F->cm], tau[F->cp]) : tau[F->cm]);
        }
        P->groups.push_back(g);
      }
    }
    P->color_ptr.push_back((int64_t)P->groups.size());
  }
  return P;
}

}  // namespace

// the kernels, once per instruction set (the setup above is compiled once)
#pragma GCC push_options
#pragma GCC target("avx2,fma")
namespace k_avx2 {
#include "cpu_baseline_kernels.inc"
}  // namespace k_avx2
#pragma GCC pop_options
#pragma GCC push_options
#pragma GCC target("avx512f,avx512dq,avx512vl,avx512bw,avx2,fma")
namespace k_avx512 {
#include "cpu_baseline_kernels.inc"
}  // namespace k_avx512
#pragma GCC pop_options

namespace {
bool use_avx512() {
  __builtin_cpu_init();
  return __builtin_cpu_supports("avx512f") && __builtin_cpu_supports("avx512dq");
}
}  // namespace

extern "C" {

// plan for the DG SIPG P_p operator of the mesh (cells with vertices in
// ascending order; DoF (k, i) is row k*n_loc + i, node order of orc_basis)
void* orc_dgb_create(const orc_mesh* m, int p, int nthreads) {
  if (p < 1 || p > 3) return nullptr;
  for (int64_t k = 0; k < m->n_cells; ++k)
    for (int a = 0; a < 3; ++a)
      if (m->cells[4 * k + a] >= m->cells[4 * k + a + 1]) return nullptr;  // ascending vertex order required
  Plan* P = build(m, p, nthreads);
  P->avx512 = use_avx512();
  return P;
}

int orc_dgb_vmult(void* plan, const double* src, double* dst) {
  const Plan& P = *static_cast<Plan*>(plan);
  omp_set_num_threads(P.nthreads);
  if (P.avx512) k_avx512::vmult(P, src, dst);
  else k_avx2::vmult(P, src, dst);
  return 0;
}

// the reference's timing protocol (CG/cells_batched/application.cc:1870-1889):
// `reps` repetitions of `n` back-to-back products; seconds per repetition
int orc_dgb_time(void* plan, const double* src, double* dst, int reps, int n, double* seconds) {
  for (int r = 0; r < reps; ++r) {
    const double t0 = omp_get_wtime();
    for (int k = 0; k < n; ++k) orc_dgb_vmult(plan, src, dst);
    seconds[r] = omp_get_wtime() - t0;
  }
  return 0;
}

int orc_dgb_info(void* plan, int64_t* out) {
  const Plan& P = *static_cast<Plan*>(plan);
  out[0] = P.nq;
  out[1] = P.nqf;
  out[2] = (int64_t)P.groups.size();
  out[3] = (int64_t)P.color_ptr.size() - 1;
  out[4] = P.nthreads;
  out[5] = P.avx512 ? 512 : 256;
  return 0;
}

// quadrature rules used, for the exactness test: kind 0 tetrahedron (4
// barycentric coordinates per point), 1 triangle (3 per point); returns n
int orc_dgb_rule(int kind, int p, double* bc, double* w) {
  const Rule r = kind == 0 ? tet_rule(p) : tri_rule(p);
  const int nb = kind == 0 ? 4 : 3;
  for (size_t q = 0; q < r.w.size(); ++q) {
    if (bc)
      for (int a = 0; a < nb; ++a) bc[q * nb + a] = r.bc[q][a];
    if (w) w[q] = r.w[q];
  }
  return (int)r.w.size();
}

void orc_dgb_destroy(void* plan) { delete static_cast<Plan*>(plan); }

}  // extern "C"
