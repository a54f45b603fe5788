// Hybrid multigrid: level hierarchy, transfers, Chebyshev-Jacobi smoothing,
// FP64 dense coarse solve, mixed-precision V-cycle, (flexible) CG and the
// right-hand side.  See mg.h for the reference counterparts.
#include "mg.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <stdexcept>

#include "amg.h"
#include "matrix.h"

namespace smf {

void launch_invert_diagonal(int precision, int64_t n, void* d, hipStream_t s);
void launch_rhs_cells(int64_t ncells, int nl, int nq, const double* geo12, const double* qtab, int kind, int nf,
                      const double* feat, double* out, hipStream_t s);

void Csr::upload(int precision) {
  d_ptr.upload(ptr);
  d_col.upload(col);
  if (precision == MF_F64) {
    d_val.upload(val);
  } else {
    std::vector<float> v(val.begin(), val.end());
    d_val.upload(v);
  }
}

std::vector<double> dense_inverse_spd(std::vector<double> A, int64_t n);

namespace {

const int kPerm[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};

// Kuhn topology of size n whose vertex (i,j,k) is the fine vertex (s i, s j, s k)
Mesh coarse_kuhn(const Mesh& fine, int n) {
  const int nf = fine.kuhn_n, s = nf / n;
  Mesh m = kuhn_cube(n, 0.0, 0);
  const int64_t n1 = n + 1, f1 = nf + 1;
  for (int64_t k = 0; k < n1; ++k)
    for (int64_t j = 0; j < n1; ++j)
      for (int64_t i = 0; i < n1; ++i) {
        const int64_t v = i + n1 * (j + n1 * k), fv = s * i + f1 * (s * j + f1 * s * k);
        for (int d = 0; d < 3; ++d) m.x[3 * v + d] = fine.x[3 * fv + d];
      }
  return m;
}

std::unique_ptr<Operator> make_cg_operator(const Mesh& mesh, int degree, int precision, const mf_options& opt) {
  auto op = std::make_unique<Operator>();
  op->kind = MF_CG_POISSON;
  op->precision = precision;
  op->degree = degree;
  op->opt = opt;
  op->opt.host_only = 0;
  op->mesh = mesh;
  op->cg.value_bytes = precision == MF_F64 ? 8 : 4;
  build_cg_plan(op->mesh, degree, op->opt, op->cg);
  make_cg_shape(op->cg);
  upload_cg(*op);
  return op;
}

std::unique_ptr<Operator> make_dg_operator(const Mesh& mesh, int degree, int precision, const mf_options& opt) {
  auto op = std::make_unique<Operator>();
  op->kind = MF_DG_SIPG;
  op->precision = precision;
  op->degree = degree;
  op->ncomp = 1;
  op->opt = opt;
  op->opt.host_only = 0;
  op->mesh = mesh;
  build_dg_plan(op->mesh, degree, 1, op->opt, false, op->dg);
  upload_dg(*op);
  return op;
}

std::vector<int64_t> inverse_perm(const std::vector<int64_t>& perm) {
  std::vector<int64_t> inv(perm.size());
  for (size_t k = 0; k < perm.size(); ++k) inv[perm[k]] = (int64_t)k;
  return inv;
}

// Solve J xhat = X - L0 for the 3x3 lattice Jacobian (columns L_b - L_0).
void solve3(const double J[3][3], const double r[3], double x[3]) {
  const double det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
                     J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                     J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
  for (int c = 0; c < 3; ++c) {
    double M[3][3];
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) M[a][b] = b == c ? r[a] : J[a][b];
    x[c] = (M[0][0] * (M[1][1] * M[2][2] - M[1][2] * M[2][1]) - M[0][1] * (M[1][0] * M[2][2] - M[1][2] * M[2][0]) +
            M[0][2] * (M[1][0] * M[2][1] - M[1][1] * M[2][0])) /
           det;
  }
}

// Prolongation from the coarse level to the fine level by reference-cell
// interpolation on the unperturbed lattice: fine DoF value = coarse function
// at the fine node's lattice position (MGTwoLevelTransfer semantics,
// application_ph.cc:1902-1907).  Dirichlet rows and columns are dropped.
Csr build_prolongation(const Operator& fine, const Operator& coarse, bool same_mesh) {
  const CGPlan& fp = fine.cg;
  const CGPlan& cp = coarse.cg;
  const Lagrange ffe(fp.p), cfe(cp.p);
  const std::vector<int64_t> cinv = inverse_perm(cp.cell_perm);
  const Mesh& fm = fine.mesh;
  const Mesh& cm = coarse.mesh;
  std::vector<std::vector<std::pair<int32_t, double>>> rows(fp.n_dofs);
  std::vector<char> done(fp.n_dofs, 0);
  std::vector<double> cval(cp.n_loc);
  for (int64_t k = 0; k < fp.n_cells; ++k) {
    const int64_t c = fp.cell_perm[k];
    for (int i = 0; i < fp.n_loc; ++i) {
      const int64_t d = fp.cell_dofs[k * fp.n_loc + i];
      if (d >= fp.n_free || done[d]) continue;
      done[d] = 1;
      const auto xr = ffe.node_point(i);
      int64_t cc;
      double xhat[3];
      if (same_mesh) {
        cc = c;
        for (int a = 0; a < 3; ++a) xhat[a] = xr[a];
      } else {
        // lattice position of the fine node, in coarse lattice units
        double X[3];
        const int32_t* v = &fm.cells[4 * c];
        for (int a = 0; a < 3; ++a) {
          const double L0 = fm.lattice[3 * v[0] + a];
          double s = L0;
          for (int b = 0; b < 3; ++b) s += (fm.lattice[3 * v[b + 1] + a] - L0) * xr[b];
          X[a] = 0.5 * s;
        }
        const int nc = cm.kuhn_n;
        int h[3];
        double t[3];
        for (int a = 0; a < 3; ++a) {
          h[a] = std::min(std::max((int)std::floor(X[a]), 0), nc - 1);
          t[a] = X[a] - h[a];
        }
        int ord[3] = {0, 1, 2};
        std::sort(ord, ord + 3, [&](int a, int b) { return t[a] > t[b] || (t[a] == t[b] && a < b); });
        int s = 0;
        for (; s < 6; ++s)
          if (kPerm[s][0] == ord[0] && kPerm[s][1] == ord[1] && kPerm[s][2] == ord[2]) break;
        cc = 6ll * (h[0] + (int64_t)nc * (h[1] + (int64_t)nc * h[2])) + s;
        const int32_t* w = &cm.cells[4 * cc];
        double J[3][3], r[3];
        for (int a = 0; a < 3; ++a) {
          const double L0 = cm.lattice[3 * w[0] + a];
          for (int b = 0; b < 3; ++b) J[a][b] = cm.lattice[3 * w[b + 1] + a] - L0;
          r[a] = X[a] - L0;
        }
        solve3(J, r, xhat);
      }
      cfe.eval(xhat, cval.data(), nullptr);
      const int64_t kc = cinv[cc];
      for (int j = 0; j < cp.n_loc; ++j) {
        const int64_t dc = cp.cell_dofs[kc * cp.n_loc + j];
        if (dc >= cp.n_free || std::fabs(cval[j]) < 1e-13) continue;
        rows[d].push_back({(int32_t)dc, cval[j]});
      }
    }
  }
  Csr P;
  P.nrows = fp.n_dofs;
  P.ncols = cp.n_dofs;
  P.ptr.assign(1, 0);
  for (int64_t d = 0; d < fp.n_dofs; ++d) {
    std::sort(rows[d].begin(), rows[d].end());
    for (auto& e : rows[d]) {
      P.col.push_back(e.first);
      P.val.push_back(e.second);
    }
    P.ptr.push_back((int64_t)P.col.size());
  }
  return P;
}

// c-transfer: injection of the CG P_k space into the DG P_k space on the
// same mesh (DG row (k, i) takes the CG DoF at the same support point), as
// the DG/CG MGTwoLevelTransfer of application_phc.cc:3416-3552.
Csr build_c_prolongation(const Operator& dg, const Operator& cg) {
  const DGPlan& dp = dg.dg;
  const CGPlan& cp = cg.cg;
  if (dp.p != cp.p || dp.n_cells != cp.n_cells) throw std::runtime_error("c-transfer needs equal degree and mesh");
  const int nl = dp.n_loc;
  const std::vector<int64_t> cinv = inverse_perm(cp.cell_perm);
  Csr P;
  P.nrows = dp.n_dofs;
  P.ncols = cp.n_dofs;
  P.ptr.resize(P.nrows + 1);
  P.col.resize(P.nrows);
  P.val.assign(P.nrows, 1.0);
  for (int64_t k = 0; k < dp.n_cells; ++k) {
    const int64_t kc = cinv[dp.cell_perm[k]];
    double h2 = 0;  // squared cell size for the matching tolerance
    for (int a = 0; a < 3; ++a) {
      const double d = dp.dof_points[3 * (k * nl + 1) + a] - dp.dof_points[3 * (k * nl) + a];
      h2 += d * d;
    }
    for (int i = 0; i < nl; ++i) {
      const double* X = &dp.dof_points[3 * (k * nl + i)];
      int best = -1;
      double bd = 1e300;
      for (int j = 0; j < nl; ++j) {
        const double* Y = &cp.dof_points[3 * cp.cell_dofs[kc * nl + j]];
        const double d = (X[0] - Y[0]) * (X[0] - Y[0]) + (X[1] - Y[1]) * (X[1] - Y[1]) + (X[2] - Y[2]) * (X[2] - Y[2]);
        if (d < bd) {
          bd = d;
          best = j;
        }
      }
      if (!(bd <= 1e-16 * h2)) throw std::runtime_error("c-transfer: DG node without a CG node");
      P.col[k * nl + i] = (int32_t)cp.cell_dofs[kc * nl + best];
      P.ptr[k * nl + i + 1] = k * nl + i + 1;
    }
  }
  P.ptr[0] = 0;
  return P;
}

Csr transpose(const Csr& A) {
  Csr T;
  T.nrows = A.ncols;
  T.ncols = A.nrows;
  T.ptr.assign(T.nrows + 1, 0);
  for (int32_t c : A.col) T.ptr[c + 1]++;
  for (int64_t r = 0; r < T.nrows; ++r) T.ptr[r + 1] += T.ptr[r];
  T.col.resize(A.col.size());
  T.val.resize(A.val.size());
  std::vector<int64_t> pos(T.ptr.begin(), T.ptr.end() - 1);
  for (int64_t r = 0; r < A.nrows; ++r)
    for (int64_t k = A.ptr[r]; k < A.ptr[r + 1]; ++k) {
      const int64_t p = pos[A.col[k]]++;
      T.col[p] = (int32_t)r;
      T.val[p] = A.val[k];
    }
  return T;
}

// dense FP64 matrix of a CG operator from its element matrices (identity on
// Dirichlet rows), then its inverse by Cholesky
std::vector<double> dense_inverse(const Operator& op) {
  const CGPlan& pl = op.cg;
  const int64_t n = pl.n_dofs;
  std::vector<double> A(n * n, 0.0);
  const int nl = pl.n_loc, nq = pl.n_q;
  const double* G = pl.shape.data();
  const double* W = G + nl * nq * 3;
  for (int64_t k = 0; k < pl.n_cells; ++k) {
    double m[6];
    if (!pl.curved)
      for (int t = 0; t < 6; ++t) m[t] = pl.metric[(size_t)t * pl.n_cells + k];
    for (int i = 0; i < nl; ++i) {
      const int64_t di = pl.cell_dofs[k * nl + i];
      if (di >= pl.n_free) continue;
      for (int j = 0; j < nl; ++j) {
        const int64_t dj = pl.cell_dofs[k * nl + j];
        if (dj >= pl.n_free) continue;
        double s = 0;
        for (int q = 0; q < nq; ++q) {
          const double* a = &G[(q * nl + i) * 3];
          const double* b = &G[(q * nl + j) * 3];
          double w = W[q];
          if (pl.curved) {  // per-point metric, weight included
            for (int t = 0; t < 6; ++t) m[t] = pl.qmetric[((size_t)q * 6 + t) * pl.n_cells + k];
            w = 1.0;
          }
          s += w * (a[0] * (m[0] * b[0] + m[1] * b[1] + m[2] * b[2]) +
                       a[1] * (m[1] * b[0] + m[3] * b[1] + m[4] * b[2]) +
                       a[2] * (m[2] * b[0] + m[4] * b[1] + m[5] * b[2]));
        }
        A[di * n + dj] += s;
      }
    }
  }
  for (int64_t i = pl.n_free; i < n; ++i) A[i * n + i] = 1.0;
  if (pl.nitsche)
    for (size_t r = 0; r < pl.bnd_rows.size(); ++r)
      for (int32_t k = pl.bnd_ptr[r]; k < pl.bnd_ptr[r + 1]; ++k) A[pl.bnd_rows[r] * n + pl.bnd_col[k]] += pl.bnd_val[k];
  return dense_inverse_spd(std::move(A), n);
}

}  // namespace

// inverse of a dense symmetric positive definite row-major n x n matrix
std::vector<double> dense_inverse_spd(std::vector<double> A, int64_t n) {
  // Cholesky A = L L^T (in place, lower; rows of a column in parallel)
  for (int64_t j = 0; j < n; ++j) {
    double s = A[j * n + j];
    for (int64_t k = 0; k < j; ++k) s -= A[j * n + k] * A[j * n + k];
    if (!(s > 0)) throw std::runtime_error("coarse matrix not positive definite");
    const double l = std::sqrt(s);
    A[j * n + j] = l;
#pragma omp parallel for schedule(static) if (n - j > 256)
    for (int64_t i = j + 1; i < n; ++i) {
      double t = A[i * n + j];
      for (int64_t k = 0; k < j; ++k) t -= A[i * n + k] * A[j * n + k];
      A[i * n + j] = t / l;
    }
  }
  // inverse: solve L L^T X = I, columns in parallel (results independent of the thread count)
  std::vector<double> invt(n * n);  // invt[c * n + i] = inv[i][c]
#pragma omp parallel
  {
    std::vector<double> y(n);
#pragma omp for schedule(dynamic, 16)
    for (int64_t c = 0; c < n; ++c) {
      for (int64_t i = 0; i < n; ++i) {
        double t = (i == c) ? 1.0 : 0.0;
        for (int64_t k = 0; k < i; ++k) t -= A[i * n + k] * y[k];
        y[i] = t / A[i * n + i];
      }
      double* xc = &invt[c * n];
      for (int64_t i = n - 1; i >= 0; --i) {
        double t = y[i];
        for (int64_t k = i + 1; k < n; ++k) t -= A[k * n + i] * xc[k];
        xc[i] = t / A[i * n + i];
      }
    }
  }
  // the inverse is symmetric: invt is its row-major form up to rounding;
  // return it transposed back so that entry (i, c) is the solve's result
  std::vector<double> inv(n * n);
  for (int64_t c = 0; c < n; ++c)
    for (int64_t i = 0; i < n; ++i) inv[i * n + c] = invt[c * n + i];
  return inv;
}

namespace {

// largest eigenvalue of a symmetric tridiagonal matrix by Sturm bisection
double tridiag_max_eig(const std::vector<double>& a, const std::vector<double>& b) {
  const int n = (int)a.size();
  double lo = 1e300, hi = -1e300;
  for (int i = 0; i < n; ++i) {
    const double r = (i > 0 ? std::fabs(b[i - 1]) : 0) + (i + 1 < n ? std::fabs(b[i]) : 0);
    lo = std::min(lo, a[i] - r);
    hi = std::max(hi, a[i] + r);
  }
  auto count_below = [&](double x) {
    int c = 0;
    double q = a[0] - x;
    if (q < 0) ++c;
    for (int i = 1; i < n; ++i) {
      q = a[i] - x - b[i - 1] * b[i - 1] / (q != 0 ? q : 1e-300);
      if (q < 0) ++c;
    }
    return c;
  };
  for (int it = 0; it < 200; ++it) {
    const double mid = 0.5 * (lo + hi);
    if (count_below(mid) >= n) hi = mid;
    else lo = mid;
  }
  return hi;
}

template <typename T>
T* P_(DevBuf& b) {
  return b.as<T>();
}

template <typename T>
void level_vmult(const MGLevel& L, T* dst, const T* src, hipStream_t s) {
  launch_op(*L.op, 0, dst, src, s);
}

template <typename T>
double device_dot(Multigrid& mg, int64_t n, const T* x, const T* y, hipStream_t s) {
  blas_dot<T>(n, x, y, mg.d_partial.as<double>(), mg.d_scalar.as<double>(), s);
  double h;
  SMF_HIP(hipMemcpyAsync(&h, mg.d_scalar.p, sizeof(double), hipMemcpyDeviceToHost, s));
  SMF_HIP(hipStreamSynchronize(s));
  return h;
}

// lambda_max of D^-1 A by CG iterations (Lanczos), as PreconditionChebyshev's
// eigenvalue estimate
template <typename T>
double estimate_lambda_max(Multigrid& mg, MGLevel& L, int iters, hipStream_t s) {
  const int64_t n = L.op->m();
  T* x = L.x.as<T>();
  T* r = L.r.as<T>();
  T* z = L.t.as<T>();
  T* p = L.d.as<T>();
  T* q = L.b.as<T>();
  // r = 1 on free rows (constrained rows are identity rows of eigenvalue 1)
  blas_fill<T>(n, T(1), r, s);
  blas_fill<T>(n, T(0), x, s);
  blas_scale<T>(n, L.inv_diag.as<T>(), r, z, s);
  SMF_HIP(hipMemcpyAsync(p, z, n * sizeof(T), hipMemcpyDeviceToDevice, s));
  double rz = device_dot<T>(mg, n, r, z, s);
  std::vector<double> alpha, beta;
  for (int it = 0; it < iters && rz > 0; ++it) {
    level_vmult<T>(L, q, p, s);
    const double pq = device_dot<T>(mg, n, p, q, s);
    const double a = rz / pq;
    blas_axpby<T>(n, -a, q, 1.0, r, s);
    blas_scale<T>(n, L.inv_diag.as<T>(), r, z, s);
    const double rz_new = device_dot<T>(mg, n, r, z, s);
    const double b = rz_new / rz;
    alpha.push_back(a);
    beta.push_back(b);
    blas_axpby<T>(n, 1.0, z, b, p, s);
    rz = rz_new;
  }
  const int m = (int)alpha.size();
  if (m == 0) return 1.0;
  std::vector<double> da(m), ob(std::max(m - 1, 0));
  for (int k = 0; k < m; ++k) {
    da[k] = 1.0 / alpha[k] + (k > 0 ? beta[k - 1] / alpha[k - 1] : 0.0);
    if (k + 1 < m) ob[k] = std::sqrt(beta[k]) / alpha[k];
  }
  return tridiag_max_eig(da, ob);
}

// Chebyshev smoothing of A x = b on level L (Saad, Iterative Methods, Alg.
// 12.1, on the Jacobi-preconditioned system), `degree` steps, x = L.x;
// zero_start skips the first residual product.  CG levels apply each step
// inside the operator's kernels, in place; scalar CG levels run the steps
// as a ChebyRun (one kernel per step, the interface rows of a step completed
// by the next step's kernel, one closing kernel; Nitsche boundary terms added
// per cell inside the kernel); DG
// levels with the four-lane kernel write x + d to L.t and swap L.t with L.x
// (the neighbour gathers of other cells still read the old x), so L.x holds
// the iterate after every step.
template <typename T>
void chebyshev(Multigrid& mg, MGLevel& L, const T* b, bool zero_start, hipStream_t s) {
  const int64_t n = L.op->m();
  const double theta = 0.5 * (L.lambda_max + L.lambda_min), delta = 0.5 * (L.lambda_max - L.lambda_min);
  const double sigma = theta / delta;
  double rho = 1.0 / sigma;
  T* d = L.d.as<T>();
  const T* dinv = L.inv_diag.as<T>();
  const bool dg = L.op->is_dg();
  const bool fused = !dg || dg_has_epilogue(*L.op);
  RowEpilogue<T> ep;
  ep.kind = 1;
  ep.b = b;
  ep.dinv = dinv;
  ep.d = d;
  const bool use_run = !dg && cg_has_cheby_run(*L.op) && mg.opt.fuse_steps != 0;
  ChebyRun<T> run;
  if (use_run) {
    run.b = b;
    run.dinv = dinv;
    run.x = L.x.as<T>();
    run.d = d;
    for (int k = 0; k < 2; ++k) {
      run.part[k] = L.run_part[k].as<T>();
      run.xi[k] = L.run_xi[k].as<T>();
      run.di[k] = L.run_di[k].as<T>();
    }
  }
  for (int k = 0; k < mg.opt.smoother_degree; ++k) {
    T* x = L.x.as<T>();
    T c1 = T(0), c2;
    if (k == 0) {
      c2 = T(1.0 / theta);
      if (zero_start) {
        blas_cheby_zero<T>(n, c2, dinv, b, d, x, s);
        continue;
      }
    } else {
      const double rho_new = 1.0 / (2.0 * sigma - rho);
      c1 = T(rho_new * rho);
      c2 = T(2.0 * rho_new / delta);
      rho = rho_new;
    }
    if (use_run) {
      launch_cg_cheby_step<T>(*L.op, run, c1, c2, k == 0, s);
    } else if (fused) {
      ep.first = k == 0;
      ep.c1 = c1;
      ep.c2 = c2;
      ep.x = x;
      if (dg) {
        ep.xout = L.t.as<T>();
        launch_dg_epilogue<T>(*L.op, x, ep, s);
        std::swap(L.x.p, L.t.p);
      } else {
        launch_cg_epilogue<T>(*L.op, x, ep, s);
      }
    } else {
      T* ax = L.t.as<T>();
      level_vmult<T>(L, ax, x, s);
      blas_cheby_step<T>(n, c1, c2, dinv, b, ax, d, x, k == 0, s);
    }
  }
  if (use_run) launch_cg_cheby_close<T>(*L.op, run, s);
}

}  // namespace

// dst = P src (dir 0) or R src (dir 1) of level L, or dst += with add: the
// row-blocked CSR-stream product when the level has it (DG -> CG restriction
// at 48^3: 6.6 M entries in 0.9 M rows), else one thread per row
template <typename T>
void level_transfer(const MGLevel& L, int dir, T* dst, const T* src, bool add, hipStream_t s) {
  const SysMatrix& M = dir == 0 ? L.Pm : L.Rm;
  if (M.on_device) return launch_spmv(M, dst, src, s, add);
  const Csr& A = dir == 0 ? L.P : L.R;
  blas_csr<T>(A.nrows, A.d_ptr.as<int64_t>(), A.d_col.as<int32_t>(), A.d_val.as<T>(), src, dst, add, s);
}
template void level_transfer<double>(const MGLevel&, int, double*, const double*, bool, hipStream_t);
template void level_transfer<float>(const MGLevel&, int, float*, const float*, bool, hipStream_t);

namespace {

template <typename T>
void vcycle(Multigrid& mg, int l, hipStream_t s) {
  MGLevel& L = mg.levels[l];
  const int64_t n = L.op->m();
  T* x = L.x.as<T>();
  T* b = L.b.as<T>();
  if (l == 0 && mg.amg) {
    // one FP64 AMG V-cycle
    Amg& amg = *mg.amg;
    blas_convert<double, T>(n, b, amg.lv[0].b.as<double>(), s);
    amg_vcycle(amg, s);
    blas_convert<T, double>(n, amg.lv[0].x.as<double>(), x, s);
    return;
  }
  if (l == 0) {
    // FP64 dense coarse solve
    blas_convert<double, T>(n, b, mg.d_cb.as<double>(), s);
    blas_dense_gemv(mg.n_coarse_free, mg.n_coarse, mg.d_coarse_inv.as<double>(), mg.d_cb.as<double>(),
                    mg.d_cx.as<double>(), s);
    blas_convert<T, double>(n, mg.d_cx.as<double>(), x, s);
    return;
  }
  MGLevel& C = mg.levels[l - 1];
  chebyshev<T>(mg, L, b, true, s);
  x = L.x.as<T>();
  // r = b - A x ; b_c = R r
  T* r = L.r.as<T>();
  if (!L.op->is_dg() || dg_has_epilogue(*L.op)) {
    RowEpilogue<T> ep;
    ep.kind = 2;
    ep.b = b;
    ep.r = r;
    if (L.op->is_dg()) launch_dg_epilogue<T>(*L.op, x, ep, s);
    else launch_cg_epilogue<T>(*L.op, x, ep, s);
  } else {
    level_vmult<T>(L, r, x, s);
    blas_axpby_t<T>(n, T(1), b, T(-1), r, s);
  }
  level_transfer<T>(L, 1, C.b.as<T>(), r, false, s);
  vcycle<T>(mg, l - 1, s);
  level_transfer<T>(L, 0, x, C.x.as<T>(), true, s);  // x += P x_c
  chebyshev<T>(mg, L, b, false, s);
}

}  // namespace

std::unique_ptr<Multigrid> build_multigrid(const Operator& fine, const MGOptions& opt) {
  const bool dg = fine.kind == MF_DG_SIPG;
  if (fine.kind != MF_CG_POISSON && !dg)
    throw std::runtime_error("multigrid needs a scalar CG Poisson or DG SIPG fine operator");
  if (fine.kind == MF_CG_POISSON && fine.cg.nitsche)
    throw std::runtime_error("multigrid below a Nitsche CG operator: build it from the DG operator");
  if (dg && fine.dg.curved)
    throw std::runtime_error("multigrid below a curved DG operator is not built (its CG levels would need curved "
                             "Nitsche boundaries)");
  auto mg = std::make_unique<Multigrid>();
  mg->opt = opt;
  mg->fine_precision = fine.precision;
  mg->n_fine = fine.m();
  const int prec = opt.level_precision;
  mf_options lopt = fine.opt;
  // DG: the CG levels carry the Nitsche boundary (no strong constraints),
  // application_phc.cc:3519-3525
  if (dg) lopt.boundary = 1;
  // levels, finest first: [DG P_k, c-level CG P_k,] P_k-1..P_1 on the fine
  // mesh, then P1 on coarser Kuhn meshes
  std::vector<std::unique_ptr<Operator>> ops;
  std::vector<int> kn;
  if (dg) {
    ops.push_back(make_dg_operator(fine.mesh, fine.degree, prec, fine.opt));
    kn.push_back(fine.mesh.kuhn_n);
    if (ops[0]->dg.cell_perm != fine.dg.cell_perm) throw std::runtime_error("level numbering differs from the fine operator");
  }
  ops.push_back(make_cg_operator(fine.mesh, fine.degree, prec, lopt));
  kn.push_back(fine.mesh.kuhn_n);
  if (!dg && ops[0]->cg.cell_dofs != fine.cg.cell_dofs)
    throw std::runtime_error("level numbering differs from the fine operator");
  for (int p = fine.degree - 1; p >= 1; --p) {
    ops.push_back(make_cg_operator(fine.mesh, p, prec, lopt));
    kn.push_back(fine.mesh.kuhn_n);
  }
  // h-coarsening stops at the first level the dense FP64 coarse solver
  // takes (each level below costs a V-cycle a dozen latency-bound launches)
  int n = fine.mesh.kuhn_n;
  if (n > 0) {
    while (n % 2 == 0 && n / 2 >= 2 && ops.back()->m() > opt.max_coarse_dofs) {
      n /= 2;
      ops.push_back(make_cg_operator(coarse_kuhn(fine.mesh, n), 1, prec, lopt));
      kn.push_back(n);
    }
  }
  const bool use_amg = opt.coarse_solver == 2 || (opt.coarse_solver == 0 && ops.back()->m() > opt.max_coarse_dofs);
  if (!use_amg && ops.back()->m() > opt.max_coarse_dofs)
    throw std::runtime_error("coarsest level too large for the dense coarse solver (needs a Kuhn mesh hierarchy "
                             "or the AMG coarse solver)");
  // assemble levels coarsest first
  const size_t nl = ops.size();
  mg->levels.resize(nl);
  for (size_t i = 0; i < nl; ++i) {
    MGLevel& L = mg->levels[nl - 1 - i];
    L.op = std::move(ops[i]);
    L.degree = L.op->degree;
    L.kuhn_n = kn[i];
  }
  const size_t w = prec == MF_F64 ? 8 : 4;
  mg->d_partial.alloc(blas_reduction_parts() * sizeof(double));
  mg->d_scalar.alloc(sizeof(double));
  hipStream_t s = nullptr;
  for (size_t l = 0; l < nl; ++l) {
    MGLevel& L = mg->levels[l];
    const int64_t m = L.op->m();
    for (DevBuf* b : {&L.inv_diag, &L.x, &L.b, &L.r, &L.t, &L.d}) b->alloc(m * w);
    if (cg_has_cheby_run(*L.op)) {
      const CGPlan& pl = L.op->cg;
      for (int k = 0; k < 2; ++k) {
        L.run_part[k].alloc(std::max<int64_t>(pl.n_ghost_total, 1) * w);
        L.run_xi[k].alloc(std::max<int64_t>(pl.n_free - pl.n_int, 1) * w);
        L.run_di[k].alloc(std::max<int64_t>(pl.n_free - pl.n_int, 1) * w);
      }
    }
    launch_op(*L.op, 1, L.inv_diag.p, nullptr, s);
    launch_invert_diagonal(prec, m, L.inv_diag.p, s);
    if (l > 0) {
      const bool same = L.op->mesh.nc == mg->levels[l - 1].op->mesh.nc && L.kuhn_n == mg->levels[l - 1].kuhn_n;
      if (L.op->is_dg()) L.P = build_c_prolongation(*L.op, *mg->levels[l - 1].op);
      else L.P = build_prolongation(*L.op, *mg->levels[l - 1].op, same);
      L.R = transpose(L.P);
      L.P.upload(prec);
      L.R.upload(prec);
      upload_blocked(L.P.ptr, L.P.col, L.P.val, prec, L.Pm);
      upload_blocked(L.R.ptr, L.R.col, L.R.val, prec, L.Rm);
    }
  }
  SMF_HIP(hipDeviceSynchronize());
  // smoother intervals: [1.2 lmax / range, 1.2 lmax]
  for (size_t l = 1; l < nl; ++l) {
    MGLevel& L = mg->levels[l];
    const double lmax = prec == MF_F64 ? estimate_lambda_max<double>(*mg, L, opt.eig_cg_iterations, s)
                                       : estimate_lambda_max<float>(*mg, L, opt.eig_cg_iterations, s);
    L.lambda_max = 1.2 * lmax;
    L.lambda_min = L.lambda_max / opt.smoothing_range;
  }
  // coarse solver
  const Operator& c0 = *mg->levels[0].op;
  mg->n_coarse = (int)c0.m();
  mg->d_cb.alloc(std::max<int64_t>(c0.m(), 1) * sizeof(double));
  mg->d_cx.alloc(std::max<int64_t>(c0.m(), 1) * sizeof(double));
  if (use_amg) {
    // one AMG V-cycle on the coarsest level's assembled FP64 matrix
    if (c0.is_dg()) throw std::runtime_error("AMG coarse solver: the coarsest level must be a CG level");
    SysMatrix sm;
    assemble_cg_matrix(c0, sm);
    Csr A0;
    A0.nrows = A0.ncols = sm.n_rows;
    A0.ptr = std::move(sm.ptr);
    A0.col = std::move(sm.col);
    A0.val = std::move(sm.val);
    mg->amg = build_amg(A0, opt.max_coarse_dofs, opt.amg_threshold);
    mg->amg->smoother_degree = opt.smoother_degree;  // its finest level, as on the geometric levels
    SMF_HIP(hipDeviceSynchronize());
    return mg;
  }
  std::vector<double> inv;
  {
    // the dense inverse needs the FP64 element data of the coarsest level
    inv = dense_inverse(c0);
  }
  // strong Dirichlet rows make A = [[A_ff, 0], [0, I]] (free DoFs first), so
  // only the leading block of the inverse is kept
  const int64_t nf = !c0.is_dg() && !c0.cg.nitsche ? c0.cg.n_free : c0.m();
  mg->n_coarse_free = (int)nf;
  if (nf < (int64_t)mg->n_coarse) {
    const int64_t n = mg->n_coarse;
    for (int64_t i = 0; i < nf; ++i)
      for (int64_t j = 0; j < nf; ++j) inv[i * nf + j] = inv[i * n + j];
    inv.resize(nf * nf);
  }
  mg->d_coarse_inv.upload(inv);
  SMF_HIP(hipDeviceSynchronize());
  return mg;
}

void mg_vmult(Multigrid& mg, void* z, const void* r, hipStream_t s) {
  MGLevel& top = mg.levels.back();
  const int64_t n = top.op->m();
  const bool f32 = mg.opt.level_precision == MF_F32;
  // copy_to_mg
  if (mg.fine_precision == MF_F64) {
    if (f32) blas_convert<float, double>(n, static_cast<const double*>(r), top.b.as<float>(), s);
    else SMF_HIP(hipMemcpyAsync(top.b.p, r, n * 8, hipMemcpyDeviceToDevice, s));
  } else {
    if (f32) SMF_HIP(hipMemcpyAsync(top.b.p, r, n * 4, hipMemcpyDeviceToDevice, s));
    else blas_convert<double, float>(n, static_cast<const float*>(r), top.b.as<double>(), s);
  }
  const int ltop = (int)mg.levels.size() - 1;
  if (f32) vcycle<float>(mg, ltop, s);
  else vcycle<double>(mg, ltop, s);
  // copy_from_mg
  if (mg.fine_precision == MF_F64) {
    if (f32) blas_convert<double, float>(n, top.x.as<float>(), static_cast<double*>(z), s);
    else SMF_HIP(hipMemcpyAsync(z, top.x.p, n * 8, hipMemcpyDeviceToDevice, s));
  } else {
    if (f32) SMF_HIP(hipMemcpyAsync(z, top.x.p, n * 4, hipMemcpyDeviceToDevice, s));
    else blas_convert<float, double>(n, top.x.as<double>(), static_cast<float*>(z), s);
  }
  mg.vcycles++;
}

int64_t cg_workspace_doubles(int64_t n) {
  const int64_t nv = (n + 31) / 32 * 32;  // 256-byte aligned vectors
  return 4 * nv + (blas_reduction_parts() + 31) / 32 * 32 + 32 + kS_N;
}

CgWorkspace cg_workspace_at(double* base, int64_t n) {
  const int64_t nv = (n + 31) / 32 * 32;
  CgWorkspace ws;
  ws.r = base;
  ws.z = base + nv;
  ws.p = base + 2 * nv;
  ws.q = base + 3 * nv;
  ws.partial = base + 4 * nv;
  ws.scalar = ws.partial + (blas_reduction_parts() + 31) / 32 * 32;
  ws.S = ws.scalar + 32;
  return ws;
}

SolveStats solve_cg(const Operator& A, Multigrid* mg, double* x, const double* b, double rel_tol, int max_iter,
                    bool flexible, hipStream_t s) {
  if (A.precision != MF_F64) throw std::runtime_error("the outer solver runs in FP64");
  const int64_t n = A.m();
  DevBuf work;
  work.alloc(cg_workspace_doubles(n) * sizeof(double));
  return solve_cg_ws(A, mg, x, b, rel_tol, max_iter, flexible, cg_workspace_at(work.as<double>(), n), s);
}

SolveStats solve_cg_ws(const Operator& A, Multigrid* mg, double* x, const double* b, double rel_tol, int max_iter,
                       bool flexible, const CgWorkspace& ws, hipStream_t s) {
  if (A.precision != MF_F64) throw std::runtime_error("the outer solver runs in FP64");
  const int64_t n = A.m();
  double *r = ws.r, *z = ws.z, *p = ws.p, *q = ws.q, *partial = ws.partial, *scalar = ws.scalar;
  auto dot = [&](const double* u, const double* v) {
    blas_dot<double>(n, u, v, partial, scalar, s);
    double h;
    SMF_HIP(hipMemcpyAsync(&h, scalar, sizeof(double), hipMemcpyDeviceToHost, s));
    SMF_HIP(hipStreamSynchronize(s));
    return h;
  };
  auto precond = [&](double* zz, const double* rr) {
    if (mg) mg_vmult(*mg, zz, rr, s);
    else SMF_HIP(hipMemcpyAsync(zz, rr, n * 8, hipMemcpyDeviceToDevice, s));
  };
  SolveStats st;
  blas_fill<double>(n, 0.0, x, s);
  SMF_HIP(hipMemcpyAsync(r, b, n * 8, hipMemcpyDeviceToDevice, s));
  st.b_norm = std::sqrt(dot(b, b));
  st.r0 = st.b_norm;
  st.rn = st.r0;
  const double tol = rel_tol * st.b_norm;
  if (st.r0 <= tol) {
    st.converged = 1;
    return st;
  }
  // alpha, beta and the dot products stay in device memory (S); the host
  // reads |r|^2 once per iteration for the stopping test
  double* Sd = ws.S;
  blas_fill<double>(kS_N, 0.0, Sd, s);  // a fill kernel (no memset node under graph capture)
  precond(z, r);
  SMF_HIP(hipMemcpyAsync(p, z, n * 8, hipMemcpyDeviceToDevice, s));
  blas_dot<double>(n, r, z, partial, Sd + kS_RZ, s);
  for (int it = 1; it <= max_iter; ++it) {
    launch_op(A, 0, q, p, s);
    blas_dot<double>(n, p, q, partial, Sd + kS_PQ, s);
    blas_krylov_scalars(Sd, 0, s);                             // alpha = rz / pq
    blas_axpby_dev(n, Sd, kS_ALPHA, 1.0, p, -1, x, s);         // x += alpha p
    blas_axpby_dev(n, Sd, kS_ALPHA, -1.0, q, -1, r, s);        // r -= alpha q
    blas_dot<double>(n, r, r, partial, Sd + kS_RR, s);
    double rr;
    SMF_HIP(hipMemcpyAsync(&rr, Sd + kS_RR, sizeof(double), hipMemcpyDeviceToHost, s));
    SMF_HIP(hipStreamSynchronize(s));
    st.rn = std::sqrt(rr);
    st.iterations = it;
    if (st.rn <= tol || !std::isfinite(st.rn)) {
      st.converged = st.rn <= tol;
      break;
    }
    precond(z, r);
    blas_dot<double>(n, r, z, partial, Sd + kS_RZN, s);
    // flexible CG (Notay 2000): beta = z_new.(r_new - r_old)/rz_old with
    // r_new - r_old = -alpha q
    if (flexible) blas_dot<double>(n, z, q, partial, Sd + kS_ZQ, s);
    blas_krylov_scalars(Sd, flexible ? 2 : 1, s);
    blas_axpby_dev(n, Sd, -1, 1.0, z, kS_BETA, p, s);  // p = z + beta p
  }
  SMF_HIP(hipStreamSynchronize(s));
  return st;
}

void assemble_rhs(const Operator& op, int kind, uint64_t seed, int n_features, double corr_length,
                  std::vector<double>& b) {
  const bool dg = op.kind == MF_DG_SIPG;
  if (op.kind != MF_CG_POISSON && !dg) throw std::runtime_error("right-hand side assembly is for scalar CG and DG operators");
  if (dg && op.dg.curved) throw std::runtime_error("right-hand side assembly: affine DG cells only");
  const CGPlan& pl = op.cg;
  const Mesh& m = op.mesh;
  const int nl = dg ? op.dg.n_loc : pl.n_loc;
  const int64_t ncell = dg ? op.dg.n_cells : pl.n_cells;
  const int deg = dg ? op.dg.p : pl.p;
  // per cell x0 and J (row-major), internal order, in the operator's local vertex order
  std::vector<double> geo(12 * ncell);
  for (int64_t k = 0; k < ncell; ++k) {
    const int32_t* v = dg ? &op.dg.cell_verts[4 * k] : &m.cells[4 * pl.cell_perm[k]];
    for (int a = 0; a < 3; ++a) {
      geo[12 * k + a] = m.x[3 * v[0] + a];
      for (int c = 0; c < 3; ++c) geo[12 * k + 3 + 3 * a + c] = m.x[3 * v[c + 1] + a] - m.x[3 * v[0] + a];
    }
  }
  const Quadrature q = tet_rule(6);
  const Lagrange fe(deg);
  std::vector<double> qtab(q.size() * (4 + nl)), val(nl);
  for (int k = 0; k < q.size(); ++k) {
    fe.eval(q.pts[k].data(), val.data(), nullptr);
    for (int a = 0; a < 3; ++a) qtab[k * (4 + nl) + a] = q.pts[k][a];
    qtab[k * (4 + nl) + 3] = q.w[k];
    for (int i = 0; i < nl; ++i) qtab[k * (4 + nl) + 4 + i] = val[i];
  }
  std::vector<double> feat;
  if (kind == 1) {
    // omega ~ N(0, I / l^2) by Box-Muller from splitmix64, phase ~ U[0, 2 pi)
    SplitMix64 rng(seed);
    auto unif = [&]() { return ((rng.next() >> 11) + 0.5) * 0x1.0p-53; };
    for (int f = 0; f < n_features; ++f) {
      for (int a = 0; a < 3; ++a) {
        const double u1 = unif(), u2 = unif();
        feat.push_back(std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2) / corr_length);
      }
      feat.push_back(2.0 * M_PI * unif());
    }
  } else {
    feat.assign(4, 0.0);
  }
  DevBuf dgeo, dq, dfeat, dout;
  dgeo.upload(geo);
  dq.upload(qtab);
  dfeat.upload(feat);
  dout.alloc(ncell * nl * sizeof(double));
  launch_rhs_cells(ncell, nl, q.size(), dgeo.as<double>(), dq.as<double>(), kind, kind == 1 ? n_features : 0,
                   dfeat.as<double>(), dout.as<double>(), nullptr);
  std::vector<double> local(ncell * nl);
  SMF_HIP(hipMemcpy(local.data(), dout.p, local.size() * sizeof(double), hipMemcpyDeviceToHost));
  if (dg) {  // cell-contiguous DoFs; homogeneous boundary data adds no Nitsche term
    b = local;
    return;
  }
  b.assign(pl.n_dofs, 0.0);
  for (int64_t k = 0; k < pl.n_cells; ++k)
    for (int i = 0; i < nl; ++i) {
      const int64_t d = pl.cell_dofs[k * nl + i];
      if (d < pl.n_free) b[d] += local[k * nl + i];
    }
}

void exact_solution(const Operator& op, std::vector<double>& u) {
  const double a0 = 3.0 * M_PI;
  if (op.kind == MF_DG_SIPG) {
    const DGPlan& dp = op.dg;
    u.resize(dp.n_dofs);
    for (int64_t d = 0; d < dp.n_dofs; ++d) {
      const double* X = &dp.dof_points[3 * d];
      u[d] = std::sin(a0 * X[0]) * std::sin(a0 * X[1]) * std::sin(a0 * X[2]);
    }
    return;
  }
  const CGPlan& pl = op.cg;
  u.assign(pl.n_dofs, 0.0);
  const double a = 3.0 * M_PI;
  for (int64_t d = 0; d < pl.n_dofs; ++d) {
    const double* X = &pl.dof_points[3 * d];
    u[d] = d < pl.n_free ? std::sin(a * X[0]) * std::sin(a * X[1]) * std::sin(a * X[2]) : 0.0;
  }
}

}  // namespace smf
