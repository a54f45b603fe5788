// Smoothed-aggregation AMG (Vanek, Mandel and Brezina, "Algebraic multigrid
// by smoothed aggregation for second and fourth order elliptic problems",
// Computing 56, 1996), the coarse solver in place of the reference's Trilinos
// ML V-cycle (CG/cells_batched/application_hp.cc:1684-1731, 1918-1930).
//
// Set-up on the host, per level:
//   * coupling (i, j) is strong when |a_ij| >= theta sqrt(|a_ii a_jj|)
//     (ML's aggregation threshold); rows without strong couplings (identity
//     rows of constrained DoFs among them) are isolated and belong to no
//     aggregate;
//   * aggregation in three passes over the rows in order: (1) a row whose
//     strong neighbours are all unaggregated starts an aggregate with them;
//     (2) a remaining row joins the pass-1 aggregate of its strongest strong
//     neighbour; (3) the rest start aggregates with their unaggregated strong
//     neighbours;
//   * tentative prolongator T: column j = indicator of aggregate j scaled to
//     unit norm (the constants, the near-null space of the Laplacian);
//   * P = (I - omega D^-1 A) T with omega = 4 / (3 lambda_max(D^-1 A));
//   * A_c = P^T A P (Galerkin, two sparse products).
// Coarsening stops at max_dense rows (dense FP64 inverse) or when a level
// shrinks by less than 10%.
//
// Solve on the device: a V-cycle with Chebyshev-Jacobi smoothing (the
// geometric levels' degree, range 30, the same recurrence) and CSR-stream
// products (row-per-thread CSR for a level whose rows exceed a row block).
#include "amg.h"

#include <algorithm>
#include <cmath>
#include <stdexcept>
#include <string>

namespace smf {

std::vector<double> dense_inverse_spd(std::vector<double> A, int64_t n);

namespace {

Csr csr_transpose(const Csr& A) {
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

// C = A B (Gustavson, columns ascending per row, sums in A's entry order)
Csr csr_multiply(const Csr& A, const Csr& B) {
  if (A.ncols != B.nrows) throw std::runtime_error("AMG: product of incompatible matrices");
  Csr C;
  C.nrows = A.nrows;
  C.ncols = B.ncols;
  C.ptr.assign(C.nrows + 1, 0);
  std::vector<int64_t> mark(B.ncols, -1);
  std::vector<double> acc(B.ncols, 0.0);
  std::vector<int32_t> cols;
  for (int64_t i = 0; i < A.nrows; ++i) {
    cols.clear();
    for (int64_t k = A.ptr[i]; k < A.ptr[i + 1]; ++k) {
      const int32_t j = A.col[k];
      const double a = A.val[k];
      for (int64_t m = B.ptr[j]; m < B.ptr[j + 1]; ++m) {
        const int32_t c = B.col[m];
        if (mark[c] != i) {
          mark[c] = i;
          acc[c] = 0.0;
          cols.push_back(c);
        }
        acc[c] += a * B.val[m];
      }
    }
    std::sort(cols.begin(), cols.end());
    for (int32_t c : cols) {
      C.col.push_back(c);
      C.val.push_back(acc[c]);
    }
    C.ptr[i + 1] = (int64_t)C.col.size();
  }
  return C;
}

std::vector<double> diagonal(const Csr& A) {
  std::vector<double> d(A.nrows, 0.0);
  for (int64_t i = 0; i < A.nrows; ++i)
    for (int64_t k = A.ptr[i]; k < A.ptr[i + 1]; ++k)
      if (A.col[k] == i) d[i] = A.val[k];
  return d;
}

// largest eigenvalue of D^-1 A by power iteration from a fixed start vector
double lambda_max_jacobi(const Csr& A, const std::vector<double>& d) {
  const int64_t n = A.nrows;
  std::vector<double> v(n), w(n);
  for (int64_t i = 0; i < n; ++i) v[i] = 1.0 + 0.5 * std::sin(0.7 * (double)i);
  double lam = 0;
  for (int it = 0; it < 30; ++it) {
    double nv = 0;
    for (double x : v) nv += x * x;
    nv = std::sqrt(nv);
    if (!(nv > 0)) break;
    for (auto& x : v) x /= nv;
    double num = 0;
    for (int64_t i = 0; i < n; ++i) {
      double s = 0;
      for (int64_t k = A.ptr[i]; k < A.ptr[i + 1]; ++k) s += A.val[k] * v[A.col[k]];
      w[i] = s / d[i];
      num += v[i] * w[i];
    }
    lam = num;
    v.swap(w);
  }
  return lam;
}

// aggregate of every row (-1: isolated); returns the number of aggregates.
// Coupling (i, j) is strong when |a_ij| >= theta sqrt(|a_ii a_jj|) (nonzero
// for theta = 0); a row without strong couplings is isolated.
int32_t aggregate(const Csr& A, const std::vector<double>& d, double theta, std::vector<int32_t>& agg) {
  const int64_t n = A.nrows;
  agg.assign(n, -1);
  std::vector<uint8_t> strong(A.col.size(), 0);
  for (int64_t i = 0; i < n; ++i)
    for (int64_t k = A.ptr[i]; k < A.ptr[i + 1]; ++k) {
      const int32_t j = A.col[k];
      strong[k] = j != i && A.val[k] != 0.0 && std::fabs(A.val[k]) >= theta * std::sqrt(std::fabs(d[i] * d[j]));
    }
  std::vector<uint8_t> iso(n, 1);
  for (int64_t i = 0; i < n; ++i)
    for (int64_t k = A.ptr[i]; k < A.ptr[i + 1]; ++k)
      if (strong[k]) iso[i] = 0;
  int32_t na = 0;
  for (int64_t i = 0; i < n; ++i) {  // pass 1
    if (agg[i] >= 0 || iso[i]) continue;
    bool fresh = true;
    for (int64_t k = A.ptr[i]; k < A.ptr[i + 1] && fresh; ++k)
      if (strong[k] && agg[A.col[k]] >= 0) fresh = false;
    if (!fresh) continue;
    agg[i] = na;
    for (int64_t k = A.ptr[i]; k < A.ptr[i + 1]; ++k)
      if (strong[k]) agg[A.col[k]] = na;
    ++na;
  }
  const std::vector<int32_t> agg1 = agg;
  for (int64_t i = 0; i < n; ++i) {  // pass 2
    if (agg[i] >= 0 || iso[i]) continue;
    int32_t best = -1;
    double bv = 0;
    for (int64_t k = A.ptr[i]; k < A.ptr[i + 1]; ++k) {
      const int32_t j = A.col[k];
      if (strong[k] && agg1[j] >= 0 && std::fabs(A.val[k]) > bv) {
        bv = std::fabs(A.val[k]);
        best = agg1[j];
      }
    }
    agg[i] = best;
  }
  for (int64_t i = 0; i < n; ++i) {  // pass 3
    if (agg[i] >= 0 || iso[i]) continue;
    agg[i] = na;
    for (int64_t k = A.ptr[i]; k < A.ptr[i + 1]; ++k) {
      const int32_t j = A.col[k];
      if (strong[k] && agg[j] < 0 && !iso[j]) agg[j] = na;
    }
    ++na;
  }
  return na;
}

template <typename F>
void cheby(Amg& amg, AmgLevel& L, bool zero_start, hipStream_t s, F&& apply) {
  const int64_t n = L.A.nrows;
  const double lmax = L.lambda_max, lmin = lmax / amg.smoothing_range;
  const double theta = 0.5 * (lmax + lmin), delta = 0.5 * (lmax - lmin), sigma = theta / delta;
  double rho = 1.0 / sigma;
  double* x = L.x.as<double>();
  for (int k = 0; k < amg.smoother_degree; ++k) {
    double c1 = 0, c2;
    if (k == 0) {
      c2 = 1.0 / theta;
      if (zero_start) {
        blas_cheby_zero<double>(n, c2, L.dinv.as<double>(), L.b.as<double>(), L.d.as<double>(), x, s);
        continue;
      }
    } else {
      const double rn = 1.0 / (2.0 * sigma - rho);
      c1 = rn * rho;
      c2 = 2.0 * rn / delta;
      rho = rn;
    }
    apply(L.t.as<double>(), x);
    blas_cheby_step<double>(n, c1, c2, L.dinv.as<double>(), L.b.as<double>(), L.t.as<double>(), L.d.as<double>(), x,
                            k == 0, s);
  }
}

void vcycle_level(Amg& amg, size_t l, hipStream_t s) {
  AmgLevel& L = amg.lv[l];
  const int64_t n = L.A.nrows;
  if (l + 1 == amg.lv.size()) {
    blas_dense_gemv((int)n, (int)n, amg.d_inv.as<double>(), L.b.as<double>(), L.x.as<double>(), s);
    return;
  }
  auto apply = [&](double* y, const double* x) {
    if (L.M.on_device) launch_spmv(L.M, y, x, s);
    else blas_csr<double>(n, L.A.d_ptr.as<int64_t>(), L.A.d_col.as<int32_t>(), L.A.d_val.as<double>(), x, y, false, s);
  };
  AmgLevel& C = amg.lv[l + 1];
  cheby(amg, L, true, s, apply);
  apply(L.r.as<double>(), L.x.as<double>());
  blas_axpby<double>(n, 1.0, L.b.as<double>(), -1.0, L.r.as<double>(), s);  // r = b - A x
  if (L.Rm.on_device)
    launch_spmv(L.Rm, C.b.as<double>(), L.r.as<double>(), s);
  else
    blas_csr<double>(L.R.nrows, L.R.d_ptr.as<int64_t>(), L.R.d_col.as<int32_t>(), L.R.d_val.as<double>(),
                     L.r.as<double>(), C.b.as<double>(), false, s);
  vcycle_level(amg, l + 1, s);
  blas_csr<double>(L.P.nrows, L.P.d_ptr.as<int64_t>(), L.P.d_col.as<int32_t>(), L.P.d_val.as<double>(),
                   C.x.as<double>(), L.x.as<double>(), true, s);
  cheby(amg, L, false, s, apply);
}

}  // namespace

std::unique_ptr<Amg> build_amg(const Csr& A0, int max_dense, double strength) {
  auto amg = std::make_unique<Amg>();
  amg->strength = strength;
  amg->lv.emplace_back();
  amg->lv.back().A.nrows = A0.nrows;
  amg->lv.back().A.ncols = A0.ncols;
  amg->lv.back().A.ptr = A0.ptr;
  amg->lv.back().A.col = A0.col;
  amg->lv.back().A.val = A0.val;
  while (amg->lv.back().A.nrows > max_dense) {
    AmgLevel& L = amg->lv.back();
    const Csr& A = L.A;
    const std::vector<double> d = diagonal(A);
    for (double v : d)
      if (!(v > 0)) throw std::runtime_error("AMG: matrix with a non-positive diagonal entry");
    L.lambda_max = lambda_max_jacobi(A, d);
    std::vector<int32_t> agg;
    // a level that barely coarsens retries with a weaker strength threshold
    // (theta / 4 per retry, 0 below 1e-3: then every nonzero coupling is strong)
    double theta = amg->strength;
    int32_t na = aggregate(A, d, theta, agg);
    while ((na == 0 || na > 0.9 * (double)A.nrows) && theta > 0.0) {
      theta = theta > 1e-3 ? 0.25 * theta : 0.0;
      na = aggregate(A, d, theta, agg);
    }
    if (na == 0 || na > 0.9 * (double)A.nrows) break;
    // tentative prolongator: unit-norm aggregate indicators
    std::vector<int64_t> size(na, 0);
    for (int32_t a : agg)
      if (a >= 0) ++size[a];
    Csr T;
    T.nrows = A.nrows;
    T.ncols = na;
    T.ptr.assign(A.nrows + 1, 0);
    for (int64_t i = 0; i < A.nrows; ++i) {
      if (agg[i] >= 0) {
        T.col.push_back(agg[i]);
        T.val.push_back(1.0 / std::sqrt((double)size[agg[i]]));
      }
      T.ptr[i + 1] = (int64_t)T.col.size();
    }
    // P = (I - omega D^-1 A) T
    const double omega = 4.0 / (3.0 * L.lambda_max);
    Csr AT = csr_multiply(A, T);
    Csr P;
    P.nrows = A.nrows;
    P.ncols = na;
    P.ptr.assign(A.nrows + 1, 0);
    for (int64_t i = 0; i < A.nrows; ++i) {
      // merge row i of T (one entry at most) with -omega/d_i row i of A T
      int64_t k = AT.ptr[i];
      const int64_t k1 = AT.ptr[i + 1];
      const bool has_t = T.ptr[i + 1] > T.ptr[i];
      const int32_t tc = has_t ? T.col[T.ptr[i]] : -1;
      const double tv = has_t ? T.val[T.ptr[i]] : 0.0;
      bool t_done = !has_t;
      for (; k < k1; ++k) {
        const int32_t c = AT.col[k];
        if (!t_done && tc < c) {
          P.col.push_back(tc);
          P.val.push_back(tv);
          t_done = true;
        }
        double v = -omega / d[i] * AT.val[k];
        if (!t_done && tc == c) {
          v += tv;
          t_done = true;
        }
        P.col.push_back(c);
        P.val.push_back(v);
      }
      if (!t_done) {
        P.col.push_back(tc);
        P.val.push_back(tv);
      }
      P.ptr[i + 1] = (int64_t)P.col.size();
    }
    Csr R = csr_transpose(P);
    Csr Ac = csr_multiply(R, csr_multiply(A, P));
    L.P = std::move(P);
    L.R = std::move(R);
    amg->lv.emplace_back();
    amg->lv.back().A = std::move(Ac);
  }
  AmgLevel& last = amg->lv.back();
  const int64_t nd = last.A.nrows;
  if (nd > (int64_t)max_dense)
    throw std::runtime_error("AMG: coarsening stalled at " + std::to_string(nd) + " rows, above the dense coarse "
                             "solver's limit of " + std::to_string(max_dense) + " (max_coarse_dofs)");
  std::vector<double> dense(nd * nd, 0.0);
  for (int64_t i = 0; i < nd; ++i)
    for (int64_t k = last.A.ptr[i]; k < last.A.ptr[i + 1]; ++k) dense[i * nd + last.A.col[k]] = last.A.val[k];
  const std::vector<double> inv = dense_inverse_spd(std::move(dense), nd);
  amg->n_dense = (int)nd;
  amg->d_inv.upload(inv);
  // device data
  for (auto& L : amg->lv) {
    const int64_t n = L.A.nrows;
    L.M.precision = MF_F64;
    L.M.n_rows = n;
    L.M.ptr = L.A.ptr;
    L.M.col = L.A.col;
    L.M.val = L.A.val;
    if (n > 0 && make_row_blocks(L.M)) upload_matrix(L.M);
    else L.A.upload(MF_F64);
    if (!L.P.ptr.empty()) {
      L.P.upload(MF_F64);
      L.Rm.precision = MF_F64;
      L.Rm.n_rows = L.R.nrows;
      L.Rm.ptr = L.R.ptr;
      L.Rm.col = L.R.col;
      L.Rm.val = L.R.val;
      if (L.R.nrows > 0 && make_row_blocks(L.Rm)) upload_matrix(L.Rm);
      else L.R.upload(MF_F64);
    }
    std::vector<double> dinv = diagonal(L.A);
    for (auto& v : dinv) v = 1.0 / v;
    L.dinv.upload(dinv);
    L.lambda_max *= 1.1;
    for (DevBuf* b : {&L.x, &L.b, &L.r, &L.d, &L.t}) b->alloc(std::max<int64_t>(n, 1) * sizeof(double));
  }
  return amg;
}

void amg_vcycle(Amg& amg, hipStream_t s) { vcycle_level(amg, 0, s); }

}  // namespace smf
