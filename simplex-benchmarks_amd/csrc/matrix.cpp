// Sparse system matrix of a CG operator: pattern and element-matrix
// assembly on the host, CSR upload for the device product.
//
// Follows Operator::get_system_matrix (CG/cells_batched/application_hp.cc:
// 1411-1438: DoFTools::make_sparsity_pattern with the constraints, then
// MatrixFreeTools::compute_matrix over the cell integral) and the timing
// comparison of CG/cells_batched/application.cc:1944-1988.  Constrained
// couplings are left out of the pattern (their values are zero), so a
// Dirichlet row holds its unit diagonal only.
#include "matrix.h"

#include <algorithm>
#include <stdexcept>
#include <string>

namespace smf {

namespace {

// element-matrix row i of cell k: out[j] = sum_q w_q G_iq . (M G_jq), M the
// cell's metric |det J| J^-1 J^-T (the cell integral of the kernels)
void element_row(const CGPlan& pl, int64_t k, int i, double* out) {
  const int nl = pl.n_loc, nq = pl.n_q;
  const double* G = pl.shape.data();
  const double* W = G + (size_t)nl * nq * 3;
  double m[6];
  if (!pl.curved)
    for (int t = 0; t < 6; ++t) m[t] = pl.metric[(size_t)t * pl.n_cells + k];
  double g[64 * 3];  // M G_iq per quadrature point (n_q <= 64)
  for (int q = 0; q < nq; ++q) {
    const double* a = &G[(q * nl + i) * 3];
    double w = W[q];
    if (pl.curved) {  // per-point metric, weight included
      for (int t = 0; t < 6; ++t) m[t] = pl.qmetric[((size_t)q * 6 + t) * pl.n_cells + k];
      w = 1.0;
    }
    g[3 * q + 0] = w * (m[0] * a[0] + m[1] * a[1] + m[2] * a[2]);
    g[3 * q + 1] = w * (m[1] * a[0] + m[3] * a[1] + m[4] * a[2]);
    g[3 * q + 2] = w * (m[2] * a[0] + m[4] * a[1] + m[5] * a[2]);
  }
  for (int j = 0; j < nl; ++j) {
    double s = 0;
    for (int q = 0; q < nq; ++q) {
      const double* b = &G[(q * nl + j) * 3];
      s += g[3 * q] * b[0] + g[3 * q + 1] * b[1] + g[3 * q + 2] * b[2];
    }
    out[j] = s;
  }
}

}  // namespace

void assemble_cg_matrix(const Operator& op, SysMatrix& A) {
  if (op.is_dg()) throw std::runtime_error("system matrix: CG operators only (the reference assembles CG matrices)");
  const CGPlan& pl = op.cg;
  const int nl = pl.n_loc, nc = op.ncomp;
  const int64_t n = pl.n_dofs, nf = pl.n_free;
  if (pl.n_q > 64) throw std::runtime_error("system matrix: more than 64 quadrature points");
  if (n * nc >= (int64_t(1) << 31)) throw std::runtime_error("system matrix: more than 2^31 - 1 rows");

  // free DoF -> its (cell, node) slots, cells ascending
  std::vector<int64_t> aptr(nf + 1, 0);
  for (int64_t k = 0; k < pl.n_cells; ++k)
    for (int i = 0; i < nl; ++i) {
      const int64_t d = pl.cell_dofs[k * nl + i];
      if (d < nf) ++aptr[d + 1];
    }
  for (int64_t d = 0; d < nf; ++d) aptr[d + 1] += aptr[d];
  std::vector<int64_t> adj(aptr[nf]);
  {
    std::vector<int64_t> pos(aptr.begin(), aptr.end() - 1);
    for (int64_t k = 0; k < pl.n_cells; ++k)
      for (int i = 0; i < nl; ++i) {
        const int64_t d = pl.cell_dofs[k * nl + i];
        if (d < nf) adj[pos[d]++] = k * nl + i;
      }
  }
  // Nitsche boundary rows by DoF
  std::vector<int32_t> brow(pl.nitsche ? n : 0, -1);
  for (size_t r = 0; r < pl.bnd_rows.size(); ++r) brow[pl.bnd_rows[r]] = (int32_t)r;

  // pattern: free row -> sorted free columns of its cells (+ boundary columns)
  auto row_cols = [&](int64_t r, std::vector<int32_t>& cols) {
    cols.clear();
    for (int64_t a = aptr[r]; a < aptr[r + 1]; ++a) {
      const int64_t k = adj[a] / nl;
      for (int j = 0; j < nl; ++j) {
        const int64_t d = pl.cell_dofs[k * nl + j];
        if (d < nf) cols.push_back((int32_t)d);
      }
    }
    if (pl.nitsche && brow[r] >= 0)
      for (int32_t e = pl.bnd_ptr[brow[r]]; e < pl.bnd_ptr[brow[r] + 1]; ++e) cols.push_back(pl.bnd_col[e]);
    std::sort(cols.begin(), cols.end());
    cols.erase(std::unique(cols.begin(), cols.end()), cols.end());
  };
  std::vector<int64_t> sptr(n + 1, 0);
#pragma omp parallel
  {
    std::vector<int32_t> cols;
#pragma omp for schedule(dynamic, 1024)
    for (int64_t r = 0; r < n; ++r) {
      if (r < nf) {
        row_cols(r, cols);
        sptr[r + 1] = (int64_t)cols.size();
      } else {
        sptr[r + 1] = 1;
      }
    }
  }
  for (int64_t r = 0; r < n; ++r) sptr[r + 1] += sptr[r];
  const int64_t snnz = sptr[n];
  std::vector<int32_t> scol(snnz);
  std::vector<double> sval(snnz, 0.0);
#pragma omp parallel
  {
    std::vector<int32_t> cols;
    double ke[64];
#pragma omp for schedule(dynamic, 1024)
    for (int64_t r = 0; r < n; ++r) {
      const int64_t o = sptr[r];
      if (r >= nf) {
        scol[o] = (int32_t)r;
        sval[o] = 1.0;
        continue;
      }
      row_cols(r, cols);
      std::copy(cols.begin(), cols.end(), scol.begin() + o);
      int32_t* c0 = scol.data() + o;
      int32_t* c1 = scol.data() + sptr[r + 1];
      double* v = sval.data() + o;
      for (int64_t a = aptr[r]; a < aptr[r + 1]; ++a) {
        const int64_t k = adj[a] / nl;
        element_row(pl, k, (int)(adj[a] % nl), ke);
        for (int j = 0; j < nl; ++j) {
          const int64_t d = pl.cell_dofs[k * nl + j];
          if (d < nf) v[std::lower_bound(c0, c1, (int32_t)d) - c0] += ke[j];
        }
      }
      if (pl.nitsche && brow[r] >= 0)
        for (int32_t e = pl.bnd_ptr[brow[r]]; e < pl.bnd_ptr[brow[r] + 1]; ++e)
          v[std::lower_bound(c0, c1, pl.bnd_col[e]) - c0] += pl.bnd_val[e];
    }
  }

  // components: block diagonal, component c in rows / columns [c n, (c+1) n)
  // (FEM rows have at most a few hundred entries, far below a row block)
  A.precision = op.precision;
  A.n_rows = n * nc;
  if (nc == 1) {  // scalar: the arrays as they are (no second copy)
    A.ptr = std::move(sptr);
    A.col = std::move(scol);
    A.val = std::move(sval);
  } else {
    A.ptr.assign(A.n_rows + 1, 0);
    A.col.resize(snnz * nc);
    A.val.resize(snnz * nc);
    for (int c = 0; c < nc; ++c) {
      for (int64_t r = 0; r <= n; ++r) A.ptr[c * n + r] = c * snnz + sptr[r];
      for (int64_t e = 0; e < snnz; ++e) {
        A.col[c * snnz + e] = (int32_t)(c * n + scol[e]);
        A.val[c * snnz + e] = sval[e];
      }
    }
  }
  if (!make_row_blocks(A))
    throw std::runtime_error("system matrix: a row has more than " + std::to_string(kSpmvBlockEntries) + " entries");
}

bool make_row_blocks(SysMatrix& A) {
  // consecutive rows, greedily, up to kSpmvBlockEntries entries and
  // kSpmvBlockRows rows per block
  A.blk.assign(1, 0);
  int64_t e = 0, rows = 0;
  for (int64_t r = 0; r < A.n_rows; ++r) {
    const int64_t len = A.ptr[r + 1] - A.ptr[r];
    if (len > kSpmvBlockEntries) {
      A.blk.clear();
      return false;
    }
    if (e + len > kSpmvBlockEntries || rows == kSpmvBlockRows) {
      A.blk.push_back((int32_t)r);
      e = rows = 0;
    }
    e += len;
    ++rows;
  }
  if (A.n_rows > 0) A.blk.push_back((int32_t)A.n_rows);
  return true;
}

bool upload_blocked(const std::vector<int64_t>& ptr, const std::vector<int32_t>& col, const std::vector<double>& val,
                    int precision, SysMatrix& A) {
  A.precision = precision;
  A.n_rows = ptr.empty() ? 0 : (int64_t)ptr.size() - 1;
  A.ptr = ptr;
  A.col = col;
  A.val = val;
  if (A.n_rows == 0 || !make_row_blocks(A)) return false;
  upload_matrix(A);
  return true;
}

void upload_matrix(SysMatrix& A) {
  A.d_ptr.upload(A.ptr);
  A.d_col.upload(A.col);
  A.d_blk.upload(A.blk);
  if (A.precision == MF_F64) {
    A.d_val.upload(A.val);
  } else {
    std::vector<float> v(A.val.begin(), A.val.end());
    A.d_val.upload(v);
  }
  A.on_device = true;
}

}  // namespace smf
