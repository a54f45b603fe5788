/* CPU oracle for the simplexmf operators -- TEST INFRASTRUCTURE ONLY.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg load
 * this file's library, as the checker.  See oracle.h.
 *
 * Pinning (tests/test_oracle_*.py): the basis by the Kronecker property at the
 * nodes, partition of unity and closed-form P1/P2 values; the quadrature by
 * exact monomial moments; the CG operator by the P1 closed-form element
 * stiffness, P2 integrals of the basis, symmetry, linear-function consistency,
 * the energy identity, the matrix-based CSR product and brute force on one
 * hexahedron.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define MAXL 20   /* max local nodes (P3) */
#define MAXQ 216  /* max quadrature points (6^3) */

/* ------------------------------------------------------------------------ */
/* nodal basis from the node-position convention (independent of closed forms)
 * Node order: 4 vertices; per edge (0,1)(0,2)(0,3)(1,2)(1,3)(2,3) its p-1
 * nodes from the lower to the higher vertex; for p=3 one node per face,
 * face f opposite vertex f.  (Convention of the library, restated.)        */

static const double VX[4][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
static const int EDGES[6][2] = {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {1, 3}, {2, 3}};

static int nloc_of(int p) { return (p + 1) * (p + 2) * (p + 3) / 6; }

static void node_points(int p, double (*x)[3]) {
  int n = 0;
  for (int k = 0; k < 4; ++k, ++n)
    for (int d = 0; d < 3; ++d) x[n][d] = VX[k][d];
  for (int e = 0; e < 6; ++e)
    for (int s = 0; s < p - 1; ++s, ++n)
      for (int d = 0; d < 3; ++d) {
        const double a = VX[EDGES[e][0]][d], b = VX[EDGES[e][1]][d];
        x[n][d] = a + (s + 1.0) / p * (b - a);
      }
  if (p == 3)
    for (int f = 0; f < 4; ++f, ++n)
      for (int d = 0; d < 3; ++d) {
        double s = 0;
        for (int k = 0; k < 4; ++k)
          if (k != f) s += VX[k][d];
        x[n][d] = s / 3.0;
      }
}

/* monomial exponents with total degree <= p */
static int monomials(int p, int (*e)[3]) {
  int n = 0;
  for (int a = 0; a <= p; ++a)
    for (int b = 0; a + b <= p; ++b)
      for (int c = 0; a + b + c <= p; ++c) {
        e[n][0] = a;
        e[n][1] = b;
        e[n][2] = c;
        ++n;
      }
  return n;
}

static long double ipow(long double x, int k) {
  long double r = 1.0L;
  for (int i = 0; i < k; ++i) r *= x;
  return r;
}

/* coefficient matrices C[p][k][i]: phi_i = sum_k C[k][i] m_k, solved and
 * evaluated in extended precision (the P3 monomial Vandermonde matrix has a
 * condition number near 1e4) */
static long double g_coef[4][MAXL][MAXL];
static int g_coef_ready[4];

static int build_coef(int p) {
  if (g_coef_ready[p]) return 0;
  const int n = nloc_of(p);
  double x[MAXL][3];
  int e[MAXL][3];
  node_points(p, x);
  if (monomials(p, e) != n) return -1;
  /* Gauss-Jordan on [V | I] with V[j][k] = m_k(x_j) */
  long double a[MAXL][2 * MAXL];
  for (int j = 0; j < n; ++j) {
    for (int k = 0; k < n; ++k) a[j][k] = ipow(x[j][0], e[k][0]) * ipow(x[j][1], e[k][1]) * ipow(x[j][2], e[k][2]);
    for (int k = 0; k < n; ++k) a[j][n + k] = (j == k) ? 1.0 : 0.0;
  }
  for (int col = 0; col < n; ++col) {
    int piv = col;
    for (int r = col + 1; r < n; ++r)
      if (fabsl(a[r][col]) > fabsl(a[piv][col])) piv = r;
    if (fabsl(a[piv][col]) < 1e-14L) return -1;
    for (int k = 0; k < 2 * n; ++k) {
      long double t = a[col][k];
      a[col][k] = a[piv][k];
      a[piv][k] = t;
    }
    const long double inv = 1.0L / a[col][col];
    for (int k = 0; k < 2 * n; ++k) a[col][k] *= inv;
    for (int r = 0; r < n; ++r)
      if (r != col) {
        const long double f = a[r][col];
        if (f != 0.0L)
          for (int k = 0; k < 2 * n; ++k) a[r][k] -= f * a[col][k];
      }
  }
  /* V^-1 = C: coefficient of monomial k in basis i is (V^-1)[k][i] */
  for (int k = 0; k < n; ++k)
    for (int i = 0; i < n; ++i) g_coef[p][k][i] = a[k][n + i];
  g_coef_ready[p] = 1;
  return 0;
}

int orc_basis(int p, const double* x, double* values, double* grads) {
  if (p < 1 || p > 3) return -1;
#pragma omp critical(orc_coef)
  build_coef(p);
  const int n = nloc_of(p);
  int e[MAXL][3];
  monomials(p, e);
  long double mv[MAXL], mg[MAXL][3];
  const long double x0 = x[0], x1 = x[1], x2 = x[2];
  for (int k = 0; k < n; ++k) {
    const int a = e[k][0], b = e[k][1], c = e[k][2];
    mv[k] = ipow(x0, a) * ipow(x1, b) * ipow(x2, c);
    mg[k][0] = a ? a * ipow(x0, a - 1) * ipow(x1, b) * ipow(x2, c) : 0.0L;
    mg[k][1] = b ? b * ipow(x0, a) * ipow(x1, b - 1) * ipow(x2, c) : 0.0L;
    mg[k][2] = c ? c * ipow(x0, a) * ipow(x1, b) * ipow(x2, c - 1) : 0.0L;
  }
  for (int i = 0; i < n; ++i) {
    long double v = 0, g0 = 0, g1 = 0, g2 = 0;
    for (int k = 0; k < n; ++k) {
      const long double ck = g_coef[p][k][i];
      v += ck * mv[k];
      g0 += ck * mg[k][0];
      g1 += ck * mg[k][1];
      g2 += ck * mg[k][2];
    }
    if (values) values[i] = (double)v;
    if (grads) {
      grads[3 * i] = (double)g0;
      grads[3 * i + 1] = (double)g1;
      grads[3 * i + 2] = (double)g2;
    }
  }
  return 0;
}

/* ------------------------------------------------------------------------ */
/* quadrature: Gauss-Legendre and the collapsed (Duffy) product rule          */

int orc_gauss01(int n, double* x, double* w) {
  if (n < 1 || n > 32) return -1;
  for (int i = 0; i < n; ++i) {
    double z = cos(M_PI * (i + 0.75) / (n + 0.5)), dp = 1.0;
    for (int it = 0; it < 100; ++it) {
      double p0 = 1.0, p1 = z;
      for (int k = 2; k <= n; ++k) {
        const double p2 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
        p0 = p1;
        p1 = p2;
      }
      if (n == 1) {
        p1 = z;
        p0 = 1.0;
      }
      dp = n * (z * p1 - p0) / (z * z - 1.0);
      const double dz = p1 / dp;
      z -= dz;
      if (fabs(dz) < 1e-16) break;
    }
    {
      double p0 = 1.0, p1 = z;
      for (int k = 2; k <= n; ++k) {
        const double p2 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
        p0 = p1;
        p1 = p2;
      }
      if (n == 1) {
        p1 = z;
        p0 = 1.0;
      }
      dp = n * (z * p1 - p0) / (z * z - 1.0);
    }
    x[i] = 0.5 * (1.0 - z);
    w[i] = 1.0 / ((1.0 - z * z) * dp * dp); /* = 0.5 * 2/((1-z^2) P'^2) */
  }
  return 0;
}

int orc_tet_rule(int n, double* pts, double* w) {
  double x[32], wx[32];
  if (orc_gauss01(n, x, wx)) return -1;
  int k = 0;
  for (int a = 0; a < n; ++a)
    for (int b = 0; b < n; ++b)
      for (int c = 0; c < n; ++c, ++k) {
        const double u = x[a], v = x[b], t = x[c];
        pts[3 * k] = u;
        pts[3 * k + 1] = v * (1.0 - u);
        pts[3 * k + 2] = t * (1.0 - u) * (1.0 - v);
        w[k] = wx[a] * wx[b] * wx[c] * (1.0 - u) * (1.0 - u) * (1.0 - v);
      }
  return 0;
}

/* ------------------------------------------------------------------------ */
/* per-degree tables and geometry                                            */

typedef struct tables {
  int p, nl, nq;
  double w[MAXQ];
  double val[MAXQ][MAXL];
  double grad[MAXQ][MAXL][3];
} tables;

static int make_tables(int p, tables* t) {
  const int n1 = p + 2;
  double pts[3 * MAXQ];
  t->p = p;
  t->nl = nloc_of(p);
  t->nq = n1 * n1 * n1;
  if (orc_tet_rule(n1, pts, t->w)) return -1;
  for (int q = 0; q < t->nq; ++q)
    if (orc_basis(p, pts + 3 * q, t->val[q], &t->grad[q][0][0])) return -1;
  return 0;
}

/* J = [x1-x0 x2-x0 x3-x0]; returns det J, jit = J^-T */
static double geometry(const orc_mesh* m, int64_t c, double jit[3][3]) {
  const int32_t* v = m->cells + 4 * c;
  double J[3][3];
  for (int d = 0; d < 3; ++d)
    for (int k = 0; k < 3; ++k) J[d][k] = m->vertices[3 * v[k + 1] + d] - m->vertices[3 * v[0] + d];
  const double det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
                     J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                     J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
  /* cofactor matrix / det = J^-T */
  jit[0][0] = (J[1][1] * J[2][2] - J[1][2] * J[2][1]) / det;
  jit[0][1] = -(J[1][0] * J[2][2] - J[1][2] * J[2][0]) / det;
  jit[0][2] = (J[1][0] * J[2][1] - J[1][1] * J[2][0]) / det;
  jit[1][0] = -(J[0][1] * J[2][2] - J[0][2] * J[2][1]) / det;
  jit[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) / det;
  jit[1][2] = -(J[0][0] * J[2][1] - J[0][1] * J[2][0]) / det;
  jit[2][0] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) / det;
  jit[2][1] = -(J[0][0] * J[1][2] - J[0][2] * J[1][0]) / det;
  jit[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) / det;
  return det;
}

/* One cell of the reference's do_cell_integral_masked_gather
 * (CG/cells_batched/application.cc:1236-1416): gradients at the quadrature
 * points (1314-1319), affine D_e = w_q JxW jac^T jac with jac = J^-T
 * (1333-1360), integration (1381-1386). */
static void cg_cell(const tables* t, const double jit[3][3], double det, const double* u, double* y) {
  double metric[3][3];
  for (int d = 0; d < 3; ++d)
    for (int f = 0; f < 3; ++f) {
      double s = 0;
      for (int e = 0; e < 3; ++e) s += jit[e][d] * jit[e][f];
      metric[d][f] = s * fabs(det);
    }
  for (int i = 0; i < t->nl; ++i) y[i] = 0.0;
  for (int q = 0; q < t->nq; ++q) {
    double g[3] = {0, 0, 0};
    for (int i = 0; i < t->nl; ++i)
      for (int d = 0; d < 3; ++d) g[d] += t->grad[q][i][d] * u[i];
    double r[3];
    for (int d = 0; d < 3; ++d) r[d] = t->w[q] * (metric[d][0] * g[0] + metric[d][1] * g[1] + metric[d][2] * g[2]);
    for (int i = 0; i < t->nl; ++i)
      y[i] += t->grad[q][i][0] * r[0] + t->grad[q][i][1] * r[1] + t->grad[q][i][2] * r[2];
  }
}

/* local stiffness matrix (for the matrix-based and diagonal paths) */
static void cg_cell_matrix(const tables* t, const double jit[3][3], double det, double* K) {
  double rg[MAXQ][MAXL][3];
  for (int q = 0; q < t->nq; ++q)
    for (int i = 0; i < t->nl; ++i)
      for (int d = 0; d < 3; ++d) {
        double s = 0;
        for (int e = 0; e < 3; ++e) s += jit[d][e] * t->grad[q][i][e];
        rg[q][i][d] = s;
      }
  for (int i = 0; i < t->nl; ++i)
    for (int j = 0; j < t->nl; ++j) {
      double s = 0;
      for (int q = 0; q < t->nq; ++q)
        s += t->w[q] * fabs(det) * (rg[q][i][0] * rg[q][j][0] + rg[q][i][1] * rg[q][j][1] + rg[q][i][2] * rg[q][j][2]);
      K[i * t->nl + j] = s;
    }
}

static tables g_tables[4];
static int g_tables_ready[4];

static const tables* get_tables(int p) {
  if (p < 1 || p > 3) return NULL;
  int ok = 1;
#pragma omp critical(orc_tables)
  {
    if (!g_tables_ready[p]) {
      if (make_tables(p, &g_tables[p]) == 0) g_tables_ready[p] = 1;
      else ok = 0;
    }
  }
  return ok ? &g_tables[p] : NULL;
}

/* ------------------------------------------------------------------------ */
/* CG operator                                                               */

static void cg_apply_cell(const tables* t, const orc_mesh* m, int64_t c, const int64_t* cell_dofs,
                          const uint8_t* con, const double* src, double* dst) {
  double jit[3][3], u[MAXL], y[MAXL];
  const double det = geometry(m, c, jit);
  const int64_t* dofs = cell_dofs + c * t->nl;
  for (int i = 0; i < t->nl; ++i) u[i] = con[dofs[i]] ? 0.0 : src[dofs[i]];
  cg_cell(t, jit, det, u, y);
  for (int i = 0; i < t->nl; ++i)
    if (!con[dofs[i]]) dst[dofs[i]] += y[i];
}

int orc_cg_vmult(const orc_mesh* m, int p, const int64_t* cell_dofs, const uint8_t* con,
                 int64_t n_dofs, const double* src, double* dst) {
  const tables* t = get_tables(p);
  if (!t) return -1;
  memset(dst, 0, n_dofs * sizeof(double));
  for (int64_t c = 0; c < m->n_cells; ++c) cg_apply_cell(t, m, c, cell_dofs, con, src, dst);
  for (int64_t i = 0; i < n_dofs; ++i)
    if (con[i]) dst[i] = src[i];
  return 0;
}

int orc_cg_colors(const orc_mesh* m, int p, const int64_t* cell_dofs, int64_t n_dofs,
                  int32_t* color_ptr, int64_t* color_cells) {
  const int nl = nloc_of(p);
  uint64_t* mask = (uint64_t*)calloc((size_t)n_dofs * 4, sizeof(uint64_t));
  int32_t* col = (int32_t*)malloc((size_t)m->n_cells * sizeof(int32_t));
  if (!mask || !col) return -1;
  int ncol = 0;
  for (int64_t c = 0; c < m->n_cells; ++c) {
    uint64_t u[4] = {0, 0, 0, 0};
    for (int i = 0; i < nl; ++i)
      for (int k = 0; k < 4; ++k) u[k] |= mask[4 * cell_dofs[c * nl + i] + k];
    int cc = -1;
    for (int k = 0; k < 4 && cc < 0; ++k)
      if (~u[k]) cc = 64 * k + __builtin_ctzll(~u[k]);
    if (cc < 0) {
      free(mask);
      free(col);
      return -1;
    }
    col[c] = cc;
    if (cc + 1 > ncol) ncol = cc + 1;
    for (int i = 0; i < nl; ++i) mask[4 * cell_dofs[c * nl + i] + (cc >> 6)] |= 1ull << (cc & 63);
  }
  for (int k = 0; k <= ncol; ++k) color_ptr[k] = 0;
  for (int64_t c = 0; c < m->n_cells; ++c) color_ptr[col[c] + 1]++;
  for (int k = 0; k < ncol; ++k) color_ptr[k + 1] += color_ptr[k];
  {
    int32_t pos[257];
    for (int k = 0; k < ncol; ++k) pos[k] = color_ptr[k];
    for (int64_t c = 0; c < m->n_cells; ++c) color_cells[pos[col[c]]++] = c;
  }
  free(mask);
  free(col);
  return ncol;
}

int orc_cg_vmult_omp(const orc_mesh* m, int p, const int64_t* cell_dofs, const uint8_t* con,
                     int64_t n_dofs, const int32_t* color_ptr, const int64_t* color_cells,
                     int n_colors, const double* src, double* dst) {
  const tables* t = get_tables(p);
  if (!t) return -1;
#pragma omp parallel
  {
#pragma omp for schedule(static)
    for (int64_t i = 0; i < n_dofs; ++i) dst[i] = 0.0;
    for (int k = 0; k < n_colors; ++k) {
#pragma omp for schedule(static)
      for (int64_t j = color_ptr[k]; j < color_ptr[k + 1]; ++j)
        cg_apply_cell(t, m, color_cells[j], cell_dofs, con, src, dst);
    }
#pragma omp for schedule(static)
    for (int64_t i = 0; i < n_dofs; ++i)
      if (con[i]) dst[i] = src[i];
  }
  return 0;
}

int orc_cg_diagonal(const orc_mesh* m, int p, const int64_t* cell_dofs, const uint8_t* con,
                    int64_t n_dofs, double* diag) {
  const tables* t = get_tables(p);
  if (!t) return -1;
  double K[MAXL * MAXL], jit[3][3];
  memset(diag, 0, n_dofs * sizeof(double));
  for (int64_t c = 0; c < m->n_cells; ++c) {
    const double det = geometry(m, c, jit);
    cg_cell_matrix(t, jit, det, K);
    for (int i = 0; i < t->nl; ++i) {
      const int64_t d = cell_dofs[c * t->nl + i];
      if (!con[d]) diag[d] += K[i * t->nl + i];
    }
  }
  for (int64_t i = 0; i < n_dofs; ++i)
    if (con[i]) diag[i] = 1.0;
  return 0;
}

typedef struct trip {
  int64_t r, c;
  double v;
} trip;

static int trip_cmp(const void* a, const void* b) {
  const trip* x = (const trip*)a;
  const trip* y = (const trip*)b;
  if (x->r != y->r) return x->r < y->r ? -1 : 1;
  if (x->c != y->c) return x->c < y->c ? -1 : 1;
  return 0;
}

int orc_cg_matrix_vmult(const orc_mesh* m, int p, const int64_t* cell_dofs, const uint8_t* con,
                        int64_t n_dofs, const double* src, double* dst) {
  const tables* t = get_tables(p);
  if (!t) return -1;
  const int nl = t->nl;
  const int64_t cap = m->n_cells * nl * nl + n_dofs;
  trip* tr = (trip*)malloc(cap * sizeof(trip));
  if (!tr) return -1;
  int64_t nt = 0;
  double K[MAXL * MAXL], jit[3][3];
  for (int64_t c = 0; c < m->n_cells; ++c) {
    const double det = geometry(m, c, jit);
    cg_cell_matrix(t, jit, det, K);
    const int64_t* dofs = cell_dofs + c * nl;
    for (int i = 0; i < nl; ++i)
      for (int j = 0; j < nl; ++j)
        if (!con[dofs[i]] && !con[dofs[j]]) tr[nt++] = (trip){dofs[i], dofs[j], K[i * nl + j]};
  }
  for (int64_t i = 0; i < n_dofs; ++i)
    if (con[i]) tr[nt++] = (trip){i, i, 1.0};
  qsort(tr, nt, sizeof(trip), trip_cmp);
  /* compress into CSR and multiply */
  int64_t* rowptr = (int64_t*)calloc(n_dofs + 1, sizeof(int64_t));
  int64_t* colidx = (int64_t*)malloc((nt ? nt : 1) * sizeof(int64_t));
  double* val = (double*)malloc((nt ? nt : 1) * sizeof(double));
  int64_t nnz = 0;
  for (int64_t k = 0; k < nt; ++k) {
    if (nnz > 0 && colidx[nnz - 1] == tr[k].c && k > 0 && tr[k - 1].r == tr[k].r) {
      val[nnz - 1] += tr[k].v;
    } else {
      colidx[nnz] = tr[k].c;
      val[nnz] = tr[k].v;
      rowptr[tr[k].r + 1]++;
      nnz++;
    }
  }
  for (int64_t i = 0; i < n_dofs; ++i) rowptr[i + 1] += rowptr[i];
  for (int64_t i = 0; i < n_dofs; ++i) {
    double s = 0;
    for (int64_t k = rowptr[i]; k < rowptr[i + 1]; ++k) s += val[k] * src[colidx[k]];
    dst[i] = s;
  }
  free(tr);
  free(rowptr);
  free(colidx);
  free(val);
  return 0;
}

int orc_cg_dense(const orc_mesh* m, int p, const int64_t* cell_dofs, const uint8_t* con,
                 int64_t n_dofs, double* A) {
  const tables* t = get_tables(p);
  if (!t) return -1;
  const int nl = t->nl;
  double K[MAXL * MAXL], jit[3][3];
  memset(A, 0, (size_t)n_dofs * n_dofs * sizeof(double));
  for (int64_t c = 0; c < m->n_cells; ++c) {
    const double det = geometry(m, c, jit);
    cg_cell_matrix(t, jit, det, K);
    const int64_t* dofs = cell_dofs + c * nl;
    for (int i = 0; i < nl; ++i)
      for (int j = 0; j < nl; ++j)
        if (!con[dofs[i]] && !con[dofs[j]]) A[dofs[i] * n_dofs + dofs[j]] += K[i * nl + j];
  }
  for (int64_t i = 0; i < n_dofs; ++i)
    if (con[i]) A[i * n_dofs + i] = 1.0;
  return 0;
}
