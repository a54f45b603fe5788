/* CPU oracle, DG SIPG part -- TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Restates the reference's deal.II path of the SIPG operator
 * (DG/cells_batched/application.cc:1348-1490): cell integrals
 * int grad u . grad v; per interior face, with n the outward normal of the
 * first cell and both traces taken at the same physical quadrature points,
 *   jump = u- - u+, avg = (grad u- + grad u+).n / 2,
 *   v-: (sigma jump - avg),  v+: -(sigma jump - avg),
 *   grad v-.n and grad v+.n: -jump/2     (1396-1423);
 * boundary faces with the mirror u+ = -u-, grad u+.n = grad u-.n (1462-1484);
 * penalty per cell from face areas and volume (1193-1245).
 * Traces are evaluated by mapping the physical face point back to each
 * cell's reference coordinates, so no face orientation tables are involved.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define MAXL 20
#define MAXQF 64

static int nloc_of(int p) { return (p + 1) * (p + 2) * (p + 3) / 6; }

/* penalty variant: 0 = (p+1)(p+3)/3 max(tau-, tau+) (reference,
 * DG/cells_batched/application.cc:1193-1245); 1 = (p+1)(p+2)/h_F with
 * h_F = |K|/|F| on each side, the larger value over the two sides */
static int g_penalty_kind = 0;
int orc_dg_set_penalty_kind(int kind) {
  if (kind != 0 && kind != 1) return -1;
  g_penalty_kind = kind;
  return 0;
}

/* planted mistakes, for the tests that show a check catches them:
 * 1 min(tau-, tau+) instead of max; 2 average of the normal derivatives
 * without the factor 1/2; 3 lumped (row-sum) mass matrix */
static int g_plant = 0;
int orc_dg_set_plant(int plant) {
  if (plant < 0 || plant > 3) return -1;
  g_plant = plant;
  return 0;
}

typedef struct geo {
  double x0[3], J[3][3], Jinv[3][3], det;
} geo;

static void cell_geo(const orc_mesh* m, int64_t c, geo* g) {
  const int32_t* v = m->cells + 4 * c;
  for (int d = 0; d < 3; ++d) {
    g->x0[d] = m->vertices[3 * v[0] + d];
    for (int k = 0; k < 3; ++k) g->J[d][k] = m->vertices[3 * v[k + 1] + d] - g->x0[d];
  }
  double(*J)[3] = g->J;
  g->det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) - J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
           J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
  const double id = 1.0 / g->det;
  g->Jinv[0][0] = (J[1][1] * J[2][2] - J[1][2] * J[2][1]) * id;
  g->Jinv[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) * id;
  g->Jinv[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) * id;
  g->Jinv[1][0] = (J[1][2] * J[2][0] - J[1][0] * J[2][2]) * id;
  g->Jinv[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * id;
  g->Jinv[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * id;
  g->Jinv[2][0] = (J[1][0] * J[2][1] - J[1][1] * J[2][0]) * id;
  g->Jinv[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * id;
  g->Jinv[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * id;
}

/* physical values and gradients of the cell basis at physical point x */
static void eval_phys(int p, const geo* g, const double* x, double* val, double (*grad)[3]) {
  double xr[3], gr[MAXL * 3];
  for (int a = 0; a < 3; ++a) {
    xr[a] = 0;
    for (int b = 0; b < 3; ++b) xr[a] += g->Jinv[a][b] * (x[b] - g->x0[b]);
  }
  orc_basis(p, xr, val, gr);
  const int n = nloc_of(p);
  for (int i = 0; i < n; ++i)
    for (int a = 0; a < 3; ++a) {
      double s = 0;
      for (int b = 0; b < 3; ++b) s += g->Jinv[b][a] * gr[3 * i + b]; /* J^-T grad_ref */
      grad[i][a] = s;
    }
}

/* faces: key = sorted vertex triple; entries (cell, local face) */
typedef struct fent {
  int32_t v[3];
  int64_t cell;
  int lf;
} fent;

static int fent_cmp(const void* a, const void* b) {
  const fent* x = (const fent*)a;
  const fent* y = (const fent*)b;
  for (int k = 0; k < 3; ++k)
    if (x->v[k] != y->v[k]) return x->v[k] < y->v[k] ? -1 : 1;
  if (x->cell != y->cell) return x->cell < y->cell ? -1 : 1;
  return 0;
}

static void sort3(int32_t* v) {
  int32_t t;
  if (v[0] > v[1]) { t = v[0]; v[0] = v[1]; v[1] = t; }
  if (v[1] > v[2]) { t = v[1]; v[1] = v[2]; v[2] = t; }
  if (v[0] > v[1]) { t = v[0]; v[0] = v[1]; v[1] = t; }
}

static fent* build_faces(const orc_mesh* m) {
  fent* f = (fent*)malloc(4 * m->n_cells * sizeof(fent));
  for (int64_t c = 0; c < m->n_cells; ++c)
    for (int lf = 0; lf < 4; ++lf) {
      fent* e = &f[4 * c + lf];
      int k = 0;
      for (int t = 0; t < 4; ++t)
        if (t != lf) e->v[k++] = m->cells[4 * c + t];
      sort3(e->v);
      e->cell = c;
      e->lf = lf;
    }
  qsort(f, 4 * m->n_cells, sizeof(fent), fent_cmp);
  return f;
}

static double tri_area(const orc_mesh* m, const int32_t* v, double* n) {
  double a[3], b[3];
  for (int d = 0; d < 3; ++d) {
    a[d] = m->vertices[3 * v[1] + d] - m->vertices[3 * v[0] + d];
    b[d] = m->vertices[3 * v[2] + d] - m->vertices[3 * v[0] + d];
  }
  n[0] = a[1] * b[2] - a[2] * b[1];
  n[1] = a[2] * b[0] - a[0] * b[2];
  n[2] = a[0] * b[1] - a[1] * b[0];
  const double l = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
  for (int d = 0; d < 3; ++d) n[d] /= l;
  return 0.5 * l;
}

int orc_dg_tau(const orc_mesh* m, double* tau) {
  fent* f = build_faces(m);
  const int64_t nf = 4 * m->n_cells;
  double* surf = (double*)calloc(m->n_cells, sizeof(double));
  for (int64_t k = 0; k < nf;) {
    int64_t j = k + 1;
    while (j < nf && f[j].v[0] == f[k].v[0] && f[j].v[1] == f[k].v[1] && f[j].v[2] == f[k].v[2]) ++j;
    double n[3];
    const double area = tri_area(m, f[k].v, n);
    const double factor = (j - k == 1) ? 1.0 : 0.5;
    for (int64_t t = k; t < j; ++t) surf[f[t].cell] += factor * area;
    k = j;
  }
  for (int64_t c = 0; c < m->n_cells; ++c) {
    geo g;
    cell_geo(m, c, &g);
    tau[c] = surf[c] / (fabs(g.det) / 6.0);
  }
  free(surf);
  free(f);
  return 0;
}

/* triangle rule (collapsed Gauss, n^2 points), barycentric weights summing to 1 */
static int tri_rule(int n, double (*bc)[3], double* w) {
  double x[32], wx[32];
  if (orc_gauss01(n, x, wx)) return -1;
  int k = 0;
  for (int a = 0; a < n; ++a)
    for (int b = 0; b < n; ++b, ++k) {
      const double u = x[a], v = x[b] * (1.0 - x[a]);
      bc[k][0] = 1.0 - u - v;
      bc[k][1] = u;
      bc[k][2] = v;
      w[k] = 2.0 * wx[a] * wx[b] * (1.0 - x[a]); /* sums to 1 */
    }
  return k;
}

typedef struct dgctx {
  int p, nl, ncomp, parallel;
  double mass, laplace, factor;
  const orc_mesh* m;
  const int64_t* cd;
  int64_t n_dofs;
  double* tau;
} dgctx;

/* apply (or take the diagonal of, when diag != NULL) the operator */
static int dg_apply(const dgctx* c, const double* src, double* dst, double* diag) {
  const orc_mesh* m = c->m;
  const int p = c->p, nl = c->nl, nc = c->ncomp;
  const int64_t N = c->n_dofs;
  const int64_t ntot = N * nc;
  if (dst) memset(dst, 0, ntot * sizeof(double));
  if (diag) memset(diag, 0, ntot * sizeof(double));
  double tpts[3 * 216], tw[216];
  const int nq1 = p + 2;
  orc_tet_rule(nq1, tpts, tw);
  const int nq = nq1 * nq1 * nq1;
  /* cells (each writes only its own rows) */
#pragma omp parallel for schedule(dynamic, 64) if (c->parallel)
  for (int64_t k = 0; k < m->n_cells; ++k) {
    geo g;
    cell_geo(m, k, &g);
    const int64_t* dofs = c->cd + k * nl;
    for (int q = 0; q < nq; ++q) {
      double x[3], val[MAXL], grad[MAXL][3];
      for (int a = 0; a < 3; ++a) {
        x[a] = g.x0[a];
        for (int b = 0; b < 3; ++b) x[a] += g.J[a][b] * tpts[3 * q + b];
      }
      eval_phys(p, &g, x, val, grad);
      const double jxw = tw[q] * fabs(g.det);
      for (int comp = 0; comp < nc; ++comp) {
        const int64_t off = (int64_t)comp * N;
        if (dst) {
          double u = 0, gu[3] = {0, 0, 0};
          for (int i = 0; i < nl; ++i) {
            const double s = src[off + dofs[i]];
            u += s * val[i];
            for (int a = 0; a < 3; ++a) gu[a] += s * grad[i][a];
          }
          for (int i = 0; i < nl; ++i)
            dst[off + dofs[i]] += jxw * (c->laplace * (gu[0] * grad[i][0] + gu[1] * grad[i][1] + gu[2] * grad[i][2]) +
                                         c->mass * (g_plant == 3 ? src[off + dofs[i]] : u) * val[i]);
        } else {
          for (int i = 0; i < nl; ++i)
            diag[off + dofs[i]] += jxw * (c->laplace * (grad[i][0] * grad[i][0] + grad[i][1] * grad[i][1] + grad[i][2] * grad[i][2]) +
                                          c->mass * val[i] * val[i]);
        }
      }
    }
  }
  /* faces: group the sorted entries, then evaluate every face (atomic adds
   * when run in parallel) */
  fent* f = build_faces(m);
  const int64_t nf = 4 * m->n_cells;
  double bc[MAXQF][3], fw[MAXQF];
  const int nqf = tri_rule(p + 2, bc, fw);
  int64_t* starts = (int64_t*)malloc((nf + 1) * sizeof(int64_t));
  int64_t ngroups = 0;
  for (int64_t k = 0; k < nf;) {
    int64_t j = k + 1;
    while (j < nf && f[j].v[0] == f[k].v[0] && f[j].v[1] == f[k].v[1] && f[j].v[2] == f[k].v[2]) ++j;
    if (j - k > 2) {
      free(f);
      free(starts);
      return -1;
    }
    starts[ngroups++] = k;
    k = j;
  }
  starts[ngroups] = nf;
#pragma omp parallel for schedule(dynamic, 64) if (c->parallel)
  for (int64_t g = 0; g < ngroups; ++g) {
    const int64_t k = starts[g], j = starts[g + 1];
    const int interior = (j - k == 2);
    const int64_t cm = f[k].cell, cp = interior ? f[k + 1].cell : -1;
    geo gm, gp;
    cell_geo(m, cm, &gm);
    if (interior) cell_geo(m, cp, &gp);
    double n[3];
    const double area = tri_area(m, f[k].v, n);
    /* orient n outward of cm: away from cm's opposite vertex */
    {
      const int32_t opp = m->cells[4 * cm + f[k].lf];
      double s = 0;
      for (int d = 0; d < 3; ++d) s += n[d] * (m->vertices[3 * f[k].v[0] + d] - m->vertices[3 * opp + d]);
      if (s < 0)
        for (int d = 0; d < 3; ++d) n[d] = -n[d];
    }
    double sigma;
    if (g_penalty_kind == 1) {
      const double hm = fabs(gm.det) / 6.0 / area, hp = interior ? fabs(gp.det) / 6.0 / area : hm;
      sigma = (c->p + 1.0) * (c->p + 2.0) * fmax(1.0 / hm, 1.0 / hp);
    } else {
      sigma = interior ? c->factor * (g_plant == 1 ? fmin(c->tau[cm], c->tau[cp]) : fmax(c->tau[cm], c->tau[cp]))
                       : c->factor * c->tau[cm];
    }
    const int64_t* dm = c->cd + cm * nl;
    const int64_t* dp = interior ? c->cd + cp * nl : NULL;
    for (int q = 0; q < nqf; ++q) {
      double x[3];
      for (int d = 0; d < 3; ++d)
        x[d] = bc[q][0] * m->vertices[3 * f[k].v[0] + d] + bc[q][1] * m->vertices[3 * f[k].v[1] + d] +
               bc[q][2] * m->vertices[3 * f[k].v[2] + d];
      const double jxw = fw[q] * area;
      double vm[MAXL], gmx[MAXL][3], vp[MAXL], gpx[MAXL][3], dnm[MAXL], dnp[MAXL];
      eval_phys(p, &gm, x, vm, gmx);
      for (int i = 0; i < nl; ++i) dnm[i] = gmx[i][0] * n[0] + gmx[i][1] * n[1] + gmx[i][2] * n[2];
      if (interior) {
        eval_phys(p, &gp, x, vp, gpx);
        for (int i = 0; i < nl; ++i) dnp[i] = gpx[i][0] * n[0] + gpx[i][1] * n[1] + gpx[i][2] * n[2];
      }
      for (int comp = 0; comp < nc; ++comp) {
        const int64_t off = (int64_t)comp * N;
        if (dst) {
          double um = 0, dum = 0, up = 0, dup = 0;
          for (int i = 0; i < nl; ++i) {
            um += src[off + dm[i]] * vm[i];
            dum += src[off + dm[i]] * dnm[i];
          }
          if (interior) {
            for (int i = 0; i < nl; ++i) {
              up += src[off + dp[i]] * vp[i];
              dup += src[off + dp[i]] * dnp[i];
            }
          } else {
            up = -um; /* mirror principle */
            dup = dum;
          }
          const double jump = um - up, avg = (g_plant == 2 ? 1.0 : 0.5) * (dum + dup);
          const double tv = (sigma * jump - avg) * jxw * c->laplace;
          const double tg = -0.5 * jump * jxw * c->laplace;
          for (int i = 0; i < nl; ++i) {
            const double v = tv * vm[i] + tg * dnm[i];
#pragma omp atomic
            dst[off + dm[i]] += v;
          }
          if (interior)
            for (int i = 0; i < nl; ++i) {
              const double v = -tv * vp[i] + tg * dnp[i];
#pragma omp atomic
              dst[off + dp[i]] += v;
            }
        } else {
          /* diagonal: u = phi_i on one side only */
          for (int i = 0; i < nl; ++i) {
            double jump = interior ? vm[i] : 2.0 * vm[i];
            double avg = interior ? 0.5 * dnm[i] : dnm[i];
            const double v = c->laplace * jxw * ((sigma * jump - avg) * vm[i] - 0.5 * jump * dnm[i]);
#pragma omp atomic
            diag[off + dm[i]] += v;
          }
          if (interior)
            for (int i = 0; i < nl; ++i) {
              const double jump = -vp[i], avg = 0.5 * dnp[i];
              const double v = c->laplace * jxw * (-(sigma * jump - avg) * vp[i] - 0.5 * jump * dnp[i]);
#pragma omp atomic
              diag[off + dp[i]] += v;
            }
        }
      }
    }
  }
  free(starts);
  free(f);
  return 0;
}

static int dg_run(const orc_mesh* m, int p, const int64_t* cd, int64_t n_dofs, int ncomp, double mass,
                  double laplace, const double* src, double* dst, double* diag, int parallel) {
  if (p < 1 || p > 3) return -1;
  dgctx c;
  c.parallel = parallel;
  c.p = p;
  c.nl = nloc_of(p);
  c.ncomp = ncomp;
  c.mass = mass;
  c.laplace = laplace;
  c.factor = (p + 1.0) * (p + 3.0) / 3.0;
  c.m = m;
  c.cd = cd;
  c.n_dofs = n_dofs;
  c.tau = (double*)malloc(m->n_cells * sizeof(double));
  orc_dg_tau(m, c.tau);
  const int rc = dg_apply(&c, src, dst, diag);
  free(c.tau);
  return rc;
}

int orc_dg_vmult(const orc_mesh* m, int p, const int64_t* cd, int64_t n_dofs, int ncomp, double mass,
                 double laplace, const double* src, double* dst) {
  return dg_run(m, p, cd, n_dofs, ncomp, mass, laplace, src, dst, NULL, 0);
}

int orc_dg_vmult_omp(const orc_mesh* m, int p, const int64_t* cd, int64_t n_dofs, int ncomp, double mass,
                     double laplace, const double* src, double* dst) {
  return dg_run(m, p, cd, n_dofs, ncomp, mass, laplace, src, dst, NULL, 1);
}

int orc_dg_diagonal(const orc_mesh* m, int p, const int64_t* cd, int64_t n_dofs, int ncomp, double mass,
                    double laplace, double* diag) {
  return dg_run(m, p, cd, n_dofs, ncomp, mass, laplace, NULL, NULL, diag, 0);
}
