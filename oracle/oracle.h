/* CPU oracle for the simplexmf operators.
 *
 * TEST INFRASTRUCTURE ONLY: tests/, __graft_entry__.smoke() and bench.py's
 * CPU-baseline leg call this library as the checker; the product never links
 * or loads it.
 *
 * It restates the reference algorithm in plain C, independently of the
 * product's tables: its own nodal basis (Vandermonde inversion over monomials
 * from the node-position convention), its own quadrature (collapsed
 * Gauss-Legendre product rules of high degree), its own geometry.
 */
#ifndef SMF_ORACLE_H
#define SMF_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_mesh {
  int64_t n_vertices;
  const double* vertices; /* 3*n_vertices */
  int64_t n_cells;
  const int32_t* cells;   /* 4*n_cells, in the order of the cell_dofs table */
} orc_mesh;

/* Nodal basis of degree p on the unit tetrahedron at x: values[n_loc], grads[3*n_loc]. */
int orc_basis(int p, const double* x, double* values, double* grads);
/* Collapsed Gauss-Legendre rule with n points per direction (n^3 points):
 * pts[3*n^3], w[n^3]. Exact for polynomials of degree <= 2n-3. */
int orc_tet_rule(int n, double* pts, double* w);
/* Gauss-Legendre on [0,1] */
int orc_gauss01(int n, double* x, double* w);

/* CG Poisson vmult (CG/cells_batched/application.cc:1236-1416 + 1173-1175):
 * dst = A src with identity rows on constrained DoFs. */
int orc_cg_vmult(const orc_mesh* m, int p, const int64_t* cell_dofs, const uint8_t* constrained,
                 int64_t n_dofs, const double* src, double* dst);
/* same, OpenMP over cell colours (CPU baseline); colors[n_cells] from orc_cg_colors */
int orc_cg_vmult_omp(const orc_mesh* m, int p, const int64_t* cell_dofs, const uint8_t* constrained,
                     int64_t n_dofs, const int32_t* color_ptr, const int64_t* color_cells,
                     int n_colors, const double* src, double* dst);
/* greedy colouring of cells sharing a DoF; returns number of colours, fills
 * color_ptr[n_colors+1] (caller gives room for 256+1) and color_cells[n_cells] */
int orc_cg_colors(const orc_mesh* m, int p, const int64_t* cell_dofs, int64_t n_dofs,
                  int32_t* color_ptr, int64_t* color_cells);
/* diagonal of A (constrained rows: 1) (application_hp.cc:1386-1409) */
int orc_cg_diagonal(const orc_mesh* m, int p, const int64_t* cell_dofs, const uint8_t* constrained,
                    int64_t n_dofs, double* diag);
/* independent matrix-based check: assemble A as CSR from element matrices and
 * multiply (the reference's own matrix comparison, application.cc:1944-1988) */
int orc_cg_matrix_vmult(const orc_mesh* m, int p, const int64_t* cell_dofs,
                        const uint8_t* constrained, int64_t n_dofs, const double* src, double* dst);
/* dense A for tiny meshes (n_dofs^2 doubles, row-major) */
int orc_cg_dense(const orc_mesh* m, int p, const int64_t* cell_dofs, const uint8_t* constrained,
                 int64_t n_dofs, double* A);

/* ---- DG SIPG (DG/cells_batched/application.cc:1170-1490) ----------------
 * DoFs: cell_dofs[c*n_loc + i] (any numbering).  Interior faces and boundary
 * faces are found from the mesh; penalty
 *   tau_K = (sum_{F bdry} |F| + 1/2 sum_{F int} |F|) / |K|,
 *   sigma_F = (p+1)(p+3)/3 * max(tau-, tau+),  sigma_B = (p+1)(p+3)/3 * tau_K.
 * mass != 0 adds mass * int u v (Helmholtz-like operator); ncomp components
 * share geometry, component c of DoF d is row c*n_dofs + d.                */
int orc_dg_vmult(const orc_mesh* m, int p, const int64_t* cell_dofs, int64_t n_dofs, int ncomp,
                 double mass, double laplace, const double* src, double* dst);
/* same, OpenMP over cells and faces (face contributions by atomic adds; CPU baseline) */
int orc_dg_vmult_omp(const orc_mesh* m, int p, const int64_t* cell_dofs, int64_t n_dofs, int ncomp,
                     double mass, double laplace, const double* src, double* dst);
int orc_dg_diagonal(const orc_mesh* m, int p, const int64_t* cell_dofs, int64_t n_dofs, int ncomp,
                    double mass, double laplace, double* diag);
/* penalty variant for the DG functions: 0 reference (default), 1 (p+1)(p+2)/h_F, h_F = |K|/|F| */
int orc_dg_set_penalty_kind(int kind);
/* planted mistakes for tests (0 none, 1 min for max penalty, 2 average
 * without 1/2, 3 lumped mass); vmult only */
int orc_dg_set_plant(int plant);
/* per-cell penalty tau_K (n_cells) */
int orc_dg_tau(const orc_mesh* m, double* tau);

/* ---- CPU baseline (cpu_baseline.cpp): the reference's optimised DG SIPG
 * vmult restated for CPUs -- 8-lane vectors of cells or faces, batches of 4
 * vectors per shape matrix, OpenMP over cell batches and coloured face
 * batches.  Cells must list their vertices in ascending order; DoF (k, i)
 * is row k*n_loc + i in orc_basis node order.  nthreads <= 0: all threads. */
void* orc_dgb_create(const orc_mesh* m, int p, int nthreads);
int orc_dgb_vmult(void* plan, const double* src, double* dst);
/* reps repetitions of n back-to-back products, seconds per repetition */
int orc_dgb_time(void* plan, const double* src, double* dst, int reps, int n, double* seconds);
/* out[6]: cell points, face points, face groups, colours, threads, vector bits */
int orc_dgb_info(void* plan, int64_t* out);
/* the rules used: kind 0 tetrahedron (4 barycentric coordinates per point),
 * 1 triangle (3 per point); returns the number of points */
int orc_dgb_rule(int kind, int p, double* bc, double* w);
void orc_dgb_destroy(void* plan);

/* ---- learned-smoother network (oracle_mlp.c) ---------------------------- */
int orc_mlp_n_params(void);
uint16_t orc_bf16(float f);
/* y[n][10] of x[n][20]; bf16 != 0 rounds weights and layer inputs to bfloat16 */
int orc_mlp_forward(const double* params, const double* x, int64_t n, int bf16, double* y);
int orc_mlp_loss(const double* params, const double* x, const double* t, int64_t n, double* loss);
int orc_mlp_loss_grad(const double* params, const double* x, const double* t, int64_t n, double* grad,
                      double* loss);
int orc_adam(double* p, const double* g, double* m, double* v, int64_t n, double lr, double b1, double b2,
             double eps, int64_t step);

#ifdef __cplusplus
}
#endif
#endif
