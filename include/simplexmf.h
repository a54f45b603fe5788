/* simplexmf — matrix-free tetrahedral finite-element operators on MI355X.
 *
 * C ABI.  Every entry point returns 0 on success and a negative code on
 * failure; mf_last_error() then returns a thread-local message.  Vectors are
 * device pointers (FP64 or FP32 as chosen at creation), streams are
 * hipStream_t values passed as void* (NULL = default stream).  Calls are
 * stream-ordered; only the create/setup calls synchronise.  dst and src must
 * not alias.
 *
 * Reference interface each entry replaces (paths relative to the
 * simplex-benchmarks repository):
 *   Operator::reinit                CG/cells_batched/application.cc:1056-1128
 *                                   DG/cells_batched/application.cc:1170-1277
 *   Operator::vmult_masked_gather   CG/cells_batched/application.cc:1165-1176
 *   Operator::vmult_manual          DG/cells_batched/application.cc:1318-1332
 *   Operator::Tvmult                CG/cells_batched/application.cc:1192-1196
 *   Operator::m                     CG/cells_batched/application.cc:1130-1138
 *   Operator::compute_inverse_diagonal CG/cells_batched/application_hp.cc:1386-1409
 */
#ifndef SIMPLEXMF_H
#define SIMPLEXMF_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mf_operator* mf_op_t;

enum mf_kind {
  MF_CG_POISSON = 0,   /* scalar Poisson, continuous P_k, homogeneous Dirichlet rows = identity */
  MF_DG_SIPG = 1,      /* scalar Poisson, discontinuous P_k, symmetric interior penalty */
  MF_CG_VECTOR = 2,    /* 3-component Laplacian, continuous P_k (components batched) */
  MF_DG_VECTOR = 3,    /* 3-component SIPG Laplacian (components batched) */
  MF_DG_HELMHOLTZ = 4  /* 3-component mass*u + laplace*SIPG (components batched) */
};

enum mf_precision { MF_F64 = 0, MF_F32 = 1 };

typedef struct mf_mesh_desc {
  int64_t n_vertices;
  const double* vertices; /* 3*n_vertices, x y z per vertex */
  int64_t n_cells;
  const int32_t* cells;   /* 4*n_cells vertex ids per tetrahedron */
  int32_t kuhn_n;         /* >0: mesh came from mf_mesh_kuhn_cube(kuhn_n, ...) (enables h-transfers) */
} mf_mesh_desc;

typedef struct mf_options {
  int32_t reorder;          /* 1: hierarchical cell reordering (default); 0: keep input order */
  int32_t patch_cells;      /* cells per patch (0: default for the degree) */
  double mass_coefficient;  /* Helmholtz mass coefficient (default 1) */
  double laplace_coefficient; /* Helmholtz Laplace coefficient (default 1) */
  int32_t penalty_kind;     /* 0: (p+1)(p+3)/3 max(tau-,tau+) (reference); 1: (p+1)(p+2)/h_face with
                               h_face = |K|/|F| per side, the larger penalty of the two sides */
  int32_t host_only;        /* 1: build the host plan only (no device memory); device entry points fail */
  int32_t accumulate;       /* CG LDS accumulation: 0 auto, 1 LDS atomics per colour, 2 read-add-write per
                               colour, 3 LDS atomics per wave (wave order fixed per round) */
  int32_t dg_kernel;        /* DG kernel: 0 auto, 1 LDS tables, 2 compile-time tables + neighbour in face-3 frame,
                               3 four lanes per cell (one face each), 4 the same with the workgroup's cell
                               values staged in LDS. Auto: four lanes per cell for the product at every
                               degree, staged for P2 and for P1 FP64 and P3 FP32 (measured faster there);
                               the one-lane form for the P1 diagonal */
  int32_t boundary;         /* CG Poisson: 0 Dirichlet rows = identity (default); 1 Nitsche: no constrained
                               DoFs, boundary faces add int(2 sigma u v - du/dn v - u dv/dn) with
                               sigma = (p+1)(p+3)/3 tau_K (the CG level under a DG operator in the hybrid
                               multigrid, DG/CG c-transfer, application_phc.cc:1424-1456, 1798-1832) */
  int32_t cells_per_wave;   /* CG: cells per wavefront of a patch workgroup (0: 1024 P1, 512 P2, 128 P3);
                               waves per patch = patch_cells / cells_per_wave, at most 4 */
  int32_t manifold;         /* CG (Dirichlet boundary) and DG: 0 affine cells (default); 1 curved cells: a P2
                               geometry map whose edge nodes lie on the spherical manifold around
                               manifold_center (MappingFE(FE_SimplexP(2)) + SphericalManifold,
                               CG/cells_batched/application.cc:1613, 1668-1677, 1763-1767), integrated
                               per quadrature point (the general D_e branch, :1361-1376) with a rule of
                               degree 2p; 2 the same P2 map with straight-edge midpoints (affine cells
                               through the curved-cell path).  DG: curved faces with per-point normals
                               and J^-1 n on both sides (DG/cells_batched/application.cc:2191-2238), tau
                               from the curved volumes and face areas; no multigrid below it */
  double manifold_center[3];  /* centre of the spherical manifold (default (5, 0, 0), as the reference) */
} mf_options;

typedef struct mf_info {
  int64_t n_dofs;          /* == mf_m() */
  int64_t n_cells;
  int64_t n_constrained;   /* Dirichlet rows (CG) */
  int32_t degree, dofs_per_cell, n_q_points, n_components;
  int64_t n_patches;
  int32_t n_patch_colors, max_cell_colors, max_local_dofs, precision;
  int64_t n_interior_faces, n_boundary_faces;
  double algorithmic_bytes; /* per vmult, reference data model */
  double flops;             /* per vmult */
} mf_info;

/* ---- meshes (host) ------------------------------------------------------ */

/* Unit cube split into n^3 hexes x 6 Kuhn tetrahedra; interior vertices
 * displaced per coordinate by perturb*h*U[-1,1) from splitmix64(seed).
 * vertices: 3*(n+1)^3 doubles, cells: 4*6*n^3 int32 (caller-allocated). */
int mf_mesh_kuhn_cube(int32_t n, double perturb, uint64_t seed, double* vertices,
                      int32_t* cells);
/* External mesh: a legacy ASCII VTK UNSTRUCTURED_GRID of tetrahedra (cell
 * type 10; classic "k i0 .. ik-1" CELLS or the VTK 5 OFFSETS/CONNECTIVITY
 * layout).  Points, lines and triangles (types 1, 3, 5) are skipped, other
 * cell types rejected.  Tetrahedra of negative volume get local vertices 1
 * and 2 swapped (counted in n_flipped).  Call with vertices == cells == NULL
 * for the sizes, then with 3*n_vertices doubles and 4*n_cells int32.  The
 * cells keep the file's order; mf_create's hierarchical reordering
 * (mf_options.reorder) plays the role of the reference's reorder_grid.
 * Replaces GridIn::read_vtk + SparsityTools::reorder_hierarchical,
 * CG/cells_batched/application.cc:1699-1756. */
int mf_mesh_read_vtk(const char* path, int64_t* n_vertices, int64_t* n_cells, int64_t* n_flipped,
                     double* vertices, int32_t* cells);

/* ---- operator lifetime -------------------------------------------------- */

void mf_default_options(mf_options* opt);

/* Replaces Operator::reinit: builds DoF numbering, hierarchical patches,
 * colourings, geometry and device tables.  Synchronises the device. */
int mf_create(const mf_mesh_desc* mesh, int32_t degree, int32_t kind, int32_t precision,
              const mf_options* opt, mf_op_t* out);
int mf_destroy(mf_op_t op);

/* ---- the operator (device) ---------------------------------------------- */

/* dst = A src (overwrites dst).  CG: Dirichlet rows are identity rows,
 * dst_c = src_c (CG/cells_batched/application.cc:1173-1175). */
int mf_vmult(mf_op_t op, void* dst, const void* src, void* stream);
/* A is symmetric: Tvmult == vmult (CG/cells_batched/application.cc:1192-1196). */
int mf_tvmult(mf_op_t op, void* dst, const void* src, void* stream);
/* diag_i = e_i^T A e_i (Dirichlet rows: 1). */
int mf_compute_diagonal(mf_op_t op, void* diag, void* stream);
/* inv_i = 1/diag_i if |diag_i| > 1e-10 else 1 (application_hp.cc:1404-1407). */
int mf_compute_inverse_diagonal(mf_op_t op, void* inv, void* stream);

/* ---- queries (host) ----------------------------------------------------- */

int64_t mf_m(mf_op_t op);   /* number of rows (all components) */
int mf_get_info(mf_op_t op, mf_info* info);
/* internal cell i is input cell perm[i] (n_cells int64) */
int mf_get_cell_order(mf_op_t op, int64_t* perm);
/* per internal cell, the row of each local node: CG n_cells*dofs_per_cell
 * scalar DoFs; DG n_cells*n_components*dofs_per_cell rows, component-major
 * within the cell (row (k*n_components + c)*dofs_per_cell + i). */
int mf_get_cell_dofs(mf_op_t op, int64_t* dofs);
/* vertex ids of every internal cell in the operator's local vertex order
 * (4*n_cells int32; DG orders them ascending) */
int mf_get_cell_vertices(mf_op_t op, int32_t* verts);
/* support point of every scalar DoF (3*n_scalar_dofs doubles; DG: node
 * (k, i) at index k*dofs_per_cell + i) */
int mf_get_dof_points(mf_op_t op, double* xyz);
/* 1 for Dirichlet-constrained scalar DoFs (n_scalar_dofs bytes) */
int mf_get_constrained(mf_op_t op, uint8_t* mask);
/* the cell quadrature rule of an operator (deal.II's QGaussSimplex(p+1),
 * CG/cells_batched/application.cc:1614, is this library's own symmetric rule
 * of degree 2p-2, or 2p for curved cells and the Helmholtz mass term):
 * n_q_points points on the unit tetrahedron (3 doubles each) and weights
 * summing to 1/6 */
int mf_get_quadrature(mf_op_t op, double* points, double* weights);
/* the face rule of a DG operator (degree 2p): *n_points points (s, t) of the
 * unit triangle (2 doubles each), weights summing to 1/2.  On the face of a
 * cell with local vertices a < b < c (ascending global ids) the point is
 * x_a + s (x_b - x_a) + t (x_c - x_a) in reference coordinates; both cells of
 * a face therefore evaluate it at the same physical point.  Call with
 * points == weights == NULL for the count. */
int mf_get_face_quadrature(mf_op_t op, int32_t* n_points, double* points, double* weights);

const char* mf_last_error(void);

/* ---- hybrid multigrid and solvers (device) ---------------------------------
 * Replaces MultigridPreconditioner + SolverCG of the reference
 * (CG/cells_batched/application_ph.cc:1754-2026, application_hp.cc:1933-1992):
 * levels P_k..P_1 on the fine mesh (p-MG, decrease by one), then P1 on Kuhn
 * meshes of half size (h-MG), Chebyshev-Jacobi smoothers, an FP64 dense
 * coarse solve (in place of the reference's Trilinos ML V-cycle) and a
 * V-cycle in the level precision (FP32 by default) inside FP64 CG. */
typedef struct mf_multigrid* mf_mg_t;

typedef struct mf_mg_options {
  int32_t smoother_degree;    /* Chebyshev degree (5) */
  double smoothing_range;     /* lambda_max / lambda_min (20) */
  int32_t eig_cg_iterations;  /* CG steps of the lambda_max estimate (10) */
  int32_t level_precision;    /* MF_F32 (default) or MF_F64 */
  int32_t max_coarse_dofs;    /* limit of the dense coarse solver (4096) */
  int32_t fuse_steps;         /* 1 (default): scalar CG levels run the Chebyshev steps without a
                                 fix-up launch per step (interface rows completed by the next step);
                                 0: one operator application + fix-up per step */
  int32_t coarse_solver;      /* 0 (default): FP64 dense inverse when the coarsest level has at most
                                 max_coarse_dofs rows, else one smoothed-aggregation AMG V-cycle on its
                                 assembled FP64 matrix (the reference's MGCoarseAMG, one ML V-cycle,
                                 application_hp.cc:1684-1731; needed on meshes without a Kuhn lattice);
                                 1: dense only; 2: AMG always */
  double amg_threshold;       /* AMG aggregation threshold theta (default 0.05): coupling (i, j) is strong
                                 when |a_ij| >= theta sqrt(|a_ii a_jj|) (ML's "aggregation: threshold") */
} mf_mg_options;

#define MF_MAX_LEVELS 16
typedef struct mf_mg_info {
  int32_t n_levels;                     /* [0] = coarsest */
  int32_t degree[MF_MAX_LEVELS];
  int32_t kuhn_n[MF_MAX_LEVELS];
  int64_t n_dofs[MF_MAX_LEVELS];
  double lambda_max[MF_MAX_LEVELS];     /* Chebyshev upper bound (1.2 x estimate) */
  int64_t vcycles;                      /* V-cycles applied so far */
  int32_t level_kind[MF_MAX_LEVELS];    /* MF_CG_POISSON or MF_DG_SIPG */
  int32_t nitsche[MF_MAX_LEVELS];       /* 1: CG level with the Nitsche boundary */
  int32_t amg_levels;                   /* AMG coarse solver: its levels (0: dense coarse solve) */
  int64_t amg_rows[MF_MAX_LEVELS];      /* rows per AMG level, [0] = the coarsest geometric level */
} mf_mg_info;

typedef struct mf_solve_stats {
  int32_t iterations;
  int32_t converged;
  double initial_residual, final_residual, rhs_norm;
} mf_solve_stats;

void mf_mg_default_options(mf_mg_options* opt);
/* Builds the hierarchy below a scalar CG Poisson operator (P_k..P_1, then h)
 * or a DG SIPG operator (DG P_k, c-level CG P_k with the Nitsche boundary,
 * then p and h on CG Nitsche levels, application_phc.cc:3416-3552); h-levels
 * need kuhn_n > 0. */
int mf_mg_create(mf_op_t fine, const mf_mg_options* opt, mf_mg_t* out);
/* dst = V-cycle(src) in the fine operator's precision (PreconditionMG::vmult) */
int mf_mg_vmult(mf_mg_t mg, void* dst, const void* src, void* stream);
int mf_mg_get_info(mf_mg_t mg, mf_mg_info* info);
int mf_mg_destroy(mf_mg_t mg);
/* (F)CG on an FP64 operator from x = 0 until ||r|| <= rel_tol ||b||; mg may be
 * NULL; flexible != 0 uses Notay's flexible CG.  Synchronises the stream. */
int mf_solve_cg(mf_op_t A, mf_mg_t mg, void* x, const void* b, double rel_tol, int32_t max_iter,
                int32_t flexible, mf_solve_stats* stats, void* stream);
/* mf_solve_cg that allocates no device memory.  No counterpart in the
 * reference: for callers that cannot allocate (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  The solver's work
 * vectors, reduction partials and Krylov scalars live in a static device
 * array of the library (32 MiB per GPU, held until the process ends);
 * operators with more than 1048576 rows are refused.  All calls on one GPU
 * share that array: calls that overlap on one GPU race.  Results equal
 * mf_solve_cg's bitwise (same kernels, same order). */
int mf_solve_cg_noalloc(mf_op_t A, mf_mg_t mg, void* x, const void* b, double rel_tol, int32_t max_iter,
                        int32_t flexible, mf_solve_stats* stats, void* stream);
/* Host right-hand side b_i = int f phi_i (Dirichlet rows 0) of a CG operator:
 * kind 0 f = 3 (3 pi)^2 prod sin(3 pi x_d); kind 1 seeded Gaussian random
 * field (random Fourier features, n_features, correlation length). */
int mf_assemble_rhs(mf_op_t op, int32_t kind, uint64_t seed, int32_t n_features, double corr_length,
                    double* b);
/* Transfer between level l-1 and level l (MGTwoLevelTransfer::prolongate_and_add /
 * restrict_and_add, application_ph.cc:1902-1907): direction 0 dst(level l) = P
 * src(level l-1), 1 dst(level l-1) = R src(level l); level-precision vectors. */
int mf_mg_transfer(mf_mg_t mg, int32_t level, int32_t direction, void* dst, const void* src, void* stream);
/* support points (3*n doubles) and Dirichlet mask (n bytes) of level l's DoFs */
int mf_mg_level_dofs(mf_mg_t mg, int32_t level, double* xyz, uint8_t* constrained);
/* interpolant of prod sin(3 pi x_d) (kind-0 exact solution) */
int mf_exact_solution(mf_op_t op, double* u);

/* ---- learned per-cell smoother + flexible CG ------------------------------
 * No counterpart in the reference: a new post-smoothing step on top of its
 * MultigridPreconditioner + SolverCG (application_ph.cc:1754-2026,
 * application_hp.cc:1950-1960).  Network 20 -> 64 -> 64 -> 10 (GELU), input
 * = a cell's 10 scaled P2 residuals + h J^-T (9) + |det J|/h^3, output = the
 * cell's correction in units of |r_K| * mean(D^-1 over K); FP32 master
 * weights, BF16 MFMA inference with FP32 accumulation, FP32 training. */
typedef struct mf_learned_smoother* mf_ls_t;
#define MF_LS_N_INPUTS 20
#define MF_LS_N_OUTPUTS 10
#define MF_LS_N_PARAMS 6154 /* W1[64][20] b1[64] W2[64][64] b2[64] W3[10][64] b3[10] */

typedef struct mf_train_options {
  int32_t n_rhs;           /* random-field right-hand sides (seeds seed .. seed+n_rhs-1) */
  int32_t pairs_per_rhs;   /* preconditioned CG iterations used per RHS */
  int32_t steps_per_pair;  /* Adam steps per (residual, error) pair */
  double lr;
  int32_t n_features;
  double corr_length;
  uint64_t seed;
  int32_t loss;            /* 0: energy of the remaining error |e' - w c|_A^2/|e'|_A^2 at the
                              optimal w (the gain inside FCG); 1: mean squared e'_K/(s_K kappa_K) */
  int32_t trajectory;      /* iterates the pairs come from: 0 plain multigrid PCG, 1 flexible CG
                              with the current learned step (errors still against the plain
                              FP64 multigrid solution) */
} mf_train_options;

typedef struct mf_train_stats {
  int32_t pairs, steps;
  double first_loss;  /* loss of the first pair before its first step (energy loss: the
                         fraction of the error's A-energy left after the learned step) */
  double last_loss;   /* loss of the last pair before its first step */
  double zero_loss;   /* loss of the zero predictor on the first pair of the last RHS */
} mf_train_stats;

typedef struct mf_fcg_options {
  double stall_ratio;          /* fall back after two iterations with |r_k| > stall_ratio |r_k-1| */
  int32_t force_fallback_at;   /* iteration at which to force the fallback, -1: never */
  int32_t use_learned;         /* 0: plain multigrid FCG (same code path) */
} mf_fcg_options;

typedef struct mf_fcg_stats {
  mf_solve_stats base;
  int32_t fallback;            /* 1 when the learned step was switched off */
  int32_t fallback_iteration;  /* iteration of the switch, -1 if none */
  int32_t learned_steps;       /* preconditioner applications with the learned step */
  double mean_omega;           /* mean energy-optimal step length of the learned correction */
} mf_fcg_stats;

void mf_ls_default_train_options(mf_train_options* opt);
void mf_ls_default_fcg_options(mf_fcg_options* opt);
/* smoother for a CG P2 FP64 Poisson operator; weights initialised from seed */
int mf_ls_create(mf_op_t fine, uint64_t seed, mf_ls_t* out);
int mf_ls_destroy(mf_ls_t ls);
/* FP32 master weights (host, MF_LS_N_PARAMS floats); set refreshes the BF16 copy */
int mf_ls_set_params(mf_ls_t ls, const float* params, int64_t n);
int mf_ls_get_params(mf_ls_t ls, float* params, int64_t n);
/* versioned little-endian state blob (FP32 master, Adam moments, BF16 copy);
 * size query with buf == NULL */
int mf_ls_save(mf_ls_t ls, void* buf, uint64_t capacity, uint64_t* size);
int mf_ls_load(mf_ls_t ls, const void* buf, uint64_t size);
/* dst = assembled learned correction for the residual src (device FP64, n = mf_m) */
int mf_ls_apply(mf_ls_t ls, void* dst, const void* src, void* stream);
/* raw network on n rows: x (device, n x 20 FP32) -> y (device, n x 10 FP32), BF16 MFMA */
int mf_ls_forward(mf_ls_t ls, const void* x, int64_t n, void* y, void* stream);
/* per-cell features x (n_cells x 20) and targets t (n_cells x 10) of residual r and error e (device) */
int mf_ls_features(mf_ls_t ls, const void* r, const void* e, void* x, void* t, void* stream);
/* FP32 gradient (host, MF_LS_N_PARAMS) and loss of mean_rows |net(x) - t|^2 over n device rows */
int mf_ls_gradient(mf_ls_t ls, const void* x, const void* t, int64_t n, float* grad, double* loss, void* stream);
/* gradient on (x, t) then one Adam step (beta 0.9/0.999, eps 1e-8); loss before the step */
int mf_ls_train_step(mf_ls_t ls, const void* x, const void* t, int64_t n, double lr, double* loss, void* stream);
/* training on (residual, error) pairs of plain FP64-outer multigrid CG solves */
int mf_ls_train(mf_ls_t ls, mf_mg_t mg, const mf_train_options* opt, mf_train_stats* stats, void* stream);
/* flexible CG from x = 0: z = V(r), then z += omega L(r - A z); ls may be NULL */
int mf_solve_fcg_learned(mf_op_t A, mf_mg_t mg, mf_ls_t ls, void* x, const void* b, double rel_tol,
                         int32_t max_iter, const mf_fcg_options* opt, mf_fcg_stats* stats, void* stream);

/* Runs the device algorithm's patch/colour schedule on the host, in the
 * operator's precision, from the same tables the kernels read (mode 0: vmult,
 * 1: diagonal).  Host vectors.  A self-check of the setup for machines without
 * a GPU; not an operator path (mf_vmult never calls it). */
int mf_debug_host_apply(mf_op_t op, int32_t mode, void* dst, const void* src);

/* ---- matrix-based comparison path ----------------------------------------
 * The assembled sparse system matrix of a CG operator and its product, for
 * timing matrix-free against matrix-based vmult as the reference does.
 *   mf_get_system_matrix  Operator::get_system_matrix, CG/cells_batched/application_hp.cc:1411-1438
 *                         (pattern: DoFTools::make_sparsity_pattern, CG/cells_batched/application.cc:1946-1961)
 *   mf_csr_vmult          TrilinosWrappers::SparseMatrix::vmult, CG/cells_batched/application.cc:1964-1983
 * CSR rows follow the operator's DoF numbering (all components for
 * MF_CG_VECTOR, component-major like the vectors); columns ascend within a
 * row.  A Dirichlet row holds its diagonal 1 only and free rows hold no
 * constrained column, so A x equals mf_vmult up to summation order.  Values
 * are the element integrals summed over cells in the operator's internal cell
 * order, in FP64, then stored in the operator's precision.  DG operators are
 * rejected (the reference assembles CG matrices only).  A host_only operator
 * gives a host-only matrix (mf_csr_vmult fails on it). */
typedef struct mf_sparse* mf_csr_t;
typedef struct mf_csr_info {
  int64_t n_rows, nnz;
  int64_t n_row_blocks;    /* workgroups of mf_csr_vmult: consecutive rows, at most block_entries entries */
  int32_t precision;       /* of the stored values */
  int32_t block_entries;   /* and at most 256 rows per block */
  double algorithmic_bytes; /* per product: nnz (4 + w) + n_rows (8 + 2 w) */
} mf_csr_info;
int mf_get_system_matrix(mf_op_t op, mf_csr_t* out);
int mf_csr_vmult(mf_csr_t A, void* dst, const void* src, void* stream);   /* overwrites dst */
int mf_csr_get_info(mf_csr_t A, mf_csr_info* info);
/* host copy: ptr[n_rows + 1], col[nnz], val[nnz] (FP64 whatever the stored precision) */
int mf_csr_copy_to_host(mf_csr_t A, int64_t* ptr, int32_t* col, double* val);
int mf_csr_destroy(mf_csr_t A);

/* ---- device plumbing (host) ----------------------------------------------
 * Thin wrappers over the HIP runtime this library is linked against, so that
 * callers without their own HIP bindings (tests, benchmarks) share one runtime
 * with the operators.  No reference counterpart (the reference is CPU-only). */
int mf_device_count(int32_t* n);
int mf_device_name(char* buf, int32_t len);
int mf_malloc(void** ptr, uint64_t bytes);
int mf_free(void* ptr);
int mf_memcpy_htod(void* dst, const void* src, uint64_t bytes);
int mf_memcpy_dtoh(void* dst, const void* src, uint64_t bytes);
int mf_memcpy_dtod(void* dst, const void* src, uint64_t bytes, void* stream);
int mf_memset(void* ptr, int32_t value, uint64_t bytes, void* stream);
/* dst[i] = src[i] for n doubles with one thread per element over the whole
 * chip (unlike hipMemcpy's copy kernel, which runs on a few workgroups): the
 * plain copy of known traffic (8 n bytes read, 8 n written) against which
 * hardware byte counters are scaled */
int mf_copy_kernel(void* dst, const void* src, uint64_t n, void* stream);
int mf_synchronize(void);
int mf_stream_create(void** stream);
int mf_stream_destroy(void* stream);
int mf_stream_synchronize(void* stream);
int mf_event_create(void** event);
int mf_event_destroy(void* event);
int mf_event_record(void* event, void* stream);
int mf_event_elapsed_ms(void* start, void* stop, float* ms);

#ifdef __cplusplus
}
#endif
#endif
