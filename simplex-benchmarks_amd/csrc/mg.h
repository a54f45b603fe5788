// Hybrid (p- then h-) multigrid preconditioner and Krylov solvers.
//
// Counterpart of MultigridPreconditioner (CG/cells_batched/application_ph.cc:
// 1754-2026): level operators, MGTwoLevelTransfer per level pair,
// PreconditionChebyshev<..., DiagonalMatrix> smoothers (degree 5, smoothing
// range 20, 10 eigenvalue CG iterations), a coarse solver and SolverCG to
// 1e-10 ||b||.  The coarse solver here is an FP64 dense inverse instead of one
// Trilinos ML V-cycle (application_hp.cc:1684-1731).  Under a DG SIPG fine
// operator the hierarchy is DG P_k -> (c) CG P_k -> (p) ... P_1 -> (h), the
// CG levels with the Nitsche boundary (application_phc.cc:3416-3552).
#pragma once
#include <memory>
#include <vector>

#include "blas.h"
#include "matrix.h"
#include "op.h"

namespace smf {

struct Csr {
  int64_t nrows = 0, ncols = 0;
  std::vector<int64_t> ptr;
  std::vector<int32_t> col;
  std::vector<double> val;
  DevBuf d_ptr, d_col, d_val;
  void upload(int precision);
};

struct MGLevel {
  std::unique_ptr<Operator> op;  // level operator in the V-cycle precision
  int degree = 1, kuhn_n = 0;
  DevBuf inv_diag, x, b, r, t, d;  // level vectors
  DevBuf run_part[2], run_xi[2], run_di[2];  // ChebyRun buffers (scalar CG levels)
  Csr P, R;                        // prolongation from / restriction to the next coarser level
  // the same in row blocks for the CSR-stream product (coalesced entry loads;
  // rows of the restriction hold up to ~30 entries), when every row fits a block
  SysMatrix Pm, Rm;
  double lambda_max = 0, lambda_min = 0;
};

// dst = P src (dir 0) or R src (dir 1) of a level, dst += with add
template <typename T>
void level_transfer(const MGLevel& L, int dir, T* dst, const T* src, bool add, hipStream_t s);

struct MGOptions {
  int smoother_degree = 5;
  double smoothing_range = 20.0;
  int eig_cg_iterations = 10;
  int level_precision = MF_F32;
  int max_coarse_dofs = 4096;
  int fuse_steps = 1;  // scalar CG levels: Chebyshev steps without per-step fix-up launches
  // 0: dense FP64 inverse when the coarsest level has at most max_coarse_dofs
  // rows, else one smoothed-aggregation AMG V-cycle on its assembled matrix;
  // 1: dense only; 2: AMG always
  int coarse_solver = 0;
  // AMG aggregation threshold theta (strong: |a_ij| >= theta sqrt(a_ii a_jj));
  // 0.05 measured best on P1 levels of perturbed tetrahedral meshes (0: 14,
  // 0.05: 10, 0.1: 12, 0.2: 20 MG-CG iterations at 48^3 P2)
  double amg_threshold = 0.05;
};

struct Amg;

struct Multigrid {
  MGOptions opt;
  int fine_precision = MF_F64;
  int64_t n_fine = 0;
  std::vector<MGLevel> levels;  // [0] = coarsest
  // coarse solver: dense FP64 inverse of the coarsest operator
  int n_coarse = 0, n_coarse_free = 0;  // dense block: the leading n_coarse_free rows (the rest: identity)
  DevBuf d_coarse_inv, d_cb, d_cx;
  std::shared_ptr<Amg> amg;  // AMG coarse solver (instead of the dense inverse)
  DevBuf d_partial, d_scalar;  // reductions
  int64_t vcycles = 0;
};

std::unique_ptr<Multigrid> build_multigrid(const Operator& fine, const MGOptions& opt);
// z = V-cycle(r); r, z in the fine operator's precision (device)
void mg_vmult(Multigrid& mg, void* z, const void* r, hipStream_t s);

struct SolveStats {
  int iterations = 0;
  double r0 = 0, rn = 0, b_norm = 0;
  int converged = 0;
};
// (F)CG on the fine FP64 operator, x0 = 0 (the reference's SolverCG,
// application_hp.cc:1950-1960); mg may be null (unpreconditioned)
SolveStats solve_cg(const Operator& A, Multigrid* mg, double* x, const double* b, double rel_tol,
                    int max_iter, bool flexible, hipStream_t s);
// device work memory of solve_cg: four vectors of A.m() doubles, the
// reduction partials (blas_reduction_parts() doubles), one scalar and the
// Krylov scalar block (kS_N doubles); every entry is written before it is read
struct CgWorkspace {
  double *r, *z, *p, *q, *partial, *scalar, *S;
};
// doubles solve_cg_ws needs for n rows
int64_t cg_workspace_doubles(int64_t n);
// CgWorkspace carved out of `base` (cg_workspace_doubles(n) doubles, 256-byte aligned)
CgWorkspace cg_workspace_at(double* base, int64_t n);
// solve_cg on caller-provided work memory: allocates nothing on the device
SolveStats solve_cg_ws(const Operator& A, Multigrid* mg, double* x, const double* b, double rel_tol,
                       int max_iter, bool flexible, const CgWorkspace& ws, hipStream_t s);

// b_i = int f phi_i on a CG operator (Dirichlet rows 0): kind 0 f = 3(3pi)^2
// prod sin(3 pi x_d) (exact solution prod sin(3 pi x_d)); kind 1 a seeded
// Gaussian random field by random Fourier features (n_features, corr_length).
void assemble_rhs(const Operator& op, int kind, uint64_t seed, int n_features, double corr_length,
                  std::vector<double>& b);
// interpolant of the kind-0 exact solution at the DoF support points
void exact_solution(const Operator& op, std::vector<double>& u);

}  // namespace smf
