// Smoothed-aggregation algebraic multigrid: the coarse solver of the hybrid
// multigrid when its coarsest geometric level is too large for the dense
// solver (external meshes, which have no h-levels).
//
// Counterpart of MGCoarseAMG (CG/cells_batched/application_hp.cc:1684-1731,
// 1918-1930): one Trilinos ML smoothed-aggregation V-cycle on the coarsest
// level's assembled matrix per coarse solve.
#pragma once
#include <memory>
#include <vector>

#include "matrix.h"
#include "mg.h"

namespace smf {

struct AmgLevel {
  Csr A;                     // this level's matrix (FP64)
  SysMatrix M;               // the same for the CSR-stream product (when every row fits a row block)
  Csr P, R;                  // prolongation from the next coarser level and R = P^T
  SysMatrix Rm;              // R for the CSR-stream product (few rows of many entries)
  DevBuf dinv, x, b, r, d, t;  // FP64 level vectors
  double lambda_max = 0;     // of D^-1 A (Chebyshev upper bound, 1.1 x estimate)
};

struct Amg {
  std::vector<AmgLevel> lv;  // [0] = the matrix the hierarchy was built from
  int n_dense = 0;           // the last level is solved with a dense FP64 inverse
  DevBuf d_inv;
  int smoother_degree = 2;   // Chebyshev-Jacobi steps before and after the coarse correction
                             // (the multigrid sets its own smoother degree)
  double smoothing_range = 30.0;
  double strength = 0.0;     // aggregation threshold theta
};

// A0: symmetric positive definite FP64 CSR, columns ascending (identity rows
// of constrained DoFs allowed: they stay out of the aggregates); levels are
// added until one has at most max_dense rows
std::unique_ptr<Amg> build_amg(const Csr& A0, int max_dense, double strength);
// lv[0].x = one V-cycle applied to lv[0].b, from x = 0 (device, stream-ordered)
void amg_vcycle(Amg& amg, hipStream_t s);

}  // namespace smf
