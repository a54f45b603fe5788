// Learned per-cell smoother: a small MLP applied to every cell's local
// residual, trained on (residual, error) pairs produced by the plain
// multigrid solver, and used as an extra post-smoothing step inside flexible
// CG.  No counterpart in the reference (a new component on top of its
// MultigridPreconditioner + SolverCG, CG/cells_batched/application_ph.cc:
// 1754-2026 and application_hp.cc:1950-1960).
//
// Network: 20 inputs = 10 scaled local residuals + 9 entries of h J^-T +
// |det J| / h^3, two hidden layers of 64 GELU units, 10 outputs = the local
// correction in units of s_K * kappa_K (s_K = |r_K|_2, kappa_K = mean of
// the inverse diagonal over the cell's free DoFs).  6154 parameters, FP32
// master copy; inference with BF16 weights and FP32 accumulation on MFMA;
// training in FP32.
#pragma once
#include <cstdint>
#include <memory>
#include <vector>

#include "mg.h"

namespace smf {

constexpr int kMlpIn = 20, kMlpH = 64, kMlpOut = 10;
// parameter layout: W1[64][20] b1[64] W2[64][64] b2[64] W3[10][64] b3[10]
constexpr int kOffW1 = 0, kOffB1 = kOffW1 + kMlpH * kMlpIn, kOffW2 = kOffB1 + kMlpH,
              kOffB2 = kOffW2 + kMlpH * kMlpH, kOffW3 = kOffB2 + kMlpH, kOffB3 = kOffW3 + kMlpOut * kMlpH,
              kMlpParams = kOffB3 + kMlpOut;
// BF16 B-operand fragments of the 16x16x16 MFMA: layer 1 2x4, layer 2 4x4, layer 3 4x1
constexpr int kFragL1 = 0, kFragL2 = 8, kFragL3 = 24, kFrags = 28;

struct LearnedSmoother {
  const Operator* op = nullptr;  // fine CG P2 FP64 operator the smoother belongs to
  int64_t n_cells = 0, n_dofs = 0;
  double h = 1.0;
  DevBuf d_params, d_m, d_v;     // FP32 master weights and Adam moments
  DevBuf d_frag;                 // BF16 MFMA fragments (kFrags x 64 lanes x 4 bf16)
  DevBuf d_cell_dofs;            // 10 x n_cells (int32, component-major), -1 for Dirichlet DoFs
  DevBuf d_geo;                  // 10 x n_cells geometry features (FP32, component-major)
  DevBuf d_kappa;                // n_cells (FP32)
  DevBuf d_slot_ptr, d_slot;     // DoF -> (cell*10+i) slots, CSR
  DevBuf d_x, d_t, d_scale;      // per-cell features, targets, residual norms
  DevBuf d_corr;                 // n_cells x 10 (FP32)
  DevBuf d_dy;                   // n_cells x 10 output gradient (energy loss)
  DevBuf d_gpart, d_lpart, d_grad, d_loss;
  DevBuf d_partial, d_scalar;    // reductions
  DevBuf d_lsc;                  // learned step in the solver: c.r', c.Ac, omega, unusable flag
  DevBuf d_part2;                // partial sums of the learned step's two dot products
  double* h_lsc = nullptr;       // pinned host copy of d_lsc
  int64_t adam_step = 0;
  LearnedSmoother() = default;
  LearnedSmoother(const LearnedSmoother&) = delete;
  LearnedSmoother& operator=(const LearnedSmoother&) = delete;
  ~LearnedSmoother() {
    if (h_lsc) (void)hipHostFree(h_lsc);
  }
};

std::unique_ptr<LearnedSmoother> make_learned_smoother(const Operator& op, uint64_t seed);
void ls_set_params(LearnedSmoother& ls, const std::vector<float>& p);
std::vector<float> ls_get_params(const LearnedSmoother& ls);
// serialised state: "SMFMLP01", u32 version, u32 n_params, i64 adam_step,
// FP32 params, m, v, then the BF16 copy (u16) of the params; little endian
std::vector<uint8_t> ls_serialize(const LearnedSmoother& ls);
void ls_deserialize(LearnedSmoother& ls, const uint8_t* data, size_t size);

// corr = assembled learned correction for residual r (FP64 vectors of the fine operator)
void ls_apply(LearnedSmoother& ls, double* corr, const double* r, hipStream_t s);
// raw network on n feature rows (device, n x 20) -> y (device, n x 10), BF16 MFMA
void ls_forward(LearnedSmoother& ls, const float* x, int64_t n, float* y, hipStream_t s);
// features (n_cells x 20) and targets e_K / (s_K kappa_K) (n_cells x 10) of residual r, error e
void ls_features(LearnedSmoother& ls, const double* r, const double* e, float* x, float* t, hipStream_t s);
// FP32 gradient of the mean-squared loss over n rows; grad/loss in device memory of the smoother
void ls_gradient(LearnedSmoother& ls, const float* x, const float* t, int64_t n, hipStream_t s);
// one step of the energy loss on a pair (r', e'): L = |e' - w c|_A^2 / |e'|_A^2
// at the optimal w, c = assembled network correction of r'; w = 1 - cos^2_A(c, e').
// Leaves the gradient in the smoother; returns L before the step.
double ls_energy_gradient(LearnedSmoother& ls, const double* rp, const double* e, hipStream_t s);
// one Adam step (lr, beta 0.9/0.999, eps 1e-8) from the last gradient; refreshes the BF16 copy
void ls_adam(LearnedSmoother& ls, double lr, hipStream_t s);

struct TrainOptions {
  int loss = 0;              // 0: energy of the remaining error (FCG gain), 1: mean squared targets
  int n_rhs = 64;            // random-field right-hand sides
  int pairs_per_rhs = 6;     // preconditioned CG iterations used per RHS
  int steps_per_pair = 2;    // Adam steps per (residual, error) pair
  double lr = 1e-3;
  int n_features = 256;
  double corr_length = 0.1;
  uint64_t seed = 1;
  int trajectory = 0;        // iterates the pairs come from: 0 plain PCG, 1 FCG with the learned step
};
struct TrainStats {
  int pairs = 0, steps = 0;
  // loss before the first step of the first and the last pair; zero_loss:
  // loss of the zero predictor (1 for the energy loss)
  double first_loss = 0, last_loss = 0, zero_loss = 0;
};
TrainStats ls_train(LearnedSmoother& ls, Multigrid& mg, const TrainOptions& opt, hipStream_t s);

struct FcgOptions {
  double stall_ratio = 0.5;   // fall back when |r_k|/|r_{k-1}| exceeds this twice in a row
  int force_fallback_at = -1; // iteration at which to force the fallback (tests)
  int use_learned = 1;
};
struct FcgStats {
  SolveStats base;
  int fallback = 0, fallback_iteration = -1, learned_steps = 0;
  double mean_omega = 0;
};
// flexible CG: z = V(r); r' = r - A z; c = L(r'); z += omega c with the
// energy-optimal omega = (c.r')/(c.A c), formed on the device.  A z + omega
// A c gives A z for the search-direction update, so the outer loop needs no
// separate A p.
FcgStats solve_fcg_learned(const Operator& A, Multigrid& mg, LearnedSmoother* ls, double* x, const double* b,
                           double rel_tol, int max_iter, const FcgOptions& fo, hipStream_t s);

// launchers (mlp_kernels.hip)
void ls_launch_features(int64_t ncells, const int32_t* cdofs, const float* geo, const float* kappa, const double* r,
                        const double* e, float* x, float* t, float* scale, hipStream_t s);
void ls_launch_infer(int64_t ncells, const float* x, const void* frag, const float* params, const float* scale,
                     const float* kappa, float* corr, hipStream_t s);
void ls_launch_assemble(int64_t ndofs, const int64_t* ptr, const int32_t* slot, const float* corr, double* out,
                        hipStream_t s);
int ls_grad_blocks();
// dy != nullptr: output gradient given per cell (t unused, loss not formed)
void ls_launch_grad(int64_t n, const float* x, const float* t, const float* dy, const float* params, float* gpart,
                    float* lpart, float* grad, float* loss, hipStream_t s);
void ls_launch_dy(int64_t ncells, const int32_t* cdofs, const int64_t* ptr, const float* scale, const float* kappa,
                  const double* gc, double coef, float* dy, hipStream_t s);
// features formed in the kernel from r - az (az may be null); x_out / scale_out optional
void ls_launch_infer_fused(int64_t ncells, const int32_t* cdofs, const float* geo, const float* kappa,
                           const double* r, const double* az, float* x_out, float* scale_out, const void* frag,
                           const float* params, float* corr, hipStream_t s, double* direct = nullptr);
int ls_dot2_parts();
// sc = {c.(r - az), c.ac, omega, unusable}; z += omega c, azout += omega ac
// (az and azout may alias: az is only read before azout is written); part
// holds 2 * ls_dot2_parts() doubles
void ls_launch_omega(int64_t n, double* part, double* sc, const double* c, const double* r, const double* az,
                     const double* ac, double* z, double* azout, hipStream_t s);
void ls_launch_adam(float* p, const float* g, float* m, float* v, double lr, int64_t step, hipStream_t s);
void ls_launch_pack(const float* p, void* frag, hipStream_t s);

}  // namespace smf
