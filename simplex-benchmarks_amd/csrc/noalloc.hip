// (F)CG solve that allocates no device memory.
//
// mf_solve_cg takes its four work vectors, reduction partials and Krylov
// scalars from a fresh device allocation on every call.  Callers that cannot
// allocate (code under HIP graph capture, tight serving loops, harnesses that
// forbid hipMalloc) use mf_solve_cg_noalloc instead: the same solver and
// kernels, with the work memory in the file-scope device array `scratch`
// below.  Each GPU holds its own copy of the array (kScratchDoubles doubles,
// 32 MiB + 41 KiB) from the first call until the process ends; its device
// address is looked up once per device with hipGetSymbolAddress and kept.
// Every call on one GPU shares that GPU's array, so calls that overlap on one
// GPU race; the solver writes every entry of it before reading it.
#include <hip/hip_runtime.h>

#include <mutex>
#include <stdexcept>
#include <string>

#include "../../include/simplexmf.h"
#include "api_handles.h"
#include "blas.h"

namespace {

// the largest operator (rows) the no-allocation solve takes
constexpr int64_t kScratchRows = int64_t(1) << 20;
constexpr int64_t kScratchDoubles = 4 * kScratchRows + 1024 + 32 + 32 + 8;

static __device__ __align__(256) double scratch[kScratchDoubles];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch[kMaxDevices] = {};

// device address of `scratch` on the current device (looked up once per device)
double* scratch_on_current_device() {
  int dev = -1;
  SMF_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) throw std::runtime_error("device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch[dev]) {
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || ptr == nullptr)
      throw std::runtime_error(std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e));
    g_scratch[dev] = static_cast<double*>(ptr);
  }
  return g_scratch[dev];
}

}  // namespace

extern "C" int mf_solve_cg_noalloc(mf_op_t A, mf_mg_t mg, void* x, const void* b, double rel_tol, int32_t max_iter,
                                   int32_t flexible, mf_solve_stats* stats, void* stream) {
  MF_TRY({
    if (!A || !x || !b) throw std::runtime_error("null argument");
    require_device(A->op);
    if (A->op.kind != MF_CG_POISSON && A->op.kind != MF_DG_SIPG)
      throw std::runtime_error("solver needs a scalar CG Poisson or DG SIPG operator");
    if (mg && mg->mg->n_fine != A->op.m()) throw std::runtime_error("multigrid does not match the operator");
    const int64_t n = A->op.m();
    if (n > kScratchRows)
      throw std::runtime_error("mf_solve_cg_noalloc: operator has " + std::to_string(n) +
                               " rows; the no-allocation solve takes at most " + std::to_string(kScratchRows));
    if (smf::cg_workspace_doubles(n) > kScratchDoubles) throw std::runtime_error("scratch layout too small");
    const smf::CgWorkspace ws = smf::cg_workspace_at(scratch_on_current_device(), n);
    smf::SolveStats st = smf::solve_cg_ws(A->op, mg ? mg->mg.get() : nullptr, static_cast<double*>(x),
                                          static_cast<const double*>(b), rel_tol, max_iter, flexible != 0, ws,
                                          static_cast<hipStream_t>(stream));
    if (stats) {
      stats->iterations = st.iterations;
      stats->converged = st.converged;
      stats->initial_residual = st.r0;
      stats->final_residual = st.rn;
      stats->rhs_norm = st.b_norm;
    }
  });
}
