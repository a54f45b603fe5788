// Right-hand side b_i = int_K f phi_i, evaluated per cell on the device.
//
// The reference assembles the right-hand side of its Poisson solves with
// FEEvaluation at the quadrature points (CG/cells_batched/application_hp.cc:
// 1441-1465, f of 77-93).  Here each lane integrates one cell with a
// degree-6 rule and writes its local vector; the host adds the local vectors
// into the global one.
#include <hip/hip_runtime.h>

#include <cmath>

#include "op.h"

namespace smf {

namespace {
constexpr double kPi = 3.14159265358979323846;

// f at x: kind 0 -> 3 (3 pi)^2 prod sin(3 pi x_d); kind 1 -> random Fourier
// features sqrt(2/M) sum_m cos(omega_m . x + phase_m)
__device__ double rhs_f(int kind, const double* x, int nf, const double* __restrict__ feat) {
  if (kind == 0) {
    const double a = 3.0 * kPi;
    return 3.0 * a * a * sin(a * x[0]) * sin(a * x[1]) * sin(a * x[2]);
  }
  double s = 0.0;
  for (int m = 0; m < nf; ++m)
    s += cos(feat[4 * m] * x[0] + feat[4 * m + 1] * x[1] + feat[4 * m + 2] * x[2] + feat[4 * m + 3]);
  return sqrt(2.0 / nf) * s;
}

__global__ void rhs_cell_kernel(int64_t ncells, int nl, int nq, const double* __restrict__ geo12,
                                const double* __restrict__ qtab, int kind, int nf,
                                const double* __restrict__ feat, double* __restrict__ out) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= ncells) return;
  // geo12: x0 (3), J (9, row-major), per cell
  double g[12];
  for (int t = 0; t < 12; ++t) g[t] = geo12[12 * k + t];
  const double det = fabs(g[3] * (g[7] * g[11] - g[8] * g[10]) - g[4] * (g[6] * g[11] - g[8] * g[9]) +
                          g[5] * (g[6] * g[10] - g[7] * g[9]));
  double y[20];
  for (int i = 0; i < nl; ++i) y[i] = 0.0;
  // qtab: per point [xhat(3), w, phi(nl)]
  const int stride = 4 + nl;
  for (int q = 0; q < nq; ++q) {
    const double* Q = qtab + q * stride;
    double x[3];
    for (int a = 0; a < 3; ++a) x[a] = g[a] + g[3 + 3 * a] * Q[0] + g[4 + 3 * a] * Q[1] + g[5 + 3 * a] * Q[2];
    const double fw = rhs_f(kind, x, nf, feat) * Q[3] * det;
    for (int i = 0; i < nl; ++i) y[i] += fw * Q[4 + i];
  }
  for (int i = 0; i < nl; ++i) out[k * nl + i] = y[i];
}
}  // namespace

void launch_rhs_cells(int64_t ncells, int nl, int nq, const double* geo12, const double* qtab, int kind, int nf,
                      const double* feat, double* out, hipStream_t s) {
  if (ncells > 0)
    hipLaunchKernelGGL(rhs_cell_kernel, dim3((unsigned)((ncells + 127) / 128)), dim3(128), 0, s, ncells, nl, nq,
                       geo12, qtab, kind, nf, feat, out);
  SMF_HIP(hipGetLastError());
}

}  // namespace smf
