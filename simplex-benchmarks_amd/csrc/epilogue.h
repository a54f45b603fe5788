// Device side of RowEpilogue (op.h): applied to row i of A x by the CG and
// DG operator kernels.
#pragma once
#include <hip/hip_runtime.h>

#include "op.h"

namespace smf {

template <typename T>
__device__ inline void apply_epilogue(const RowEpilogue<T>& ep, int64_t i, T ax) {
  if (ep.bnd) ax += ep.bnd[i];
  if (ep.kind == 1) {
    const T r = ep.b[i] - ax;
    const T di = ep.first ? ep.c2 * ep.dinv[i] * r : ep.c1 * ep.d[i] + ep.c2 * ep.dinv[i] * r;
    ep.d[i] = di;
    if (ep.xout) ep.xout[i] = ep.x[i] + di;
    else ep.x[i] += di;
  } else {
    ep.r[i] = ep.b[i] - ax;
  }
}

}  // namespace smf
