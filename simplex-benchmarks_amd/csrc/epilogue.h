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

// Inputs of one row's epilogue, loaded ahead of its use so that a thread can
// keep several rows' loads in flight (epilogue_load, then epilogue_finish
// gives exactly apply_epilogue).
template <typename T>
struct EpilogueIn {
  T b, bnd, dinv, d, x;
};

template <typename T>
__device__ inline EpilogueIn<T> epilogue_load(const RowEpilogue<T>& ep, int64_t i) {
  EpilogueIn<T> in;
  in.b = ep.b[i];
  in.bnd = ep.bnd ? ep.bnd[i] : T(0);
  if (ep.kind == 1) {
    in.dinv = ep.dinv[i];
    in.d = ep.first ? T(0) : ep.d[i];
    in.x = ep.x[i];
  } else {
    in.dinv = in.d = in.x = T(0);
  }
  return in;
}

template <typename T>
__device__ inline void epilogue_finish(const RowEpilogue<T>& ep, int64_t i, T ax, const EpilogueIn<T>& in) {
  if (ep.bnd) ax += in.bnd;
  if (ep.kind == 1) {
    const T r = in.b - ax;
    const T di = ep.first ? ep.c2 * in.dinv * r : ep.c1 * in.d + ep.c2 * in.dinv * r;
    ep.d[i] = di;
    if (ep.xout) ep.xout[i] = in.x + di;
    else ep.x[i] = in.x + di;
  } else {
    ep.r[i] = in.b - ax;
  }
}

// Epilogue of rows base + j, j in [0, n), for the thread's share j = lane,
// lane + nt, ... (ax[j] in LDS): EB rows per batch, every load of a batch
// issued before its first use.
template <typename T, int EB>
__device__ inline void epilogue_rows(const RowEpilogue<T>& ep, int64_t base, int n, int lane, int nt, const T* ax) {
  for (int j0 = 0; j0 < n; j0 += EB * nt) {
    EpilogueIn<T> in[EB];
#pragma unroll
    for (int u = 0; u < EB; ++u) in[u] = epilogue_load(ep, base + min(j0 + u * nt + lane, n - 1));
#pragma unroll
    for (int u = 0; u < EB; ++u) {
      const int j = j0 + u * nt + lane;
      if (j < n) epilogue_finish(ep, base + j, ax[j], in[u]);
    }
  }
}

}  // namespace smf
