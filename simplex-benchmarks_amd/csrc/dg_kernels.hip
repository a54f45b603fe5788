// Dispatch of the DG SIPG operators to the per-degree, per-precision
// translation units (dg_p<P>_f<bits>.hip), which instantiate the kernels of
// dg_kernels.h.
#include <hip/hip_runtime.h>

#include <stdexcept>

#include "epilogue.h"
#include "op.h"

namespace smf {

template <int P, typename T>
void launch_dg_degree(const Operator& op, int mode, T* dst, const T* src, hipStream_t s);
template <int P, typename T>
void launch_dg_ep_degree(const Operator& op, const T* x, RowEpilogue<T> ep, hipStream_t s);

namespace {

template <typename T>
void launch_dg_prec(const Operator& op, int mode, T* dst, const T* src, hipStream_t s) {
  if (mode != 0 && mode != 1) throw std::runtime_error("DG: mode 0 (product) or 1 (diagonal)");
  switch (op.dg.p) {
    case 1: launch_dg_degree<1, T>(op, mode, dst, src, s); break;
    case 2: launch_dg_degree<2, T>(op, mode, dst, src, s); break;
    case 3: launch_dg_degree<3, T>(op, mode, dst, src, s); break;
    default: throw std::runtime_error("unsupported DG degree");
  }
}

}  // namespace

bool dg_has_epilogue(const Operator& op) {
  return op.kind == MF_DG_SIPG && !op.dg.curved &&
         (op.opt.dg_kernel == 3 || op.opt.dg_kernel == 4 || (op.opt.dg_kernel == 0 && op.dg.p >= 2));
}

template <typename T>
void launch_dg_epilogue(const Operator& op, const T* x, RowEpilogue<T> ep, hipStream_t s) {
  if (!dg_has_epilogue(op)) throw std::runtime_error("DG epilogue: scalar SIPG with the four-lane kernel");
  if ((op.precision == MF_F64) != (sizeof(T) == 8)) throw std::runtime_error("DG epilogue: precision mismatch");
  if (ep.kind == 1 && (ep.xout == nullptr || ep.xout == ep.x))
    throw std::runtime_error("DG epilogue: the Chebyshev step writes a separate output vector");
  if (ep.kind != 1 && ep.kind != 2) throw std::runtime_error("DG epilogue: kind 1 or 2");
  switch (op.dg.p) {
    case 1: launch_dg_ep_degree<1, T>(op, x, ep, s); break;
    case 2: launch_dg_ep_degree<2, T>(op, x, ep, s); break;
    case 3: launch_dg_ep_degree<3, T>(op, x, ep, s); break;
    default: throw std::runtime_error("unsupported DG degree");
  }
}
template void launch_dg_epilogue<double>(const Operator&, const double*, RowEpilogue<double>, hipStream_t);
template void launch_dg_epilogue<float>(const Operator&, const float*, RowEpilogue<float>, hipStream_t);

void launch_dg(const Operator& op, int mode, void* dst, const void* src, hipStream_t s) {
  if (op.precision == MF_F64)
    launch_dg_prec<double>(op, mode, static_cast<double*>(dst), static_cast<const double*>(src), s);
  else
    launch_dg_prec<float>(op, mode, static_cast<float*>(dst), static_cast<const float*>(src), s);
}

}  // namespace smf
