// DG SIPG kernels of degree 1 in FP64 (every operator kind, mode and kernel
// variant of dg_kernels.h).
#include "dg_kernels.h"

namespace smf {
template void launch_dg_degree<1, double>(const Operator&, int, double*, const double*, hipStream_t);
template void launch_dg_ep_degree<1, double>(const Operator&, const double*, RowEpilogue<double>, hipStream_t);
}  // namespace smf
