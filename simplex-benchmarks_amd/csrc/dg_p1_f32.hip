// DG SIPG kernels of degree 1 in FP32 (every operator kind, mode and kernel
// variant of dg_kernels.h).
#include "dg_kernels.h"

namespace smf {
template void launch_dg_degree<1, float>(const Operator&, int, float*, const float*, hipStream_t);
template void launch_dg_ep_degree<1, float>(const Operator&, const float*, RowEpilogue<float>, hipStream_t);
}  // namespace smf
