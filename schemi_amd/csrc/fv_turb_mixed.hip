// fv_turb_mixed.hip -- mixed-precision turbulent stage kernels: closure and
// Riemann solve in FP32, accumulation and update in FP64.  Contraction stays
// off so that the two cells of a face evaluate it to the same bits.
#include "fv_turb_impl.hpp"

namespace sg {
int launch_turb_stage_mixed(const TurbStageLaunch &L) { return launch_turb_impl<float, true>(L); }
}  // namespace sg
