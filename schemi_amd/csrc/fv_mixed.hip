// fv_mixed.hip -- mixed-precision stage kernels: reconstruction, limiter and
// Riemann solve in FP32, conserved accumulation and update in FP64.
#include "fv_launch_impl.hpp"

namespace sg {
int launch_stage_mixed(const StageLaunch &L) { return launch_stage_impl<float, true>(L); }
}  // namespace sg
