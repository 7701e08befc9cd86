// fv_turb_f64.hip -- FP64 turbulent stage kernels; compiled with
// -ffp-contract=off so that every multiply and add is rounded separately.
#include "fv_turb_impl.hpp"

namespace sg {
int launch_turb_stage_f64(const TurbStageLaunch &L) { return launch_turb_impl<double, false>(L); }
}  // namespace sg
