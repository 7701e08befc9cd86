// fv_f64.hip -- FP64 stage kernels; compiled with -ffp-contract=off so that
// every multiply and add is rounded separately, as in the reference.
#include "fv_launch_impl.hpp"

namespace sg {
int launch_stage_f64(const StageLaunch &L) { return launch_stage_impl<double, false>(L); }
}  // namespace sg
