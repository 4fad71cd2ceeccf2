// stage_main_d.hip -- FP64 stage launchers: 2-D, tangential and row-march
// 3-D kernels, and the External_Function path.
#include "stage_launch.h"

namespace fuss {
// the LDS-ring kernels are instantiated in stage_ring{4,8}_d.hip
extern template void launch_stage3p_ring<double, 4>(fuss_ctx*, Level&, const StageArgs<double>&, bool, bool);
extern template void launch_stage3p_ring<double, 8>(fuss_ctx*, Level&, const StageArgs<double>&, bool, bool);
// the register-column kernel in stage_reg_d.hip
extern template void launch_stage3r<double>(fuss_ctx*, Level&, const StageArgs<double>&, bool, bool);
template void launch_stage2d<double>(fuss_ctx*, Level&, const StageArgs<double>&, bool, bool, bool);
template void launch_stage3d<double>(fuss_ctx*, Level&, const StageArgs<double>&, bool, bool, bool);
template void launch_stage_raw<double>(fuss_ctx*, Level&, const StageArgs<double>&, int, bool);
template void launch_newstate<double>(fuss_ctx*, Level&, const StageArgs<double>&, bool, bool);
}  // namespace fuss
