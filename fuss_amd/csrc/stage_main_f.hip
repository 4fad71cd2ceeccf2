// stage_main_f.hip -- FP32 stage launchers: 2-D, tangential and row-march
// 3-D kernels, and the External_Function path.
#include "stage_launch.h"

namespace fuss {
// the LDS-ring kernels are instantiated in stage_ring{4,8}_f.hip
extern template void launch_stage3p_ring<float, 4>(fuss_ctx*, Level&, const StageArgs<float>&, bool, bool);
extern template void launch_stage3p_ring<float, 8>(fuss_ctx*, Level&, const StageArgs<float>&, bool, bool);
// the register-column kernel in stage_reg_f.hip
extern template void launch_stage3r<float>(fuss_ctx*, Level&, const StageArgs<float>&, bool, bool);
template void launch_stage2d<float>(fuss_ctx*, Level&, const StageArgs<float>&, bool, bool, bool);
template void launch_stage3d<float>(fuss_ctx*, Level&, const StageArgs<float>&, bool, bool, bool);
template void launch_stage_raw<float>(fuss_ctx*, Level&, const StageArgs<float>&, int, bool);
template void launch_newstate<float>(fuss_ctx*, Level&, const StageArgs<float>&, bool, bool);
}  // namespace fuss
