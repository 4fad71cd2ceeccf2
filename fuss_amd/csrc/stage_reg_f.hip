// stage_reg_f.hip -- FP32 register-column stage kernel (rectilinear levels).
#include "stage_reg_launch.h"

namespace fuss {
template void launch_stage3r<float>(fuss_ctx*, Level&, const StageArgs<float>&, bool, bool);
}  // namespace fuss
