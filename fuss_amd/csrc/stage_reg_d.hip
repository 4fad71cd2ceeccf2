// stage_reg_d.hip -- FP64 register-column stage kernel (rectilinear levels).
#include "stage_reg_launch.h"

namespace fuss {
template void launch_stage3r<double>(fuss_ctx*, Level&, const StageArgs<double>&, bool, bool);
}  // namespace fuss
