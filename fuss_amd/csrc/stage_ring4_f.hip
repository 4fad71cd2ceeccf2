// stage_ring4_f.hip -- FP32 LDS-ring stage kernel on 4-row columns.
#include "stage_launch.h"

namespace fuss {
template void launch_stage3p_ring<float, 4>(fuss_ctx*, Level&, const StageArgs<float>&, bool, bool);
}  // namespace fuss
