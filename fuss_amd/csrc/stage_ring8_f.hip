// stage_ring8_f.hip -- FP32 LDS-ring stage kernel on 8-row columns.
#include "stage_launch.h"

namespace fuss {
template void launch_stage3p_ring<float, 8>(fuss_ctx*, Level&, const StageArgs<float>&, bool, bool);
}  // namespace fuss
