// stage_ring8_d.hip -- FP64 LDS-ring stage kernel on 8-row columns.
#include "stage_launch.h"

namespace fuss {
template void launch_stage3p_ring<double, 8>(fuss_ctx*, Level&, const StageArgs<double>&, bool, bool);
}  // namespace fuss
