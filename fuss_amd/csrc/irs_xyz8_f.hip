// irs_xyz8_f.hip -- FP32 fused x-y-z IRS pass on 8-row columns.
#include "irs_launch.h"

namespace fuss {
template void launch_irs_xyz_ty<float, 8>(fuss_ctx*, Level&, const IrsArgs<float>&, bool);
}  // namespace fuss
