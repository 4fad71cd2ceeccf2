// irs_xyz4_f.hip -- FP32 fused x-y-z IRS pass on 4-row columns.
#include "irs_launch.h"

namespace fuss {
template void launch_irs_xyz_ty<float, 4>(fuss_ctx*, Level&, const IrsArgs<float>&, bool);
}  // namespace fuss
