// irs_xyz8_d.hip -- FP64 fused x-y-z IRS pass on 8-row columns.
#include "irs_launch.h"

namespace fuss {
template void launch_irs_xyz_ty<double, 8>(fuss_ctx*, Level&, const IrsArgs<double>&, bool);
}  // namespace fuss
