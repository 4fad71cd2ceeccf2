// irs_xyz4_d.hip -- FP64 fused x-y-z IRS pass on 4-row columns.
#include "irs_launch.h"

namespace fuss {
template void launch_irs_xyz_ty<double, 4>(fuss_ctx*, Level&, const IrsArgs<double>&, bool);
}  // namespace fuss
