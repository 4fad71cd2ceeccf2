// irs_main_d.hip -- FP64 IRS launchers: per-direction lines, x-y pass, z
// march, tridiagonal lines.
#include "irs_launch.h"

namespace fuss {
// the fused x-y-z pass is instantiated in irs_xyz{4,8}_d.hip
extern template void launch_irs_xyz_ty<double, 4>(fuss_ctx*, Level&, const IrsArgs<double>&, bool);
extern template void launch_irs_xyz_ty<double, 8>(fuss_ctx*, Level&, const IrsArgs<double>&, bool);
template void launch_irs_xyz<double>(fuss_ctx*, Level&, const IrsArgs<double>&, bool);
template void launch_irs_z<double>(fuss_ctx*, Level&, const IrsArgs<double>&, bool);
template void launch_irs_line<double>(fuss_ctx*, Level&, int, int, const IrsArgs<double>&, bool, bool);
template void launch_irs_xy<double>(fuss_ctx*, Level&, int, const IrsArgs<double>&, bool);
template int launch_irs_tridiag<double>(fuss_ctx*, Level&, IrsArgs<double>, double*, double*, bool, int*);
}  // namespace fuss
