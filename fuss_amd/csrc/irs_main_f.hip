// irs_main_f.hip -- FP32 IRS launchers: per-direction lines, x-y pass, z
// march, tridiagonal lines.
#include "irs_launch.h"

namespace fuss {
// the fused x-y-z pass is instantiated in irs_xyz{4,8}_f.hip
extern template void launch_irs_xyz_ty<float, 4>(fuss_ctx*, Level&, const IrsArgs<float>&, bool);
extern template void launch_irs_xyz_ty<float, 8>(fuss_ctx*, Level&, const IrsArgs<float>&, bool);
template void launch_irs_xyz<float>(fuss_ctx*, Level&, const IrsArgs<float>&, bool);
template void launch_irs_z<float>(fuss_ctx*, Level&, const IrsArgs<float>&, bool);
template void launch_irs_line<float>(fuss_ctx*, Level&, int, int, const IrsArgs<float>&, bool, bool);
template void launch_irs_xy<float>(fuss_ctx*, Level&, int, const IrsArgs<float>&, bool);
template int launch_irs_tridiag<float>(fuss_ctx*, Level&, IrsArgs<float>, float*, float*, bool, int*);
}  // namespace fuss
