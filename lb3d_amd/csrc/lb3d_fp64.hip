// lb3d_fp64.hip -- plain FP64 form: the reference's values and operation
// order.  Built with -ffp-contract=off so that no multiply-add is fused.
#include "lb3d_launch.h"

namespace lb3d {
Launcher *make_launcher_fp64(const lb3d_params &p) { return new LauncherT<double, false>(p); }
BatchLauncher *make_batch_launcher_fp64() { return new BatchLauncherT<double, false>(); }
}  // namespace lb3d
