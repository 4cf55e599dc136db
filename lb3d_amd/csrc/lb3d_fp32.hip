// lb3d_fp32.hip -- FP32 deviation storage: populations are held as
// n_i - w_i*rho0 and updated in FP32 (fused multiply-adds allowed).
#include "lb3d_launch.h"

namespace lb3d {
Launcher *make_launcher_fp32(const lb3d_params &p) { return new LauncherT<float, true>(p); }
BatchLauncher *make_batch_launcher_fp32() { return new BatchLauncherT<float, true>(); }
}  // namespace lb3d
