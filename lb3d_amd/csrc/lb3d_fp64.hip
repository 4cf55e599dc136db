// lb3d_fp64.hip -- plain FP64 form: the reference's values and operation
// order.  Built with -ffp-contract=off so that no multiply-add is fused.
#include "lb3d_launch.h"

namespace lb3d {
Launcher *make_launcher_fp64(const lb3d_params &p) { return new LauncherT<double, false>(p); }
BatchLauncher *make_batch_launcher_fp64() { return new BatchLauncherT<double, false>(); }
// One two-pass timestep over the buffers of b (f[cur] -> f[1-cur], b.aux as
// the moments field), with a launcher on the stack: nothing is allocated.
void step_buffers_fp64(const lb3d_params &p, const DevBufs &b, hipStream_t s)
{
    LauncherT<double, false> L(p);
    L.moments(b, s);
    L.dipole(b, s);
    L.collide(b, s);
}
}  // namespace lb3d
