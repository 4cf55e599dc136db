// lb3d_fp32.hip -- FP32 deviation storage: populations are held as
// n_i - w_i*rho0 and updated in FP32 (fused multiply-adds allowed).
#include "lb3d_launch.h"

namespace lb3d {
Launcher *make_launcher_fp32(const lb3d_params &p) { return new LauncherT<float, true>(p); }
BatchLauncher *make_batch_launcher_fp32() { return new BatchLauncherT<float, true>(); }
// One two-pass timestep over the buffers of b (f[cur] -> f[1-cur], b.aux as
// the moments field), with a launcher on the stack: nothing is allocated.
void step_buffers_fp32(const lb3d_params &p, const DevBufs &b, hipStream_t s)
{
    LauncherT<float, true> L(p);
    L.moments(b, s);
    L.dipole(b, s);
    L.collide(b, s);
}
}  // namespace lb3d
