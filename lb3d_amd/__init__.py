"""MI355X-native LB3D timestep: D3Q19 BGK collision, streaming, Shan-Chen
forces with the amphiphile dipole coupling, periodic and Lees-Edwards
boundaries, as hand-written HIP kernels for gfx950 behind a C ABI
(include/lb3d_gpu.h).  The Python host is in lb3d_amd.gpu; the Fortran host
is in lb3d_amd/fortran."""
from .gpu import (LB3D_BC_LEES_EDWARDS, LB3D_BC_PERIODIC, LB3D_FP32_DEV, LB3D_FP64, LB3DError,
                  Lattice, lib)

__all__ = ["Lattice", "LB3DError", "lib", "LB3D_FP64", "LB3D_FP32_DEV", "LB3D_BC_PERIODIC",
           "LB3D_BC_LEES_EDWARDS"]
