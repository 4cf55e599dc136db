"""lb3d_gpu_step_noalloc (one timestep over caller-owned device buffers, the
moments field in a static device array) against lb3d_gpu_run_steps on a
context: bitwise, on two calls in a row, the second with a smaller lattice."""
import ctypes

import numpy as np
import pytest

from lb3d_amd.gpu import LB3D_FP32_DEV, LB3D_FP64, lib, make_params, Lattice
from oracle.pyoracle import Oracle

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def hip():
    h = ctypes.CDLL("libamdhip64.so")
    h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    h.hipFree.argtypes = [ctypes.c_void_p]
    h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    h.hipDeviceSynchronize.argtypes = []
    return h


class DevBuf:
    """A device buffer of the test (the entry point under test allocates nothing)."""

    def __init__(self, h, arr):
        self.h, self.nbytes, self.dtype, self.shape = h, arr.nbytes, arr.dtype, arr.shape
        self.p = ctypes.c_void_p()
        assert h.hipMalloc(ctypes.byref(self.p), self.nbytes) == 0
        a = np.ascontiguousarray(arr)
        assert h.hipMemcpy(self.p, a.ctypes.data_as(ctypes.c_void_p), self.nbytes, H2D) == 0

    def get(self):
        out = np.empty(self.shape, self.dtype)
        assert self.h.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), self.p, self.nbytes, D2H) == 0
        return out

    def free(self):
        self.h.hipFree(self.p)


def device_layout(o, ncomp, precision, rho0):
    """Populations f[(c*19 + i)*ns + s] and dipoles db[k*ns + s] of the oracle's
    state in the device layout of the precision."""
    I = o.interior()  # (nz, ny, nx, site doubles)
    ns = I.shape[0] * I.shape[1] * I.shape[2]
    w = np.array([1.0 / 36.0] * 18 + [1.0 / 3.0])
    f = np.empty((ncomp, 19, ns))
    for c in range(ncomp):
        v = I[..., 24 * c:24 * c + 19].reshape(ns, 19).T
        f[c] = v if precision == LB3D_FP64 else v - w[:, None] * rho0[c]
    dt = np.float64 if precision == LB3D_FP64 else np.float32
    db = I[..., 75:78].reshape(ns, 3).T.astype(dt) if ncomp == 3 else None
    return f.astype(dt).reshape(-1), db


def to_pops(fdev, ncomp, precision, rho0, ns):
    w = np.array([1.0 / 36.0] * 18 + [1.0 / 3.0])
    f = fdev.astype(np.float64).reshape(ncomp, 19, ns)
    if precision != LB3D_FP64:
        f = f + w[None, :, None] * np.array(rho0[:ncomp])[:, None, None]
    return f.reshape(-1)


def one_case(h, dims, ncomp, precision):
    kw = dict(ncomp=ncomp, precision=precision, g_br=0.08, g_bs=-0.006, g_ss=-0.003, rho0=(0.7, 0.7, 0.7))
    o = Oracle(*dims, ncomp=ncomp, g_br=0.08, g_bs=-0.006, g_ss=-0.003)
    o.init_noise(11, amp=(0.02, 0.02, 0.01))
    ref = Lattice(*dims, **kw)
    ref.upload_sites(o.N)
    ref.bc_init()
    ref.set_plan(0)
    ref.time_loop(1)
    want = ref.populations()
    want_db = ref.download_sites()[1:-1, 1:-1, 1:-1, 75:78] if ncomp == 3 else None
    ns = int(np.prod(dims))
    f, db = device_layout(o, ncomp, precision, kw["rho0"])
    fin, fout = DevBuf(h, f), DevBuf(h, np.zeros_like(f))
    dbin = DevBuf(h, db) if ncomp == 3 else None
    dbout = DevBuf(h, np.zeros_like(db)) if ncomp == 3 else None
    p = make_params(*dims, **kw)
    rc = lib().lb3d_gpu_step_noalloc(ctypes.byref(p), fin.p, fout.p, dbin.p if dbin else None,
                                     dbout.p if dbout else None, None)
    assert rc == 0, lib().lb3d_gpu_last_error()
    assert h.hipDeviceSynchronize() == 0
    got = to_pops(fout.get(), ncomp, precision, kw["rho0"], ns)
    np.testing.assert_array_equal(got, want.reshape(-1), err_msg=f"{dims} ncomp {ncomp}")
    if ncomp == 3:
        gdb = dbout.get().astype(np.float64).reshape(3, dims[2], dims[1], dims[0]).transpose(1, 2, 3, 0)
        np.testing.assert_array_equal(gdb, want_db)
    for b in (fin, fout, dbin, dbout):
        if b:
            b.free()
    ref.close()


@pytest.mark.parametrize("ncomp,precision", [(3, LB3D_FP32_DEV), (3, LB3D_FP64), (2, LB3D_FP32_DEV)])
def test_step_noalloc_equals_run_steps(ncomp, precision):
    h = hip()
    one_case(h, (24, 20, 16), ncomp, precision)
    one_case(h, (12, 10, 8), ncomp, precision)  # the second call: a smaller lattice


def test_step_noalloc_rejects_what_it_cannot_hold():
    # ternary FP64 moments: 80 bytes per site; 100^3 sites need more than 64 MiB
    p = make_params(100, 100, 100, ncomp=3, precision=LB3D_FP64)
    one = ctypes.c_void_p(16)
    two = ctypes.c_void_p(32)
    assert lib().lb3d_gpu_step_noalloc(ctypes.byref(p), one, two, one, two, None) != 0
    assert b"scratch" in lib().lb3d_gpu_last_error()
    assert lib().lb3d_gpu_step_noalloc(ctypes.byref(p), one, one, one, two, None) != 0
