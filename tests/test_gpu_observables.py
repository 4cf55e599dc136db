"""Observables read back from the device: the Lees-Edwards x-halo planes, the
momentum field and the momentum transfer through x-planes (the reference's
shear stress, code/lb3d_io_stress.F90:38-173), against the oracle on the same
state."""
import numpy as np
import pytest

from lb3d_amd.gpu import LB3D_BC_LEES_EDWARDS, LB3D_FP32_DEV, LB3D_FP64, LB3DError, Lattice
from tests.test_gpu_parity import pair

pytestmark = pytest.mark.gpu

CX = np.array([1, -1, 0, 0, 0, 0, 1, 1, 1, 1, -1, -1, -1, -1, 0, 0, 0, 0, 0])
CY = np.array([0, 0, 1, -1, 0, 0, 1, -1, 0, 0, 1, -1, 0, 0, 1, 1, -1, -1, 0])
CZ = np.array([0, 0, 0, 0, 1, -1, 0, 0, 1, -1, 0, 0, 1, -1, 1, -1, 1, -1, 0])
G = np.array([2] * 6 + [1] * 13, dtype=float)


def oracle_halo(o):
    """x = 0 and x = nx+1 planes of the oracle's whole_N, shape (2, 2, 19, nz, ny)."""
    o.halo_refresh()
    N = o.N
    out = np.empty((2, 2, 19, o.nz, o.ny))
    for f, x in ((0, 0), (1, o.nx + 1)):
        for c in (0, 1):
            out[f, c] = N[1:-1, 1:-1, x, 24 * c:24 * c + 19].transpose(2, 0, 1)
    return out


@pytest.mark.parametrize("shear_u", [0.04, -0.3])
def test_le_halo_and_stress_match_oracle(shear_u):
    """Positive and negative offsets (lelo < 0 takes the other interpolation
    pair, and -0.3 drives the negativity clamp)."""
    o, g = pair(16, 12, 20, ncomp=2, g_br=0.09, boundary=LB3D_BC_LEES_EDWARDS, shear_u=shear_u,
                amp=(0.3, 0.3, 0.0) if shear_u < 0 else (0.01, 0.01, 0.0))
    o.run(7)
    g.time_loop(7)
    assert g.le_state()[1] == o.le_state()[1]
    want = oracle_halo(o)
    got = g.le_halo()
    d = np.abs(got - want).max()
    print(f"shear_u {shear_u}: max |halo diff| {d:.3e}")
    assert d <= 1e-12
    for planes in ((1, 16), (4, 9)):
        s_o = o.stress(*planes)
        s_g = g.shear_stress(*planes)
        np.testing.assert_allclose(s_g, s_o, rtol=1e-10, atol=1e-14)


def test_stress_periodic_matches_oracle_and_halo_needs_le():
    o, g = pair(12, 10, 8, ncomp=2, g_br=0.09)
    o.run(5)
    g.time_loop(5)
    o.halo_refresh()
    np.testing.assert_allclose(g.shear_stress(), o.stress(), rtol=1e-10, atol=1e-14)
    with pytest.raises(LB3DError):
        g.le_halo()


@pytest.mark.parametrize("ncomp", [2, 3])
def test_momentum_field(ncomp):
    """P_c = sum_i g_i n_c,i c_i of the device populations (FP64 form) and of
    the oracle's populations after the same steps."""
    o, g = pair(10, 9, 7, ncomp=ncomp, amp=(0.02, 0.02, 0.01))
    o.run(4)
    g.time_loop(4)
    P = g.momentum()
    n = o.populations()  # (C, 19, nz, ny, nx)
    want = np.stack([np.einsum("i,izyx->zyx", G * cv, n[c]) for c in range(ncomp)
                     for cv in (CX, CY, CZ)]).reshape(P.shape)
    np.testing.assert_allclose(P, want, rtol=0, atol=1e-13)


def test_fp32_le_halo_close_to_fp64():
    kw = dict(ncomp=2, g_br=0.08, boundary=LB3D_BC_LEES_EDWARDS, shear_u=0.05)
    _, g64 = pair(16, 16, 24, precision=LB3D_FP64, **kw)
    _, g32 = pair(16, 16, 24, precision=LB3D_FP32_DEV, **kw)
    g64.time_loop(20)
    g32.time_loop(20)
    h64, h32 = g64.le_halo(), g32.le_halo()
    rel = np.abs(h32 - h64).max() / np.abs(h64).max()
    print(f"FP32 vs FP64 halo planes: {rel:.3e}")
    assert rel < 1e-5
    s64, s32 = g64.shear_stress(), g32.shear_stress()
    print(f"mean shear stress FP64 {s64[2]:.6e} FP32 {s32[2]:.6e}")
    assert abs(s32[2] - s64[2]) <= 1e-3 * abs(s64[2]) + 1e-8


@pytest.mark.parametrize("ncomp", [2, 3])
def test_full_sanity_matches_oracle(ncomp):
    o, g = pair(11, 9, 7, ncomp=ncomp, amp=(0.02, 0.02, 0.01))
    o.run(6)
    g.time_loop(6)
    want, got = o.sanity_full(), g.sanity()
    for k in ("fluid_sites", "nan_sites", "negative_sites", "fast_sites"):
        assert got[k] == want[k], k
    np.testing.assert_allclose(got["mass"], want["mass"], rtol=1e-13)
    np.testing.assert_allclose(got["momentum"], want["momentum"], rtol=1e-9, atol=1e-13)
    np.testing.assert_allclose(got["vel_sum"], want["vel_sum"], rtol=1e-9, atol=1e-13)
    np.testing.assert_allclose(got["max_vel"], want["max_vel"], rtol=1e-12)
    np.testing.assert_array_equal(got["max_pos"], want["max_pos"])


@pytest.mark.parametrize("precision", [LB3D_FP64, LB3D_FP32_DEV])
def test_full_sanity_ties_keep_the_last_site(precision):
    """A uniform state: every site has the same velocity, and the maximum is
    reported at the last site of the reference's loop, (nx, ny, nz)."""
    g = Lattice(12, 10, 6, ncomp=2, precision=precision)
    g.init_noise(1, amp=(0.0, 0.0, 0.0), dipoles=False)
    s = g.sanity()
    assert s["fluid_sites"] == 720 and s["fast_sites"] == 0
    assert (s["max_pos"][:2] == [12, 10, 6]).all()
    assert (s["max_pos"][2] == 0).all()
