"""Parity of the HIP path with the CPU oracle on the same inputs, through the C ABI.

FP64 plain form: populations after N steps within 1e-11 absolute of the
oracle (the contract of the FP64 path), component masses to 1e-13 relative.
The golden cases additionally reproduce the reference's own float32 dumps.
"""
import os

import numpy as np
import pytest

from lb3d_amd.gpu import LB3D_BC_LEES_EDWARDS, LB3D_FP32_DEV, LB3D_FP64, LB3DError, Lattice
from oracle.pyoracle import Oracle

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden", "reference_goldens.npz")
POP_TOL = 1e-11
MASS_RTOL = 1e-13


def pair(nx, ny, nz, seed=1, amp=(0.01, 0.01, 0.0), rho=(0.7, 0.7, 0.7), dipoles=True, **kw):
    """An oracle lattice with seeded noise and a GPU lattice holding the same state."""
    okw = dict(kw)
    okw.pop("precision", None)
    if okw.pop("boundary", 0) == LB3D_BC_LEES_EDWARDS:
        okw["lees_edwards"] = 1
    o = Oracle(nx, ny, nz, **okw)
    o.init_noise(seed, rho_mean=rho, amp=amp, dipoles=dipoles)
    g = Lattice(nx, ny, nz, **kw)
    g.upload_sites(o.N)
    o.bc_init()
    g.bc_init()
    return o, g


def assert_pops_close(o, g, tol=POP_TOL, what=""):
    a, b = o.populations(), g.populations()
    d = np.abs(a - b).max()
    print(f"{what} max|dn| = {d:.3e}", flush=True)
    assert d <= tol, f"{what}: max |dn| {d:.3e} > {tol}"


def test_device_init_matches_oracle_init():
    for ncomp in (2, 3):
        o = Oracle(9, 7, 5, ncomp=ncomp)
        o.init_noise(3, amp=(0.01, 0.02, 0.005))
        g = Lattice(9, 7, 5, ncomp=ncomp)
        g.init_noise(3, amp=(0.01, 0.02, 0.005))
        np.testing.assert_array_equal(g.populations(), o.populations())
        if ncomp == 3:
            db = g.download_sites()[1:-1, 1:-1, 1:-1, 75:78]
            np.testing.assert_allclose(db, o.interior()[..., 75:78], rtol=0, atol=1e-15)


def test_binary_periodic_fp64_32cubed_100_steps():
    """The parity configuration: binary Shan-Chen, 32^3 periodic, 100 steps."""
    o, g = pair(32, 32, 32, ncomp=2, g_br=0.08)
    m0, _, _ = g.sanity_check()
    o.run(100)
    g.time_loop(100)
    assert g.nt == 100
    assert_pops_close(o, g, what="binary 32^3 x100")
    m1, p1, c1 = g.sanity_check()
    om, op, oc = o.sanity()
    np.testing.assert_allclose(m1[:2], m0[:2], rtol=MASS_RTOL)
    np.testing.assert_allclose(m1[:2], om[:2], rtol=MASS_RTOL)
    np.testing.assert_allclose(p1, op, rtol=0, atol=1e-12)
    np.testing.assert_array_equal(c1, oc)


def test_ternary_periodic_fp64():
    o, g = pair(20, 18, 16, ncomp=3, g_br=0.08, g_bs=-0.006, g_ss=-0.003, beta=10.0, tau_d=1.0,
                amp=(0.01, 0.01, 0.01))
    o.run(40)
    g.time_loop(40)
    assert_pops_close(o, g, what="ternary 20x18x16 x40")
    db_g = g.download_sites()[1:-1, 1:-1, 1:-1, 75:78]
    np.testing.assert_allclose(db_g, o.interior()[..., 75:78], rtol=0, atol=1e-11)
    m1, _, _ = g.sanity_check()
    om, _, _ = o.sanity()
    np.testing.assert_allclose(m1, om, rtol=MASS_RTOL)


def test_lees_edwards_binary_fp64():
    o, g = pair(16, 12, 20, ncomp=2, g_br=0.08, boundary=LB3D_BC_LEES_EDWARDS, shear_u=0.05)
    o.run(60)
    g.time_loop(60)
    assert_pops_close(o, g, what="LE 16x12x20 x60")
    so, lo, fo, uo = o.le_state()
    sg, lg, fg, ug = g.le_state()
    assert (lg, sg, fg, ug) == (lo, so, fo, uo)


def test_lees_edwards_offset_state_across_step_paths():
    """The Lees-Edwards offset state advances once per step whichever path ran
    the step: staged steps (one kernel updates it), graph-replayed and
    profiled two-pass steps (the halo-plane kernels compute it, the collide
    stores it), the tuner's trials (they leave it unchanged) and a one-pass
    step; populations and (shear_sum, lelo, lefrac, shear_u) as the oracle's."""
    o, g = pair(16, 12, 20, ncomp=2, g_br=0.08, boundary=LB3D_BC_LEES_EDWARDS, shear_u=0.07)
    g.set_plan(0)
    for _ in range(3):
        g.time_step()
    g.time_loop(5)
    g.profile_steps(2)
    g.autotune(2)
    for mode, tx, ty in ((3, 32, 8), (2, 32, 8), (1, 32, 8), (3, 64, 4)):
        try:
            g.set_plan(mode, tx, ty, 8)
            break
        except LB3DError:
            continue
    else:
        raise AssertionError("no one-pass plan for this lattice")
    g.time_loop(1)
    g.set_plan(0)
    g.time_loop(3)
    o.run(14)
    assert_pops_close(o, g, what="LE across step paths x14")
    assert g.le_state() == o.le_state()


def test_lees_edwards_negative_shear_and_clamp():
    """Negative shear (the lelo < 0 branch) and a strong shear that makes the
    Galilean correction trip the negativity clamp."""
    o, g = pair(10, 6, 14, ncomp=2, g_br=0.05, boundary=LB3D_BC_LEES_EDWARDS, shear_u=-0.3,
                amp=(0.3, 0.3, 0.0), rho=(0.5, 0.4, 0.0))
    o.run(12)
    g.time_loop(12)
    assert_pops_close(o, g, what="LE negative shear")


def test_odd_sizes_and_bdist0():
    for kw in (dict(ncomp=2, bdist=0), dict(ncomp=3, bdist=0), dict(ncomp=2, bdist=2)):
        o, g = pair(7, 5, 3, amp=(0.02, 0.02, 0.02), **kw)
        o.run(15)
        g.time_loop(15)
        assert_pops_close(o, g, what=f"7x5x3 {kw}")


def test_unequal_tau_amass_and_body_force():
    kw = dict(ncomp=3, tau=(0.8, 1.3, 1.1), amass=(1.0, 0.9, 1.2), tau_d=2.0, beta=3.0, d_0=0.9,
              g_accn=(1e-5, -2e-5, 3e-5), g_accn_min=(3, 2, 4), g_accn_max=(9, 7, 10))
    o, g = pair(12, 10, 14, amp=(0.01, 0.01, 0.01), **kw)
    o.run(20)
    g.time_loop(20)
    assert_pops_close(o, g, what="unequal tau/amass + body force")


@pytest.mark.parametrize("bdist", [1, 3, 4, 5, 6])
@pytest.mark.parametrize("ncomp", [2, 3])
def test_other_equilibrium_forms(bdist, ncomp):
    """The remaining forms of code/lb3d_bdist.F90 (1: iterated exponential,
    3: second order, 4: Guo, 5: Guo to fourth order, 6: fourth order), with a
    body-force box so that Guo's (tau - 1/2) g term is on inside and off
    outside it; FP32 deviation storage follows the FP64 form."""
    kw = dict(ncomp=ncomp, bdist=bdist, g_br=0.08, tau=(0.9, 1.2, 1.0),
              g_accn=(2e-4, -1e-4, 3e-4), g_accn_min=(2, 3, 2), g_accn_max=(6, 8, 5))
    o, g = pair(9, 10, 8, amp=(0.03, 0.03, 0.02), **kw)
    o.run(12)
    g.time_loop(12)
    assert_pops_close(o, g, what=f"bdist {bdist} ncomp {ncomp}")
    _, g32 = pair(9, 10, 8, amp=(0.03, 0.03, 0.02), precision=LB3D_FP32_DEV, **kw)
    g32.time_loop(12)
    a, b = g.populations(), g32.populations()
    rel = np.abs(a - b).max() / np.abs(a).max()
    print(f"bdist {bdist} ncomp {ncomp}: FP32 vs FP64 {rel:.3e}")
    assert rel < 2e-6


def test_staged_surface_equals_graph_replay():
    """The reference's stage calls and the captured-graph loop give identical bits."""
    for kw in (dict(ncomp=3), dict(ncomp=2, boundary=LB3D_BC_LEES_EDWARDS, shear_u=0.05)):
        o1, g1 = pair(12, 8, 10, amp=(0.01, 0.01, 0.01), **kw)
        _, g2 = pair(12, 8, 10, amp=(0.01, 0.01, 0.01), **kw)
        for _ in range(5):
            g1.time_step()
        g2.time_loop(5)
        np.testing.assert_array_equal(g1.populations(), g2.populations())
        assert g1.nt == g2.nt == 5


def test_stage_order_is_enforced():
    g = Lattice(8, 8, 8, ncomp=2)
    g.init_noise(1)
    g.bc_init()
    with pytest.raises(LB3DError) as e:
        g.collision()
    assert e.value.code == 3
    with pytest.raises(LB3DError):
        g.advection()
    g.bc_apply(6, 1)
    with pytest.raises(LB3DError):
        g.bc_apply(8, 1)
    g.advection()
    g.bc_apply(8, 1)
    g.collision()
    assert g.nt == 1


@pytest.mark.parametrize("name", ["gyroid", "leesEdwards", "bodyForce", "wettingWall", "zouHe"])
def test_reference_golden_tests_on_gpu(name):
    """Restore the reference test's checkpoint on the GPU, run its 10 steps,
    and reproduce its float32 dumps exactly (od, wd, sur, dir, vel, arr)."""
    from tests.test_oracle_golden import CASES
    gold = np.load(GOLD)
    kw = dict(CASES[name])
    obs = kw.pop("obs")
    le = kw.pop("lees_edwards")
    o = Oracle(16, 16, 16, restore=1, lees_edwards=le, shear_sum=float(gold[f"{name}/shear_sum"][0]), **kw)
    o.load_checkpoint_array(gold[f"{name}/checkpoint"])
    if f"{name}/wall" in gold:
        o.init_rock(gold[f"{name}/wall"], 0.0)  # the walls travel with rock_state in the sites
    g = Lattice(16, 16, 16, restore=1, boundary=LB3D_BC_LEES_EDWARDS if le else 0,
                shear_sum=float(gold[f"{name}/shear_sum"][0]), **kw)
    g.upload_sites(o.N)
    o.bc_init()
    g.bc_init()
    o.run(10)
    g.time_loop(10)
    assert_pops_close(o, g, tol=1e-13, what=f"golden {name}")
    # observables of the GPU state, evaluated by the oracle's dump routines
    chk = Oracle(16, 16, 16, restore=1, lees_edwards=le, shear_sum=0.0, **kw)
    chk.N[...] = g.download_sites()
    for ob in obs:
        mine = chk.dump(ob).astype(np.float32)
        np.testing.assert_array_equal(mine, gold[f"{name}/{ob}"], err_msg=f"{name}/{ob}")
    # the library's own arr observable (forces on the device, the reference's
    # sums on the host) against the oracle's dump routine and the golden
    arr = g.arrows()
    np.testing.assert_allclose(arr, chk.dump("arr"), rtol=0, atol=1e-13, err_msg=f"{name}/arr")
    np.testing.assert_array_equal(arr.astype(np.float32), gold[f"{name}/arr"], err_msg=f"{name}/arr float32")


@pytest.mark.parametrize("kw", [dict(ncomp=2, g_br=0.08), dict(ncomp=3, g_br=0.08, g_bs=-0.006, g_ss=-0.003),
                                dict(ncomp=2, g_br=0.09, boundary=LB3D_BC_LEES_EDWARDS, shear_u=0.04)])
def test_arrows_match_oracle_and_fp32(kw):
    """arr after 12 steps on a lattice that does not divide into tiles: FP64
    against the oracle's dump_arrows of the GPU state, the FP32 form within
    its rounding, and LB3D_ESTATE before the first advection."""
    dims = (13, 10, 7)
    o, g = pair(*dims, **kw, amp=(0.02, 0.02, 0.01))
    with pytest.raises(LB3DError):
        g.arrows()
    g.time_loop(12)
    okw = {k: v for k, v in kw.items() if k not in ("boundary",)}
    if kw.get("boundary") == LB3D_BC_LEES_EDWARDS:
        okw["lees_edwards"] = 1
    chk = Oracle(*dims, **okw)
    chk.N[...] = g.download_sites()
    want = chk.dump("arr")
    np.testing.assert_allclose(g.arrows(), want, rtol=0, atol=1e-13)
    _, f = pair(*dims, **kw, amp=(0.02, 0.02, 0.01), precision=LB3D_FP32_DEV)
    f.time_loop(12)
    np.testing.assert_allclose(f.arrows(), want, rtol=0, atol=2e-6)


@pytest.mark.parametrize("itype", [-1, 0, 1])
@pytest.mark.parametrize("ncomp,dims", [(3, (9, 7, 6)), (2, (8, 6, 5)), (3, (6, 5, 1))])
def test_zou_he_flux_boundary(itype, ncomp, dims):
    """The Zou-He flux boundary on the z planes (code/lb3d_invasion.F90:
    2143-2457) with fluxes large enough to show in FP64 (the reference's
    zouHe golden uses 1e-6), against the oracle; nz = 1 applies both rules to
    one plane.  FP32 deviation storage follows the FP64 form."""
    kw = dict(ncomp=ncomp, g_br=0.08, inv_fluid=12, inv_type=itype, inv_flux=(0.004, -0.002, 0.003))
    o, g = pair(*dims, amp=(0.02, 0.02, 0.01), **kw)
    o.run(12)
    g.time_loop(12)
    assert_pops_close(o, g, what=f"Zou-He itype {itype} ncomp {ncomp} {dims}")
    N = g.download_sites()
    I = o.interior()
    if ncomp == 3:
        np.testing.assert_allclose(N[1:-1, 1:-1, 1:-1, 72:75], I[..., 72:75], rtol=0, atol=1e-13)
    _, g32 = pair(*dims, amp=(0.02, 0.02, 0.01), precision=LB3D_FP32_DEV, **kw)
    g32.time_loop(12)
    a, b = g.populations(), g32.populations()
    rel = np.abs(a - b).max() / np.abs(a).max()
    print(f"Zou-He FP32 vs FP64 {rel:.3e}")
    assert rel < 2e-6
    # the one-pass plans do not take the boundary
    with pytest.raises(LB3DError):
        g.set_plan(3, 32, 4, 4)


def test_walls_fp32_close_to_fp64_and_rules():
    """The wetting-wall case in the FP32 form tracks FP64; walls are refused
    where they are not supported (ternary, Lees-Edwards, the batch) and the
    one-pass plans are unavailable with them."""
    from tests.test_oracle_golden import CASES
    gold = np.load(GOLD)
    kw = dict(CASES["wettingWall"])
    kw.pop("obs")
    kw.pop("lees_edwards")
    o = Oracle(16, 16, 16, **kw)
    o.load_checkpoint_array(gold["wettingWall/checkpoint"])
    o.init_rock(gold["wettingWall/wall"], 0.0)
    o.bc_init()
    res = {}
    for prec in (LB3D_FP64, LB3D_FP32_DEV):
        g = Lattice(16, 16, 16, precision=prec, rho0=(0.35, 0.35, 0.35), **kw)
        g.upload_sites(o.N)
        g.bc_init()
        with pytest.raises(LB3DError):
            g.set_plan(1, 32, 4, 8)
        g.time_loop(50)
        res[prec] = g.populations()
        N = g.download_sites()
        assert (N[1:-1, 1:-1, 1:-1, 48] == o.interior()[..., 48]).all()
        s = g.sanity()
        assert s["fluid_sites"] == 4096 - 256 and s["nan_sites"] == 0
    d = np.abs(res[LB3D_FP32_DEV] - res[LB3D_FP64]).max()
    print(f"walls FP32 vs FP64 after 50 steps: {d:.3e}")
    assert d < 1e-6
    t = Oracle(16, 16, 16, ncomp=3)
    t.init_rock(gold["wettingWall/wall"], 0.0)
    for bad in (dict(ncomp=3), dict(ncomp=2, boundary=LB3D_BC_LEES_EDWARDS, shear_u=0.01)):
        g = Lattice(16, 16, 16, **bad)
        with pytest.raises(LB3DError):
            g.upload_sites(t.N if bad["ncomp"] == 3 else o.N)


def test_fp32_deviation_binary_density_error():
    """FP32 deviation storage against FP64 on the same input: density error."""
    _, g64 = pair(32, 32, 32, ncomp=2)
    _, g32 = pair(32, 32, 32, ncomp=2, precision=LB3D_FP32_DEV)
    g64.time_loop(200)
    g32.time_loop(200)
    G = np.array([2] * 6 + [1] * 13, dtype=float)
    r64 = np.einsum("i,cizyx->czyx", G, g64.populations())
    r32 = np.einsum("i,cizyx->czyx", G, g32.populations())
    rel = np.abs(r32 - r64).max() / np.abs(r64).max()
    print(f"fp32 binary 32^3 x200 Linf rel rho error {rel:.3e}")
    assert rel < 1e-5
    m64, _, _ = g64.sanity_check()
    m32, _, _ = g32.sanity_check()
    np.testing.assert_allclose(m32[:2], m64[:2], rtol=1e-6)


def test_fp32_deviation_ternary_and_le():
    G = np.array([2] * 6 + [1] * 13, dtype=float)
    for kw in (dict(ncomp=3, amp=(0.01, 0.01, 0.01)),
               dict(ncomp=2, boundary=LB3D_BC_LEES_EDWARDS, shear_u=0.05)):
        _, g64 = pair(16, 14, 18, **kw)
        _, g32 = pair(16, 14, 18, precision=LB3D_FP32_DEV, **kw)
        g64.time_loop(100)
        g32.time_loop(100)
        r64 = np.einsum("i,cizyx->czyx", G, g64.populations())
        r32 = np.einsum("i,cizyx->czyx", G, g32.populations())
        rel = np.abs(r32 - r64).max() / np.abs(r64).max()
        print(f"fp32 {kw} x100 Linf rel rho error {rel:.3e}")
        assert rel < 1e-5


def test_walls_graph_dropped_when_walls_go():
    """A lattice that ran with walls, then re-initialised without them, steps
    like a fresh lattice (the captured graph of the wall kernels is dropped)."""
    from tests.test_oracle_golden import CASES
    gold = np.load(GOLD)
    kw = dict(CASES["wettingWall"])
    kw.pop("obs")
    kw.pop("lees_edwards")
    o = Oracle(16, 16, 16, **kw)
    o.load_checkpoint_array(gold["wettingWall/checkpoint"])
    o.init_rock(gold["wettingWall/wall"], 0.0)
    g = Lattice(16, 16, 16, **kw)
    g.upload_sites(o.N)
    g.bc_init()
    g.time_loop(2)
    g.init_noise(7)
    g.time_loop(4)
    fresh = Lattice(16, 16, 16, **kw)
    fresh.init_noise(7)
    fresh.time_loop(4)
    assert np.array_equal(g.populations(), fresh.populations())
