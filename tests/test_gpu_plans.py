"""Execution plans: the one-pass z-march kernel against the two-pass kernels
and the oracle, over tile shapes, z-chunks and lattice sizes that do not
divide into tiles; the autotuner leaves the lattice state unchanged."""
import numpy as np
import pytest

from lb3d_amd.gpu import LB3D_BC_LEES_EDWARDS, LB3D_FP32_DEV, LB3D_FP64, LB3DError, Lattice
from tests.test_gpu_parity import assert_pops_close, pair

pytestmark = pytest.mark.gpu

SHAPES = [(64, 4), (32, 8), (64, 8), (32, 16), (128, 4), (32, 4), (128, 2), (256, 2)]


def fused_shapes(precision):
    """(mode, tx, ty) of the one-pass kernel that exist for this form:
    mode 1 stages the look-ahead plane in LDS, mode 2 keeps it in registers,
    mode 3 pulls plane k again in the collide."""
    g = Lattice(8, 8, 8, ncomp=2, precision=precision)
    ok = []
    for mode in (1, 2, 3):
        for tx, ty in SHAPES:
            try:
                g.set_plan(mode, tx, ty, 4)
                ok.append((mode, tx, ty))
            except LB3DError:
                pass
    return ok


@pytest.mark.parametrize("precision", [LB3D_FP64, LB3D_FP32_DEV])
@pytest.mark.parametrize("dims,le", [((20, 13, 11), False), ((7, 5, 3), False), ((70, 9, 6), True),
                                      ((16, 12, 20), True)])
def test_fused_equals_two_pass_bitwise(precision, dims, le):
    kw = dict(ncomp=2, precision=precision, g_br=0.09)
    if le:
        kw.update(boundary=LB3D_BC_LEES_EDWARDS, shear_u=0.04)
    _, ref = pair(*dims, **kw)
    ref.set_plan(0)
    ref.time_loop(9)
    want = ref.populations()
    for mode, tx, ty in fused_shapes(precision):
        for lz in (1, 4, dims[2]):
            _, g = pair(*dims, **kw)
            g.set_plan(mode, tx, ty, lz)
            g.time_loop(9)
            got = g.populations()
            assert np.array_equal(got, want), \
                f"{dims} le={le} mode {mode} {tx}x{ty} lz={lz}: max {np.abs(got - want).max()}"


@pytest.mark.parametrize("ncomp,precision", [(2, LB3D_FP32_DEV), (3, LB3D_FP64)])
def test_xcd_grouping_does_not_change_results(ncomp, precision):
    """The work-group -> tile placement (chunks of xcd_group tiles per XCD)
    covers every tile once for any chunk size, including a last chunk that is
    narrower than the others (48x40x36: 270 site tiles, 34 per XCD)."""
    dims = (48, 40, 36)
    kw = dict(ncomp=ncomp, precision=precision, g_br=0.09, amp=(0.02, 0.02, 0.01))
    _, ref = pair(*dims, **kw)
    ref.set_plan(0, xcd_group=1)
    ref.time_loop(5)
    want = ref.populations()
    plans = [(0, 64, 4, 1, xg) for xg in (3, 4, 0, 1000)]
    if ncomp == 2:
        plans += [(1, 32, 8, 5, xg) for xg in (1, 3, 0)] + [(1, 128, 4, 36, 0)]
    for mode, tx, ty, lz, xg in plans:
        _, g = pair(*dims, **kw)
        g.set_plan(mode, tx, ty, lz, xcd_group=xg)
        g.time_loop(5)
        assert np.array_equal(g.populations(), want), f"plan {mode} {tx}x{ty} lz={lz} xcd_group={xg}"


def test_one_pass_plan_rejected_where_it_does_not_exist():
    for kw in (dict(ncomp=3), dict(ncomp=2, bdist=0)):
        g = Lattice(8, 8, 8, **kw)
        with pytest.raises(LB3DError):
            g.set_plan(1, 64, 4, 8)


def test_fused_fp64_matches_oracle_32cubed_100_steps():
    o, g = pair(32, 32, 32, ncomp=2, g_br=0.08)
    g.set_plan(1, 32, 4, 16)
    o.run(100)
    g.time_loop(100)
    assert_pops_close(o, g, what="fused binary 32^3 x100")


def test_fused_staged_surface_and_download_moments():
    """With the one-pass plan, advection() only advances the stage; the
    downloaded rho are still those of the last advection."""
    o, g = pair(12, 10, 8, ncomp=2)
    g.set_plan(1, 32, 4, 8)
    for _ in range(3):
        g.time_step()
    o.run(3)
    assert_pops_close(o, g, what="fused staged x3")
    N = g.download_sites()
    np.testing.assert_allclose(N[1:-1, 1:-1, 1:-1, 19], o.interior()[..., 19], rtol=0, atol=1e-14)
    np.testing.assert_allclose(N[1:-1, 1:-1, 1:-1, 43], o.interior()[..., 43], rtol=0, atol=1e-14)


@pytest.mark.parametrize("ncomp", [2, 3])
def test_autotune_keeps_state_and_picks_a_plan(ncomp):
    kw = dict(ncomp=ncomp, amp=(0.01, 0.01, 0.01))
    o, g = pair(24, 16, 12, **kw)
    o.run(2)
    g.time_loop(2)
    before = g.populations()
    rep = g.autotune(trial_steps=2)
    print(rep)
    assert "chosen:" in rep
    np.testing.assert_array_equal(g.populations(), before)
    o.run(5)
    g.time_loop(5)
    assert_pops_close(o, g, what=f"after autotune ncomp={ncomp}")


TERN_SHAPES = [(32, 4), (32, 8), (16, 16), (64, 4), (64, 8), (32, 16)]
TERN_KW = dict(ncomp=3, g_br=0.08, g_bs=-0.006, g_ss=-0.003, amp=(0.02, 0.02, 0.01))


def fused3_shapes(precision):
    """Tile shapes of the ternary one-pass kernel (mode 3) that exist for this form."""
    g = Lattice(8, 8, 8, ncomp=3, precision=precision)
    ok = []
    for tx, ty in TERN_SHAPES:
        try:
            g.set_plan(3, tx, ty, 4)
            ok.append((tx, ty))
        except LB3DError:
            pass
    return ok


def test_ternary_one_pass_shapes_exist():
    assert len(fused3_shapes(LB3D_FP32_DEV)) == 6
    assert (32, 4) in fused3_shapes(LB3D_FP64)
    g = Lattice(8, 8, 8, ncomp=3)
    with pytest.raises(LB3DError):
        g.set_plan(2, 32, 4, 4)  # the register and LDS stages are binary only


@pytest.mark.parametrize("precision", [LB3D_FP64, LB3D_FP32_DEV])
@pytest.mark.parametrize("dims,bdist", [((20, 13, 11), 2), ((7, 5, 3), 2), ((33, 9, 6), 0)])
def test_ternary_one_pass_equals_three_pass_bitwise(precision, dims, bdist):
    """Populations and the dipole field after 7 steps, for every tile shape,
    z-chunks of 1, 3 and nz, lattices that do not divide into tiles."""
    kw = dict(TERN_KW, precision=precision, bdist=bdist)
    _, ref = pair(*dims, **kw)
    ref.set_plan(0)
    ref.time_loop(7)
    want = ref.populations()
    want_db = ref.download_sites()[1:-1, 1:-1, 1:-1, 75:78]
    for tx, ty in fused3_shapes(precision):
        for lz in (1, 3, dims[2]):
            _, g = pair(*dims, **kw)
            g.set_plan(3, tx, ty, lz, xcd_group=0)
            g.time_loop(7)
            got = g.populations()
            assert np.array_equal(got, want), \
                f"{dims} bdist {bdist} {tx}x{ty} lz={lz}: max {np.abs(got - want).max()}"
            assert np.array_equal(g.download_sites()[1:-1, 1:-1, 1:-1, 75:78], want_db)


def test_ternary_one_pass_fp64_matches_oracle_and_downloads():
    """24x16x12 ternary, 20 steps against the oracle; the staged surface and
    the downloaded moments (rho, da of the last advection) as the oracle's."""
    o, g = pair(24, 16, 12, **TERN_KW)
    g.set_plan(3, 32, 8, 4)
    o.run(17)
    g.time_loop(17)
    for _ in range(3):
        g.time_step()
    o.run(3)
    assert_pops_close(o, g, what="ternary one-pass 24x16x12 x20")
    N = g.download_sites()
    I = o.interior()
    for off in (19, 43, 67, 72, 73, 74, 75, 76, 77):
        np.testing.assert_allclose(N[1:-1, 1:-1, 1:-1, off], I[..., off], rtol=0, atol=1e-13, err_msg=str(off))


def test_ternary_one_pass_graph_after_staged_step_and_reinit():
    """The dipole field follows the population parity: a graph captured after
    an odd number of staged steps, a re-initialised state and the per-kernel
    profiling all leave the ternary one-pass plan equal to the three-pass plan."""
    kw = dict(TERN_KW, precision=LB3D_FP32_DEV)
    dims = (20, 13, 11)
    _, ref = pair(*dims, **kw)
    ref.set_plan(0)
    ref.time_loop(4)
    want = ref.populations()
    want_db = ref.download_sites()[1:-1, 1:-1, 1:-1, 75:78]

    _, g = pair(*dims, **kw)
    g.set_plan(3, 32, 4, 4)
    g.time_step()  # odd parity before the graph is captured
    g.time_loop(3)
    assert np.array_equal(g.populations(), want)
    assert np.array_equal(g.download_sites()[1:-1, 1:-1, 1:-1, 75:78], want_db)

    # profile_steps advances the state like time_loop
    _, p = pair(*dims, **kw)
    p.set_plan(3, 32, 4, 4)
    p.time_step()
    p.profile_steps(3)
    assert np.array_equal(p.populations(), want)
    assert np.array_equal(p.download_sites()[1:-1, 1:-1, 1:-1, 75:78], want_db)

    # init after an odd step count, with a graph already captured
    _, h = pair(*dims, **kw)
    h.set_plan(3, 32, 4, 4)
    h.time_loop(3)
    h.init_noise(5, amp=(0.02, 0.02, 0.01))
    h.time_loop(2)
    fresh = Lattice(*dims, **{k: v for k, v in kw.items() if k != "amp"})
    fresh.set_plan(3, 32, 4, 4)
    fresh.init_noise(5, amp=(0.02, 0.02, 0.01))
    fresh.time_loop(2)
    assert np.array_equal(h.populations(), fresh.populations())
    assert np.array_equal(h.download_sites()[1:-1, 1:-1, 1:-1, 75:78],
                          fresh.download_sites()[1:-1, 1:-1, 1:-1, 75:78])


def test_autotune_keeps_state_and_plan_runs_after_it():
    """The tuner's trials write f[1-cur] and the spare dipole field only."""
    kw = dict(TERN_KW, precision=LB3D_FP32_DEV)
    dims = (24, 16, 12)
    _, ref = pair(*dims, **kw)
    ref.set_plan(0)
    ref.time_loop(5)
    want = ref.populations()
    _, g = pair(*dims, **kw)
    g.set_plan(3, 32, 8, 4)
    g.time_loop(3)
    g.autotune(1)
    g.set_plan(0)
    g.time_loop(2)
    assert np.array_equal(g.populations(), want)


@pytest.mark.parametrize("precision", [LB3D_FP64, LB3D_FP32_DEV])
def test_lees_edwards_edge_columns_follow_every_state_change(precision):
    """On the two-pass plan the Lees-Edwards halo planes read compact copies of
    the edge columns that the moments and collide passes write.  Every way the
    state changes outside those passes (one-pass steps, the staged surface,
    the tuner's trials, profiled steps, an upload) must leave the run bitwise
    equal to the straight two-pass run."""
    kw = dict(ncomp=2, precision=precision, g_br=0.09, boundary=LB3D_BC_LEES_EDWARDS, shear_u=0.04)
    dims = (20, 9, 12)
    _, ref = pair(*dims, **kw)
    ref.set_plan(0)
    ref.time_loop(12)
    want = ref.populations()
    mode, tx, ty = fused_shapes(precision)[0]
    o, g = pair(*dims, **kw)
    g.set_plan(0)
    g.time_loop(2)                 # graph-replayed two-pass steps
    g.set_plan(mode, tx, ty, 4)
    g.time_loop(3)                 # one-pass steps: the edge copies are not written
    g.set_plan(0)
    g.time_step()                  # a staged two-pass step after them
    g.time_loop(1)
    g.autotune(1)                  # the trials overwrite the edge copies
    g.set_plan(0)
    g.time_loop(2)
    g.profile_steps(2)
    g.time_loop(1)
    assert g.nt == 12
    got = g.populations()
    assert np.array_equal(got, want), f"max {np.abs(got - want).max()}"
    # an upload replaces the state under the edge copies of the last step
    g.upload_sites(o.N)
    g.bc_init()
    g.time_loop(12)
    got = g.populations()
    assert np.array_equal(got, want), f"after upload: max {np.abs(got - want).max()}"
