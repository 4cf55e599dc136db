"""Batched ensemble: lattices of uneven sizes and their own parameters in one
set of persistent launches, checked lattice by lattice against the oracle and
against single-lattice contexts."""
import numpy as np
import pytest

from lb3d_amd.gpu import LB3D_FP32_DEV, LB3D_FP64, Batch, Lattice
from oracle.pyoracle import Oracle

pytestmark = pytest.mark.gpu

DIMS = [(8, 8, 8), (13, 7, 5), (16, 12, 9), (5, 9, 11), (24, 16, 8)]


def lattice_kw(k, ncomp):
    return dict(g_br=0.06 + 0.01 * k, g_bs=-0.006 + 0.001 * k, g_ss=-0.003, beta=10.0 - k, tau_d=1.0 + 0.1 * k)


@pytest.mark.parametrize("ncomp", [2, 3])
def test_batch_fp64_matches_oracle_per_lattice(ncomp):
    per = [lattice_kw(k, ncomp) for k in range(len(DIMS))]
    b = Batch(DIMS, per_lattice=per, ncomp=ncomp, precision=LB3D_FP64)
    oracles = []
    for k, d in enumerate(DIMS):
        o = Oracle(*d, ncomp=ncomp, **per[k])
        o.init_noise(100 + k, amp=(0.02, 0.02, 0.01))
        b.upload_sites(k, o.N)
        o.bc_init()
        oracles.append(o)
    steps = 12
    b.time_loop(steps)
    assert b.nt == steps
    for k, o in enumerate(oracles):
        o.run(steps)
        d = np.abs(b.populations(k) - o.populations()).max()
        print(f"lattice {k} {DIMS[k]} max|dn| {d:.3e}")
        assert d <= 1e-11
        m, _, _ = b.sanity_check(k)
        om, _, _ = o.sanity()
        np.testing.assert_allclose(m[:ncomp], om[:ncomp], rtol=1e-13)


@pytest.mark.parametrize("precision", [LB3D_FP64, LB3D_FP32_DEV])
def test_batch_persistent_equals_sequential_and_single(precision):
    ncomp = 3
    per = [lattice_kw(k, ncomp) for k in range(len(DIMS))]

    def make():
        b = Batch(DIMS, per_lattice=per, ncomp=ncomp, precision=precision)
        for k in range(len(DIMS)):
            b.init_noise(k, 7 + k, amp=(0.01, 0.01, 0.01))
        return b

    b1, b2, b3, b4, b5, b6 = make(), make(), make(), make(), make(), make()
    b1.set_plan(0, 7)       # persistent, a small grid that strides many times
    b2.set_plan(1)          # lattice by lattice
    b3.autotune(2)          # whatever the tuner picks; the state must be unchanged
    b4.set_plan(0, 24, 3)   # persistent, chunks of 3 work items per XCD
    b5.set_plan(0, 16)      # persistent in groups: every lattice its own group
    b5.set_groups(1)
    rs = 8 if precision == LB3D_FP64 else 4
    b6.set_plan(0, 4096)    # groups of about two lattices, one work item per workgroup
    b6.set_groups(57 * rs * (8 * 8 * 8 + 13 * 7 * 5))
    bs = (b1, b2, b3, b4, b5, b6)
    for b in bs:
        b.time_loop(9)
    for k, d in enumerate(DIMS):
        g = Lattice(*d, ncomp=ncomp, precision=precision, **per[k])
        g.init_noise(7 + k, amp=(0.01, 0.01, 0.01))
        g.bc_init()
        g.time_loop(9)
        want = g.populations()
        for b in bs:
            np.testing.assert_array_equal(b.populations(k), want)
        sb, sg = b1.sanity(k), g.sanity()
        np.testing.assert_array_equal(sb["max_pos"], sg["max_pos"])
        np.testing.assert_array_equal(sb["max_vel"], sg["max_vel"])
        assert sb["fluid_sites"] == np.prod(d)


@pytest.mark.parametrize("ncomp,precision", [(3, LB3D_FP32_DEV), (2, LB3D_FP64)])
def test_batch_mixed_equilibrium_forms(ncomp, precision):
    """Lattices with bdist 0 and 2 in one batch: each collide launch serves the
    lattices of its own form (and the grouped plan groups of either form);
    every lattice equals its single-lattice run bitwise."""
    forms = [0, 2, 2, 0, 2]
    per = [dict(lattice_kw(k, ncomp), bdist=forms[k]) for k in range(len(DIMS))]

    def make():
        b = Batch(DIMS, per_lattice=per, ncomp=ncomp, precision=precision)
        for k in range(len(DIMS)):
            b.init_noise(k, 31 + k, amp=(0.01, 0.01, 0.01))
        return b

    plans = [make(), make(), make()]
    plans[0].set_plan(0, 64, 1)   # persistent, one group
    plans[1].set_plan(0, 64, 1)   # persistent, every lattice its own group
    plans[1].set_groups(1)
    plans[2].set_plan(1)          # lattice by lattice
    for b in plans:
        b.time_loop(7)
    for k, d in enumerate(DIMS):
        g = Lattice(*d, ncomp=ncomp, precision=precision, **per[k])
        g.init_noise(31 + k, amp=(0.01, 0.01, 0.01))
        g.bc_init()
        g.time_loop(7)
        want = g.populations()
        for j, b in enumerate(plans):
            np.testing.assert_array_equal(b.populations(k), want, err_msg=f"plan {j} lattice {k} bdist {forms[k]}")
        g.close()
    for b in plans:
        b.close()


CM_DIMS = [(20, 13, 11), (7, 5, 3), (33, 9, 6), (24, 24, 24), (16, 12, 9), (2, 3, 2)]


def cm_batch(forms, seed0=50):
    per = [dict(lattice_kw(k, 3), bdist=forms[k % len(forms)]) for k in range(len(CM_DIMS))]
    b = Batch(CM_DIMS, per_lattice=per, ncomp=3, precision=LB3D_FP32_DEV)
    for k in range(len(CM_DIMS)):
        b.init_noise(k, seed0 + k, amp=(0.02, 0.02, 0.01))
    return b, per


@pytest.mark.parametrize("cols,lz,thr", [(16, 32, 256), (14, 1, 256), (30, 2, 512), (6, 5, 256), (254, 64, 256),
                                         (62, 8, 512), (510, 3, 512)])
def test_collide_moments_plan_equals_three_pass_bitwise(cols, lz, thr):
    """The collide-moments plan (moments gathered from the post-collision
    populations, edge sites pulled) against the persistent three-pass plan
    and single-lattice runs: lattices that do not divide into tiles, z-chunks
    of 1 and 2 planes (every site an edge), both equilibrium forms, odd step
    counts around the captured two-step graph."""
    forms = [2, 0, 2, 2, 0, 2]
    ref, per = cm_batch(forms)
    ref.set_plan(0, 64, 1)
    ref.time_loop(7)
    b, _ = cm_batch(forms)
    b.set_collide_moments(cols, lz, 4, tile_threads=thr)
    b.time_loop(3)
    b.time_loop(4)
    assert b.nt == 7
    for k, d in enumerate(CM_DIMS):
        np.testing.assert_array_equal(b.populations(k), ref.populations(k), err_msg=f"lattice {k} {d}")
    g = Lattice(*CM_DIMS[3], ncomp=3, precision=LB3D_FP32_DEV, **per[3])
    g.init_noise(53, amp=(0.02, 0.02, 0.01))
    g.bc_init()
    g.time_loop(7)
    np.testing.assert_array_equal(b.populations(3), g.populations())


def test_collide_moments_plan_switches_and_reinit():
    """The moments carried between steps are rebuilt whenever they may be
    stale: after another plan's steps, a re-initialised lattice, the tuner."""
    forms = [2]
    ref, _ = cm_batch(forms)
    ref.set_plan(0, 64, 1)
    ref.time_loop(3)
    ref.init_noise(2, 77, amp=(0.02, 0.02, 0.01))
    ref.time_loop(6)
    b, _ = cm_batch(forms)
    b.set_collide_moments(16, 8, 16)
    b.time_loop(2)
    b.set_plan(0, 64, 1)
    b.time_loop(1)
    b.init_noise(2, 77, amp=(0.02, 0.02, 0.01))
    b.set_collide_moments(24, 4, 1, tile_threads=512)
    b.time_loop(3)
    b.autotune(1)
    b.set_collide_moments(16, 32, 16)
    ms, per_step = b.profile_steps(1)
    assert per_step > 0
    b.time_loop(2)
    assert b.nt == 6  # init_noise restarts the batch's step count
    for k in range(len(CM_DIMS)):
        np.testing.assert_array_equal(b.populations(k), ref.populations(k), err_msg=f"lattice {k}")


@pytest.mark.parametrize("ncomp,precision", [(2, LB3D_FP64), (3, LB3D_FP32_DEV), (3, LB3D_FP64)])
def test_batch_every_equilibrium_form_with_body_force(ncomp, precision):
    """Lattices of every equilibrium form of code/lb3d_bdist.F90 in one batch,
    with unequal relaxation times and a body-force box (Guo's term of bdist 4
    and 5 on inside it): the forms other than 0 and 2 share one collide
    launch; every lattice equals its single-lattice run bitwise, on the
    persistent, grouped and sequential plans, and the FP64 ones the oracle."""
    forms = [0, 1, 2, 3, 4, 5, 6]
    dims = [(8, 8, 8), (13, 7, 5), (16, 12, 9), (5, 9, 11), (10, 6, 7), (7, 7, 7), (9, 5, 6)]
    per = [dict(lattice_kw(k, ncomp), bdist=forms[k], tau=(0.9 + 0.02 * k, 1.1, 1.0),
                g_accn=(2e-4, -1e-4, 3e-4), g_accn_min=(2, 2, 2), g_accn_max=(4, 5, 4)) for k in range(len(forms))]

    def make():
        b = Batch(dims, per_lattice=per, ncomp=ncomp, precision=precision)
        for k in range(len(dims)):
            b.init_noise(k, 60 + k, amp=(0.02, 0.02, 0.01))
        return b

    plans = [make(), make(), make()]
    plans[0].set_plan(0, 64, 1)
    plans[1].set_plan(0, 64, 1)
    plans[1].set_groups(1)
    plans[2].set_plan(1)
    for b in plans:
        b.time_loop(7)
    for k, d in enumerate(dims):
        g = Lattice(*d, ncomp=ncomp, precision=precision, **per[k])
        g.init_noise(60 + k, amp=(0.02, 0.02, 0.01))
        g.bc_init()
        g.time_loop(7)
        want = g.populations()
        for j, b in enumerate(plans):
            np.testing.assert_array_equal(b.populations(k), want, err_msg=f"plan {j} lattice {k} bdist {forms[k]}")
        g.close()
        if precision == LB3D_FP64:
            o = Oracle(*d, ncomp=ncomp, **per[k])
            o.init_noise(60 + k, amp=(0.02, 0.02, 0.01))
            o.bc_init()
            o.run(7)
            assert np.abs(plans[0].populations(k) - o.populations()).max() <= 1e-11
    for b in plans:
        b.close()
