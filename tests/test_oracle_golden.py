"""The oracle against the reference's own golden tests (testLB3D.sh).

Each case restores the test's checkpoint, runs the reference's 10 timesteps and
compares the float32 dumps of t=10 (test/<name>/reference/*.xdr, committed as
tests/golden/reference_goldens.npz).  The reference compares these files
byte-exactly (testLB3D.sh:283-296); the oracle rounded to float32 must match
every value exactly.
"""
import os

import numpy as np
import pytest

from oracle.pyoracle import Oracle

GOLD = os.path.join(os.path.dirname(__file__), "golden", "reference_goldens.npz")

# parameters of test/<name>/<name>.in
CASES = {
    "gyroid": dict(ncomp=3, lees_edwards=0, g_br=0.08, g_bs=-0.006, g_ss=-0.003, beta=10.0,
                   d_0=1.0, tau_d=1.0, obs=["od", "wd", "sur", "dir", "vel", "arr"]),
    "leesEdwards": dict(ncomp=2, lees_edwards=1, g_br=0.1, shear_u=0.01, beta=10.0,
                        obs=["od", "wd", "vel", "arr"]),
    "bodyForce": dict(ncomp=3, lees_edwards=0, g_br=0.08, g_bs=-0.006, g_ss=0.003, beta=1.0,
                      d_0=1.0, tau_d=1.0, g_accn=(1e-5, 1e-5, 1e-5), g_accn_min=(10, 10, 10),
                      g_accn_max=(17, 17, 17), obs=["od", "wd", "sur", "dir", "vel", "arr"]),
    # binary with a wetting wall read from wettingWall.wall.xdr (rock_colour = 1
    # colours the rock red, g_wr = -0.1)
    "wettingWall": dict(ncomp=2, lees_edwards=0, g_br=0.12, g_wr=-0.1, rock_colour_r=1.0, beta=1.0,
                        obs=["od", "wd", "vel", "arr"]),
    # ternary with the Zou-He flux boundary on the z planes (boundary_cond = -1,
    # inv_fluid = 12, inv_type = -1, red mass flow pr = 1e-6)
    "zouHe": dict(ncomp=3, lees_edwards=0, g_br=0.08, g_bs=-0.006, g_ss=0.003, beta=1.0, d_0=1.0,
                  tau_d=1.0, inv_fluid=12, inv_type=-1, inv_flux=(1e-6, 0.0, 0.0),
                  obs=["od", "wd", "sur", "dir", "vel", "arr"]),
}


@pytest.fixture(scope="module")
def gold():
    return np.load(GOLD)


def restore_case(gold, name):
    kw = dict(CASES[name])
    kw.pop("obs")
    o = Oracle(16, 16, 16, restore=1, shear_sum=float(gold[f"{name}/shear_sum"][0]), **kw)
    o.load_checkpoint_array(gold[f"{name}/checkpoint"])
    if f"{name}/wall" in gold:
        # lb3d_ic_init stage 4: the checkpoint, then lbe_init_rock (code/lb3d_ic.F90:76-97)
        o.init_rock(gold[f"{name}/wall"], 0.0)
    o.bc_init()
    return o


@pytest.mark.parametrize("name", list(CASES))
def test_golden_dumps_exact(gold, name):
    o = restore_case(gold, name)
    o.run(10)
    assert o.nt == 10
    for obs in CASES[name]["obs"]:
        mine = o.dump(obs).astype(np.float32)
        ref = gold[f"{name}/{obs}"]
        assert mine.shape == ref.shape
        np.testing.assert_array_equal(mine, ref, err_msg=f"{name}/{obs}")


def test_golden_mass_gyroid_conserved(gold):
    """Periodic BGK conserves each component's mass to rounding."""
    o = restore_case(gold, "gyroid")
    m0, _, _ = o.sanity()
    o.run(10)
    m1, _, c = o.sanity()
    np.testing.assert_allclose(m1, m0, rtol=1e-13)
    assert c[0] == 0


def test_golden_le_shear_sum(gold):
    """LE restore bookkeeping: shear_sum is re-read, reduced by shear_u and
    re-added at init (code/lb3d_bc.F90:58-78), then +shear_u per step."""
    o = restore_case(gold, "leesEdwards")
    s0, lelo0, f0, u = o.le_state()
    assert u == pytest.approx(0.01 * 16 / 15, rel=0, abs=1e-17)
    assert s0 == pytest.approx(float(gold["leesEdwards/shear_sum"][0]), abs=1e-15)
    o.run(10)
    s1, lelo1, f1, _ = o.le_state()
    assert s1 == pytest.approx(s0 + 10 * u, abs=1e-14)
    assert lelo1 == int(2.0 * s1) % 16
    assert f1 == pytest.approx(abs(2 * s1 - int(2 * s1)), abs=1e-14)
