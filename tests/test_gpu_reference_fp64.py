"""The FP64 GPU path against the reference LB3D itself.

The seeded initial state of each case in tests/golden/reference_fp64.npz
(written by tests/golden/make_ref_fp64.py from runs of the reference built by
oracle/ref_build.sh) is uploaded through the C ABI and run 100 timesteps on
the default plan and on the two-pass plan; every population and dipole must
be within 1e-11 of the reference's (the 32^3 lattice: every density), and
each component's mass must be conserved to 1e-13 relative.
"""
import importlib.util
import json
import os

import numpy as np
import pytest

from lb3d_amd.gpu import LB3D_BC_LEES_EDWARDS, LB3D_FP64, Lattice
from oracle import refrun

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
_spec = importlib.util.spec_from_file_location("make_ref_fp64", os.path.join(GOLDEN, "make_ref_fp64.py"))
make_ref_fp64 = importlib.util.module_from_spec(_spec)
_spec.loader.exec_module(make_ref_fp64)


@pytest.fixture(scope="module")
def ref():
    return np.load(os.path.join(GOLDEN, "reference_fp64.npz"))


@pytest.mark.parametrize("plan", ["default", "two-pass"])
@pytest.mark.parametrize("name", list(make_ref_fp64.CASES))
def test_gpu_fp64_matches_reference(ref, name, plan):
    s = json.loads(ref[f"{name}/spec"].tobytes().decode())
    c = make_ref_fp64.CASES[name]
    ternary, shear = c["ternary"], c["shear"]
    o = make_ref_fp64.initial_oracle(c)
    kw = dict(c["phys"], ncomp=3 if ternary else 2, restore=1, shear_sum=0.0)
    if shear is not None:
        kw.update(boundary=LB3D_BC_LEES_EDWARDS, shear_u=shear)
    g = Lattice(*c["dims"], precision=LB3D_FP64, **kw)
    g.upload_sites(o.N)
    g.bc_init()
    if plan == "two-pass":
        g.set_plan(0)
    g.time_loop(s["steps"])
    assert g.nt == s["steps"]
    nc = kw["ncomp"]
    if f"{name}/out" in ref.files:
        want = ref[f"{name}/out"]
        got = refrun.records_from_sites(g.download_sites()[1:-1, 1:-1, 1:-1], ternary)
        cols = list(range(19 * nc)) + (list(range(60, 63)) if ternary else [])  # populations, db
        d = np.abs(got[..., cols] - want[..., cols]).max()
        print(f"{name} {plan}: max |population, dipole - reference| {d:.3e}")
        assert d <= 1e-11
    else:
        rho = g.density()
        for k, obs in enumerate(("od", "wd")):
            d = np.abs(rho[k] - ref[f"{name}/{obs}"]).max()
            print(f"{name} {plan}: max |rho_{k} - reference| {d:.3e}")
            assert d <= 1e-11
    m1 = ref[f"{name}/mass"][1]
    m = g.sanity_check()[0]
    assert np.all(np.abs(m[:nc] - m1[:nc]) <= 1e-13 * np.abs(m1[:nc]))
    g.close()
