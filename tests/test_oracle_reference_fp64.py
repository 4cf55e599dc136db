"""The oracle against the reference LB3D itself, in FP64.

tests/golden/reference_fp64.npz holds the reference's state after 100
timesteps from the oracle's seeded initial states (written by
tests/golden/make_ref_fp64.py, which runs the reference built from its own
sources by oracle/ref_build.sh).  The oracle, from the same initial state,
must match it to 1e-11 in every population, dipole and density (it matches
bit for bit where the fixtures were made), and both must conserve each
component's mass to 1e-13 relative.
"""
import importlib.util
import json
import os

import numpy as np
import pytest

from oracle import refrun

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
FIX = os.path.join(GOLDEN, "reference_fp64.npz")
_spec = importlib.util.spec_from_file_location("make_ref_fp64", os.path.join(GOLDEN, "make_ref_fp64.py"))
make_ref_fp64 = importlib.util.module_from_spec(_spec)
_spec.loader.exec_module(make_ref_fp64)
CASES, initial_oracle = make_ref_fp64.CASES, make_ref_fp64.initial_oracle


@pytest.fixture(scope="module")
def ref():
    return np.load(FIX)


def spec(ref, name):
    return json.loads(ref[f"{name}/spec"].tobytes().decode())


@pytest.mark.parametrize("name", list(CASES))
def test_oracle_matches_reference_fp64(ref, name):
    s = spec(ref, name)
    dims, ternary, shear = tuple(s["dims"]), s["ternary"], s["shear"]
    o = initial_oracle(dims, ternary, shear)
    o.bc_init()
    o.run(s["steps"])
    assert o.nt == s["steps"]
    if f"{name}/out" in ref.files:
        want = ref[f"{name}/out"]
        got = refrun.records_from_sites(o.interior(), ternary)
        d = np.abs(got - want)
        print(f"{name}: max |population - reference| {d[..., :57 if ternary else 38].max():.3e}, "
              f"all fields {d.max():.3e}")
        assert d.max() <= 1e-11
    else:
        for obs in ("od", "wd"):
            d = np.abs(o.dump(obs) - ref[f"{name}/{obs}"]).max()
            print(f"{name}: max |{obs} - reference| {d:.3e}")
            assert d <= 1e-11
    m0, m1 = ref[f"{name}/mass"]
    nc = 3 if ternary else 2
    assert np.all(np.abs(m1[:nc] - m0[:nc]) <= 1e-13 * np.abs(m0[:nc])), "reference mass"
    mo = o.sanity()[0]
    assert np.all(np.abs(mo[:nc] - m0[:nc]) <= 1e-13 * np.abs(m0[:nc])), "oracle mass"


def test_fixture_cases_differ_from_the_initial_state(ref):
    """100 steps move every case well away from its initial state, and the
    fixtures of the two shear signs differ: a fixture equal to its input
    would pin nothing."""
    for name, (dims, ternary, shear, kind) in CASES.items():
        if kind != "records":
            continue
        o = initial_oracle(dims, ternary, shear)
        start = refrun.records_from_sites(o.interior(), ternary)
        assert np.abs(ref[f"{name}/out"] - start).max() > 1e-4
    a, b = ref["le_binary_16x8x8/out"], ref["le_binary_16x8x8_negative/out"]
    assert np.abs(a - b).max() > 1e-4
