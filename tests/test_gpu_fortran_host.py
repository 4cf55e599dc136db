"""The Fortran host (ISO_C_BINDING over the C ABI) runs the reference's own
regression tests: the namelist input of test/<name>/<name>.in unchanged, the
test's XDR checkpoint restored from disk, 10 steps on the GPU, and the XDR
float32 dumps it writes compared byte for byte with the reference's goldens
(testLB3D.sh compares them with diff)."""
import os
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "lb3d_amd", "fortran", "bin", "lb3d")
GOLD = os.path.join(ROOT, "tests", "golden", "reference_goldens.npz")
INPUTS = os.path.join(ROOT, "tests", "golden", "inputs")

pytestmark = pytest.mark.gpu

CASES = {"gyroid": (3, ["od", "wd", "sur", "dir", "vel", "arr"]),
         "leesEdwards": (2, ["od", "wd", "vel", "arr"]),
         "bodyForce": (3, ["od", "wd", "sur", "dir", "vel", "arr"]),
         "wettingWall": (2, ["od", "wd", "vel", "arr"]),
         "zouHe": (3, ["od", "wd", "sur", "dir", "vel", "arr"])}


def write_checkpoint(dirname, name, gold):
    os.makedirs(dirname, exist_ok=True)
    ck = gold[f"{name}/checkpoint"].astype(">f8")
    ck.tofile(os.path.join(dirname, f"checkpoint_{name}_t00000000-0000000000_p000000.xdr"))
    topo = np.array([1, 1, 1, 0, 0, 0], dtype=">i4").tobytes() + gold[f"{name}/shear_sum"].astype(">f8").tobytes()
    with open(os.path.join(dirname, f"checktopo_{name}_t00000000-0000000000.xdr"), "wb") as f:
        f.write(topo)
    if f"{name}/wall" in gold:
        # the obstacle file next to the input (obs_folder = '.')
        gold[f"{name}/wall"].astype(">f4").tofile(os.path.join(os.path.dirname(dirname), f"{name}.wall.xdr"))


@pytest.mark.parametrize("name", list(CASES))
def test_fortran_host_runs_reference_test(tmp_path, name):
    assert os.path.exists(EXE), "build the Fortran host first (lb3d_amd/build.py)"
    gold = np.load(GOLD)
    ncomp, obs = CASES[name]
    write_checkpoint(str(tmp_path / "checkpoint"), name, gold)
    shutil.copy(os.path.join(INPUTS, f"{name}.in"), tmp_path / f"{name}.in")
    # the reference's dump_xdr.din switch: dump_format = 'xdr'
    r = subprocess.run([EXE, "-f", f"{name}.in", "-c", str(ncomp)], cwd=tmp_path,
                       capture_output=True, text=True, timeout=120)
    print(r.stdout[-3000:], r.stderr[-2000:])
    assert r.returncode == 0
    for ob in obs:
        path = tmp_path / "output" / f"{ob}_{name}_t00000010-0000000000.xdr"
        got = np.fromfile(path, dtype=">f4")
        want = gold[f"{name}/{ob}"].astype(">f4").reshape(-1)
        assert got.shape == want.shape, ob
        np.testing.assert_array_equal(got, want, err_msg=f"{name}/{ob}")


def test_fortran_host_staged_loop_equals_graph_loop(tmp_path):
    """fast_loop = .false. runs the reference's stage calls one by one; the
    dumps must be identical to the graph-replayed loop."""
    gold = np.load(GOLD)
    name = "gyroid"
    outs = {}
    for fast in (".true.", ".false."):
        d = tmp_path / fast.strip(".")
        write_checkpoint(str(d / "checkpoint"), name, gold)
        text = open(os.path.join(INPUTS, f"{name}.in")).read()
        text += f"\n&gpu_input\nfast_loop = {fast}\n/\n"
        (d / f"{name}.in").write_text(text)
        r = subprocess.run([EXE, "-f", f"{name}.in", "-c", "3"], cwd=d, capture_output=True, text=True,
                           timeout=120)
        assert r.returncode == 0, r.stdout + r.stderr
        outs[fast] = np.fromfile(d / "output" / f"od_{name}_t00000010-0000000000.xdr", dtype=">f4")
    np.testing.assert_array_equal(outs[".true."], outs[".false."])


@pytest.mark.parametrize("name,ncomp,precision", [("gyroid", 3, "fp64"), ("leesEdwards", 2, "fp64"),
                                                  ("leesEdwards", 2, "fp32")])
def test_fortran_host_restart_equivalence(tmp_path, name, ncomp, precision):
    """10 steps straight against 5 steps, the host's own XDR checkpoint (with
    checktopo and shear_sum), a restore and 5 more steps: identical dumps.
    The FP32 form keeps its deviation base fixed (&gpu_input rho0) so that
    the restored deviations are the same numbers."""
    gold = np.load(GOLD)
    base = open(os.path.join(INPUTS, f"{name}.in")).read()
    extra = f"\n&gpu_input\nprecision = '{precision}'\nrho0 = 0.35, 0.35, 0.35\n/\n"

    def run(d, text):
        d.mkdir(parents=True, exist_ok=True)
        write_checkpoint(str(d / "checkpoint"), name, gold)
        (d / f"{name}.in").write_text(text + extra)
        r = subprocess.run([EXE, "-f", f"{name}.in", "-c", str(ncomp)], cwd=d, capture_output=True,
                           text=True, timeout=120)
        assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
        return r

    straight = tmp_path / "straight"
    run(straight, base)
    first = tmp_path / "first"
    run(first, base.replace("n_checkpoint = 10000", "n_checkpoint = 5"))
    ck = first / "output" / f"checkpoint_{name}_t00000005-0000000000_p000000.xdr"
    assert ck.exists() and (first / "output" / f"checktopo_{name}_t00000005-0000000000.xdr").exists()
    second = tmp_path / "second"
    text = base.replace("srccpfolder = './checkpoint'", f"srccpfolder = '{first / 'output'}'")
    text = text.replace("restore_string = 't00000000-0000000000'", "restore_string = 't00000005-0000000000'")
    r = run(second, text)
    assert "t00000005" in r.stdout
    for ob in CASES[name][1]:
        fn = f"{ob}_{name}_t00000010-0000000000.xdr"
        np.testing.assert_array_equal(np.fromfile(second / "output" / fn, dtype=">f4"),
                                      np.fromfile(straight / "output" / fn, dtype=">f4"), err_msg=ob)
