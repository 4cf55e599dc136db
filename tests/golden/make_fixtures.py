"""Write tests/golden/reference_goldens.npz from the reference's own golden tests.

Reads, as data only, the checkpoint inputs and the float32 XDR dumps that the
reference's regression tests hold (test/{gyroid,leesEdwards,bodyForce,wettingWall,zouHe}/,
run by testLB3D.sh), and stores them unchanged in one compressed npz so that
tests run where the reference tree is absent.  Formats:
  checkpoint_*_p000000.xdr  big-endian FP64 records, x outer .. z inner,
                            per site n_r[19] n_b[19] [n_s[19] da[3] db[3]] rock
                            (code/lb3d_io_checkpoint.F90:790-873)
  checktopo_*.xdr           cdims[3] i32, ccoords[3] i32, shear_sum f64
  <obs>_*.reference-t00000010.xdr  big-endian float32, z outer .. x inner,
                            vectors interleaved (code/lb3d_io_xdrf.F90:717-912)
Usage: python tests/golden/make_fixtures.py [/path/to/reference]
"""
import os
import sys

import numpy as np

TESTS = {
    "gyroid": (64, ["od", "wd", "sur", "dir", "vel", "arr"]),
    "leesEdwards": (39, ["od", "wd", "vel", "arr"]),
    "bodyForce": (64, ["od", "wd", "sur", "dir", "vel", "arr"]),
    "wettingWall": (39, ["od", "wd", "vel", "arr"]),
    "zouHe": (64, ["od", "wd", "sur", "dir", "vel", "arr"]),
}


def main(ref):
    out = {}
    for name, (rec, obs) in TESTS.items():
        base = os.path.join(ref, "test", name)
        ck = np.fromfile(os.path.join(base, "checkpoint",
                                      f"checkpoint_{name}_t00000000-0000000000_p000000.xdr"),
                         dtype=">f8")
        out[f"{name}/checkpoint"] = ck.reshape(16, 16, 16, rec).astype("<f8")
        topo = open(os.path.join(base, "checkpoint",
                                 f"checktopo_{name}_t00000000-0000000000.xdr"), "rb").read()
        out[f"{name}/shear_sum"] = np.frombuffer(topo[24:32], dtype=">f8").astype("<f8")
        for o in obs:
            d = np.fromfile(os.path.join(base, "reference", f"{o}_{name}.reference-t00000010.xdr"),
                            dtype=">f4").astype("<f4")
            out[f"{name}/{o}"] = d.reshape(16, 16, 16, 3) if o in ("dir", "arr") else d.reshape(16, 16, 16)
    # the obstacle file of test/wettingWall: one big-endian float32 per site,
    # z outer .. x inner (read_rock_xdrf_par, code/lb3d_io_xdrf.F90:400-470)
    wall = np.fromfile(os.path.join(ref, "test", "wettingWall", "wettingWall.wall.xdr"), dtype=">f4")
    out["wettingWall/wall"] = wall.astype("<f4").reshape(16, 16, 16)
    here = os.path.dirname(os.path.abspath(__file__))
    # the namelist input files of the same tests (data, read by the Fortran host)
    os.makedirs(os.path.join(here, "inputs"), exist_ok=True)
    for name in TESTS:
        src = os.path.join(ref, "test", name, f"{name}.in")
        with open(src) as fi, open(os.path.join(here, "inputs", f"{name}.in"), "w") as fo:
            fo.write(fi.read())
    dst = os.path.join(here, "reference_goldens.npz")
    np.savez_compressed(dst, **out)
    print("wrote", dst, os.path.getsize(dst), "bytes")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
