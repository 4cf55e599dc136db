"""Write tests/golden/reference_fp64.npz: FP64 states of the reference LB3D
after 100 timesteps from the same seeded initial states the oracle and the GPU
path build.

The reference is built by oracle/ref_build.sh (amdflang + MPICH) into
oracle/_ref/ and run with `mpirun -n 1`.  Each case's initial state is the
oracle's seeded noise state (oracle.pyoracle.Oracle.init_noise, seed 7, mean
density 0.7, amplitudes 0.01), handed to the reference as an 'ori'
checkpoint (init_cond 7, restore at t = 0, shear_sum 0); the reference runs
100 steps and writes an 'ori' checkpoint, which is stored here as data:

  <case>/out      records [x, y, z, :] (n_r[19] n_b[19] [n_s[19] da[3] db[3]] rock)
                  for the small lattices
  <case>/od, /wd  red and blue densities [z, y, x] for the 32^3 lattice
  <case>/mass     per-component masses of the initial and the final state

Usage: python tests/golden/make_ref_fp64.py   (after bash oracle/ref_build.sh)
"""
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import refrun  # noqa: E402
from oracle.pyoracle import Oracle  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_fp64.npz")
STEPS = 100
SEED = 7
TERNARY_KW = dict(g_bs=-0.006, g_ss=-0.003, tau_d=1.0, beta=10.0, d_0=1.0)
# name: (dims, ternary, Lees-Edwards shear velocity or None, what is stored)
CASES = {
    "binary_12x10x8": ((12, 10, 8), False, None, "records"),
    "ternary_10x8x12": ((10, 8, 12), True, None, "records"),
    "le_binary_16x8x8": ((16, 8, 8), False, 0.05, "records"),
    "le_binary_16x8x8_negative": ((16, 8, 8), False, -0.07, "records"),
    "binary_32": ((32, 32, 32), False, None, "densities"),
}


def oracle_kw(ternary, shear):
    kw = dict(ncomp=3 if ternary else 2, g_br=0.08, bdist=2, lees_edwards=0 if shear is None else 1,
              shear_u=0.0 if shear is None else shear, restore=1, shear_sum=0.0)
    if ternary:
        kw.update(TERNARY_KW)
    return kw


def initial_oracle(dims, ternary, shear):
    o = Oracle(*dims, **oracle_kw(ternary, shear))
    o.init_noise(SEED, rho_mean=(0.7, 0.7, 0.7), amp=(0.01, 0.01, 0.01 if ternary else 0.0), dipoles=ternary)
    return o


def main():
    for t in (False, True):
        ok, why = refrun.available(t)
        if not ok:
            raise SystemExit(why)
    out = {}
    for name, (dims, ternary, shear, kind) in CASES.items():
        o = initial_oracle(dims, ternary, shear)
        start = refrun.records_from_sites(o.interior(), ternary)
        m0 = o.sanity()[0]
        kw = dict(lees_edwards=shear is not None, shear_u=0.0 if shear is None else shear)
        if ternary:
            kw.update(TERNARY_KW)
        _, rec = refrun.run(name.replace("_", ""), dims, STEPS, ternary, start=start, checkpoint_at=STEPS, **kw)
        ld = Oracle(*dims, **oracle_kw(ternary, shear))
        ld.load_checkpoint_array(rec)
        m1 = ld.sanity()[0]
        out[f"{name}/mass"] = np.stack([m0, m1])
        if kind == "records":
            out[f"{name}/out"] = rec
        else:
            out[f"{name}/od"] = ld.dump("od")
            out[f"{name}/wd"] = ld.dump("wd")
        out[f"{name}/spec"] = np.frombuffer(json.dumps(dict(dims=dims, ternary=ternary, shear=shear, steps=STEPS,
                                                            seed=SEED)).encode(), dtype=np.uint8)
        print(name, dims, "stored", kind, flush=True)
    np.savez_compressed(OUT, **out)
    print("wrote", OUT, os.path.getsize(OUT), "bytes")


if __name__ == "__main__":
    main()
