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
  <case>/spec     the case (JSON): lattice, form, physics

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
BODY = dict(g_accn=(2e-4, -1e-4, 3e-4), g_accn_min=(2, 3, 2), g_accn_max=(6, 5, 4))


def case(dims, ternary, shear=None, kind="records", **phys):
    p = dict(g_br=0.08, bdist=2)
    if ternary:
        p.update(TERNARY_KW)
    p.update(phys)
    if "g_accn_max" not in p:
        p["g_accn_max"] = tuple(dims)  # the reference's 0: the lattice's extent
    return dict(dims=tuple(dims), ternary=ternary, shear=shear, kind=kind, phys=p)


CASES = {
    "binary_12x10x8": case((12, 10, 8), False),
    "ternary_10x8x12": case((10, 8, 12), True),
    "le_binary_16x8x8": case((16, 8, 8), False, shear=0.05),
    "le_binary_16x8x8_negative": case((16, 8, 8), False, shear=-0.07),
    "binary_32": case((32, 32, 32), False, kind="densities"),
    # the other equilibrium forms (code/lb3d_bdist.F90), unequal relaxation
    # times and masses, a body-force box.  Guo's term of bdist 4 and 5 is
    # (tau - 1/2) g_accn inside the box; outside it the reference leaves its
    # local dd as the last box site of the loop set it (or unset), so those
    # two forms take a box that covers the lattice.  (bdist 1 is left out:
    # the reference's own run of it multiplies the mass 4e7-fold in 100 steps
    # of this case; the oracle follows it, the GPU tests check it over 12.)
    "binary_bdist0_unequal": case((10, 8, 6), False, bdist=0, tau=(1.0, 0.8, 1.0), amass=(1.0, 1.2, 1.0)),
    **{f"binary_bdist{b}_body": case((8, 6, 5), False, bdist=b, tau=(0.9, 1.2, 1.0), **BODY) for b in (3, 6)},
    **{f"binary_bdist{b}_body": case((8, 6, 5), False, bdist=b, tau=(0.9, 1.2, 1.0), g_accn=BODY["g_accn"],
                                     g_accn_min=(1, 1, 1), g_accn_max=(8, 6, 5)) for b in (4, 5)},
    "ternary_body_unequal": case((10, 8, 6), True, tau=(0.9, 1.2, 1.1), amass=(1.0, 0.9, 1.2), tau_d=2.0, beta=3.0,
                                 d_0=0.9, g_accn=(1e-4, -2e-4, 3e-4), g_accn_min=(3, 2, 2), g_accn_max=(8, 6, 5)),
    # the Zou-He flux boundary on the z planes (inv_fluid 12): mass flow and
    # velocity forms
    "ternary_zou_he_mass_flow": case((8, 6, 10), True, inv_fluid=12, inv_type=-1, inv_flux=(1e-3, 0.0, 0.0)),
    "binary_zou_he_velocity": case((8, 6, 10), False, inv_fluid=12, inv_type=1, inv_flux=(0.004, -0.002, 0.0)),
}


def oracle_kw(c):
    kw = dict(c["phys"], ncomp=3 if c["ternary"] else 2, restore=1, shear_sum=0.0,
              lees_edwards=0 if c["shear"] is None else 1, shear_u=0.0 if c["shear"] is None else c["shear"])
    return kw


def initial_oracle(c):
    o = Oracle(*c["dims"], **oracle_kw(c))
    o.init_noise(SEED, rho_mean=(0.7, 0.7, 0.7), amp=(0.01, 0.01, 0.01 if c["ternary"] else 0.0),
                 dipoles=c["ternary"])
    return o


def main():
    for t in (False, True):
        ok, why = refrun.available(t)
        if not ok:
            raise SystemExit(why)
    out = {}
    for name, c in CASES.items():
        dims, ternary = c["dims"], c["ternary"]
        o = initial_oracle(c)
        start = refrun.records_from_sites(o.interior(), ternary)
        m0 = o.sanity()[0]
        kw = dict(c["phys"], lees_edwards=c["shear"] is not None, shear_u=0.0 if c["shear"] is None else c["shear"])
        _, rec = refrun.run(name.replace("_", ""), dims, STEPS, ternary, start=start, checkpoint_at=STEPS, **kw)
        ld = Oracle(*dims, **oracle_kw(c))
        ld.load_checkpoint_array(rec)
        m1 = ld.sanity()[0]
        out[f"{name}/mass"] = np.stack([m0, m1])
        if c["kind"] == "records":
            out[f"{name}/out"] = rec
        else:
            out[f"{name}/od"] = ld.dump("od")
            out[f"{name}/wd"] = ld.dump("wd")
        spec = dict(c, steps=STEPS, seed=SEED)
        out[f"{name}/spec"] = np.frombuffer(json.dumps(spec).encode(), dtype=np.uint8)
        print(name, dims, "stored", c["kind"], flush=True)
    np.savez_compressed(OUT, **out)
    print("wrote", OUT, os.path.getsize(OUT), "bytes")


if __name__ == "__main__":
    main()
