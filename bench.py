#!/usr/bin/env python3
"""bench.py -- MLUPS and HBM-roofline fraction of the LB3D D3Q19 timestep on
one MI355X: the FP32-batched ensemble against the plain FP64 path.

Headline workload (the batched ternary ensemble): 48 independent periodic
ternary lattices, cube edges drawn uniformly from {16, 24, ..., 96} (numpy
default_rng(1)), g_br uniform in [0.06, 0.12], g_bs = -0.006, g_ss = -0.003,
tau = tau_d = 1, beta = 10, d_0 = 1, bdist 2, densities 0.7 + U(-0.01, 0.01)
for red and blue, 0.7 surfactant, dipoles uniform on the sphere (noise seed
1000 + k), FP32 deviation storage, all lattices advanced together by the
tuner's plan with graph-replayed steps.  The timed region is `--steps` whole
steps with every lattice resident in HBM, bracketed by stream
synchronisation; `value` is its MLUPS (sum of lattice sites x steps / wall
time).  Beside it: the 48 lattices run one after another on the plain FP64
path, the step's rate against the 480 B/site roofline with HIP-event timing
on the batch's stream, the HBM traffic per step from rocprofv3 counters, the
per-phase kernel times, and the FP32 error against FP64.

The binary spinodal lattice (256^3, FP32 deviation and FP64 plain), the
ternary 128^3 and the Lees-Edwards 256x64x64 configurations are reported in
extra keys, each with its FP32-against-FP64 error, the FP64 floor (a 2^-51
perturbation) and an FP64 control perturbed at FP32 resolution (2^-24).  The
CPU baseline is the reference LB3D itself (built by oracle/ref_build.sh, run
under mpirun) timed on a bounded sample, with the oracle (the FP64 C
restatement of the reference loop) beside it; the oracle alone where the
reference build is absent.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "MLUPS + % HBM roofline, FP32-batched vs FP64-plain D3Q19, 1 MI355X"
HBM_PEAK_GBS = 8000.0


def bytes_per_site(ncomp, ternary, real_bytes):
    """Algorithmic bytes per site update: read and write every stored
    population once, plus the ternary dipole field (3 reals)."""
    return 2 * (19 * ncomp + (3 if ternary else 0)) * real_bytes


def parse_plan(chosen):
    """set_plan arguments from the tuner's 'chosen: ...' line."""
    w = chosen.split()
    kind, shape, lz, stage, xg = w[1], w[2], w[3], w[4], w[5]
    tx, ty = (int(v) for v in shape.split("x"))
    mode = 0 if kind == "two-pass" else {"lds": 1, "regs": 2, "reload": 3}[stage]
    return dict(mode=mode, tx=tx, ty=ty, lz=int(lz.split("=")[1]), xg=int(xg.split("=")[1]))


def velocity(rho, mom):
    """Barycentric velocity sum_c P_c / sum_c rho_c, shape (3, nz, ny, nx)."""
    return mom.sum(axis=0) / rho.sum(axis=0)[None]


def errs(a, b):
    """(Linf, L2) error of a against b, relative to max|b| and ||b||."""
    d = np.abs(a - b)
    ref = np.abs(b)
    return float(d.max() / ref.max()), float(np.sqrt((d ** 2).sum() / (ref ** 2).sum()))


def run_form(n, ncomp, precision, steps, warmup, profile_steps):
    """The binary spinodal lattice in one storage form: the tuner's plan,
    `steps` timed steps (wall clock and HIP events), per-kernel times."""
    from lb3d_amd.gpu import Lattice

    g = Lattice(n, n, n, ncomp=ncomp, precision=precision, g_br=0.08, bdist=2)
    g.init_noise(1, rho_mean=(0.7, 0.7, 0.7), amp=(0.01, 0.01, 0.0), dipoles=False)
    g.bc_init()
    report = g.autotune(trial_steps=3)
    chosen = report.strip().splitlines()[-1]
    print(f"[bench] autotune: {chosen}", file=sys.stderr, flush=True)
    g.time_loop(warmup)
    g.synchronize()
    t0 = time.perf_counter()
    ev_ms = g.time_loop_timed(steps)
    dt = time.perf_counter() - t0
    kern, ms_step_prof = g.profile_steps(profile_steps)
    g.close()
    return dict(seconds=dt, ms_per_step=1e3 * dt / steps, mlups=n ** 3 * steps / dt / 1e6,
                event_ms_per_step=ev_ms / steps, kernels=kern, plan=chosen, prof_ms_per_step=ms_step_prof)


# The FP64 comparison runs: the floor starts from a mean density 2^-51
# higher (uniform); the control from the noise of the initial state scaled so
# that every site's density moves by about 0.7 x 2^-24 (random per site, like
# FP32's rounding, which a uniform shift of that size does not mimic: it
# barely excites the unstable modes).
CONTROL_DRHO = 0.7 * 2.0 ** -24
PERTURB = (("floor", 2.0 ** -51, "initial mean density x (1 + 2^-51)"),
           ("control", CONTROL_DRHO, "initial noise amplitude a -> a + 0.7 x 2^-24 (a site's density moves by "
                                     "up to 0.7 x 2^-24, FP32 resolution)"))


def perturbed(tag, amp):
    """(mean density, noise amplitudes) of the run `tag`."""
    if tag == "floor":
        return 0.7 * (1.0 + PERTURB[0][1]), amp
    if tag == "control":
        return 0.7, tuple(a + CONTROL_DRHO if a > 0 else a for a in amp)
    return 0.7, amp


def first_over(traj, key, tol=1e-5):
    """First checkpoint step at which the density error `key` exceeds tol (None: never)."""
    return next((t["step"] for t in traj if t[key] > tol), None)


def config_accuracy(name, dims, steps, kw, amp, checkpoints):
    """FP32 deviation form against FP64 on one configuration, advanced
    together to each checkpoint (the last one is `steps`): the density error
    along the way, the FP64 floor (FP64 against FP64 from an initial mean
    density 2^-51 higher) and an FP64 control perturbed at FP32 resolution
    (2^-24), the velocity error at the end, and for Lees-Edwards the x-halo
    planes and the mean shear stress."""
    from lb3d_amd.gpu import LB3D_BC_LEES_EDWARDS, LB3D_FP32_DEV, LB3D_FP64, Lattice

    le = kw.get("boundary") == LB3D_BC_LEES_EDWARDS
    runs, secs = {}, {}
    for tag, prec in (("fp64", LB3D_FP64), ("fp32", LB3D_FP32_DEV), ("floor", LB3D_FP64), ("control", LB3D_FP64)):
        g = Lattice(*dims, precision=prec, **kw)
        m, a = perturbed(tag, amp)
        g.init_noise(1, rho_mean=(m, m, m), amp=a, dipoles=True)
        g.bc_init()
        g.autotune(trial_steps=2)
        runs[tag], secs[tag] = g, 0.0
    traj, done = [], 0
    for cp in list(checkpoints) + [steps]:
        if cp <= done or cp > steps:
            continue
        for tag, g in runs.items():
            t0 = time.perf_counter()
            g.time_loop(cp - done)
            g.synchronize()
            secs[tag] += time.perf_counter() - t0
        done = cp
        rho = {tag: g.density() for tag, g in runs.items()}
        traj.append(dict(step=cp, **{tag: errs(rho[tag], rho["fp64"])[0] for tag in ("fp32", "floor", "control")}))
    res = {}
    for tag, g in runs.items():
        rho = g.density()
        r = dict(rho=rho, u=velocity(rho, g.momentum()), mass=g.sanity_check()[0][:kw["ncomp"]])
        if le:
            r["halo"] = g.le_halo()
            r["stress"] = g.shear_stress()[2]
        res[tag] = r
        g.close()
    a, b = res["fp32"], res["fp64"]
    out = dict(steps=steps, lattice=list(dims),
               mlups={t: round(float(np.prod(dims)) * steps / secs[t] / 1e6, 1) for t in ("fp64", "fp32")})
    out["density_linf_rel"], out["density_l2_rel"] = errs(a["rho"], b["rho"])
    out["velocity_linf_rel"], out["velocity_l2_rel"] = errs(a["u"], b["u"])
    for tag, rel, what in PERTURB:
        f = res[tag]
        fl = errs(f["rho"], b["rho"])
        key = "fp64_perturbed_floor" if tag == "floor" else "fp64_control_fp32_resolution"
        out[key] = dict(density_linf_rel=fl[0], density_l2_rel=fl[1],
                        velocity_linf_rel=errs(f["u"], b["u"])[0], perturbation=what)
    # pointwise errors of an unstable flow grow from the rounding of each
    # step; the control run says how far a perturbation of FP32's size grows
    # on its own, the floor run how far one of FP64's size does
    out["density_linf_over_floor"] = out["density_linf_rel"] / max(out["fp64_perturbed_floor"]["density_linf_rel"],
                                                                   1e-300)
    out["density_linf_over_control"] = out["density_linf_rel"] / max(
        out["fp64_control_fp32_resolution"]["density_linf_rel"], 1e-300)
    out["trajectory_density_linf"] = traj
    out["first_step_density_linf_over_1e-5"] = {t: first_over(traj, t) for t in ("fp32", "control", "floor")}
    out["density_std_rel"] = float(abs(a["rho"][0].std() - b["rho"][0].std()) / b["rho"][0].std())
    out["mass_rel"] = float(np.max(np.abs(a["mass"] - b["mass"]) / np.abs(b["mass"])))
    if le:
        out["le_halo_linf_rel"] = errs(a["halo"], b["halo"])[0]
        out["mean_shear_stress"] = dict(fp64=b["stress"], fp32=a["stress"],
                                        rel_diff=abs(a["stress"] - b["stress"]) / abs(b["stress"]))
    print(f"[bench] {name}: density Linf {out['density_linf_rel']:.3e} (control "
          f"{out['fp64_control_fp32_resolution']['density_linf_rel']:.3e}, floor "
          f"{out['fp64_perturbed_floor']['density_linf_rel']:.3e}) after {steps} steps; trajectory {traj}",
          file=sys.stderr, flush=True)
    return out


def traffic_probe(n, plan):
    """Child run under rocprofv3 --pmc: a few steps (plan["steps"], default 4)
    of the FP32 lattice with the given plan, then a device copy of known size
    (the counter calibration)."""
    import torch
    from lb3d_amd.gpu import LB3D_FP32_DEV, Lattice

    g = Lattice(n, n, n, ncomp=2, precision=LB3D_FP32_DEV, g_br=0.08, bdist=2)
    g.init_noise(1, rho_mean=(0.7, 0.7, 0.7), amp=(0.01, 0.01, 0.0), dipoles=False)
    g.bc_init()
    g.set_plan(plan["mode"], plan["tx"], plan["ty"], max(plan["lz"], 1), xcd_group=plan["xg"])
    g.time_loop(plan.get("steps", 4))
    g.synchronize()
    g.close()
    a = torch.ones(COPY_BYTES // 4, dtype=torch.float32, device="cuda:0")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()


COPY_BYTES = 1 << 30


def ternary_probe(n):
    """Child run for rocprofv3 counter passes: 3 steps of the FP32 ternary
    n^3 lattice on the three-pass plan, then 3 on each of two one-pass plans
    (32x8 and 64x4 tiles, z-chunks of 16)."""
    from lb3d_amd.gpu import LB3D_FP32_DEV, Lattice

    g = Lattice(n, n, n, ncomp=3, precision=LB3D_FP32_DEV, g_br=0.08, g_bs=-0.006, g_ss=-0.003, tau_d=1.0,
                beta=10.0, d_0=1.0, bdist=2)
    g.init_noise(1, rho_mean=(0.7, 0.7, 0.7), amp=(0.01, 0.01, 0.0), dipoles=True)
    g.bc_init()
    for plan in ((0,), (3, 32, 8, 16, 0), (3, 64, 4, 16, 0)):
        g.set_plan(*plan)
        g.time_loop(3)
        g.synchronize()
    g.close()


FP64_PROBE_PLANS = ((0, 64, 4, 1, 64), (3, 256, 2, 128, 64), (2, 256, 2, 128, 64), (1, 32, 8, 64, 32))


def fp64_probe():
    """Child run for rocprofv3 passes: the binary 256^3 lattice in FP64, 3
    steps on each of the plans FP64_PROBE_PLANS (set_plan arguments: the
    two-pass plan, then the one-pass kernel with the streamed populations
    pulled again, in registers, in an LDS stage)."""
    import torch  # noqa: F401  (first: see copy_calibration)
    from lb3d_amd.gpu import LB3D_FP64, Lattice

    g = Lattice(256, 256, 256, ncomp=2, precision=LB3D_FP64, g_br=0.08, bdist=2)
    g.init_noise(1, rho_mean=(0.7, 0.7, 0.7), amp=(0.01, 0.01, 0.0), dipoles=False)
    g.bc_init()
    for plan in FP64_PROBE_PLANS:
        g.set_plan(*plan)
        g.time_loop(3)
        g.synchronize()
    g.close()
    copy_calibration()


LE_DIMS = (256, 64, 64)
LE_KW = dict(ncomp=2, g_br=0.08, bdist=2, shear_u=0.05)


def le_lattice(precision):
    from lb3d_amd.gpu import LB3D_BC_LEES_EDWARDS, Lattice

    g = Lattice(*LE_DIMS, precision=precision, boundary=LB3D_BC_LEES_EDWARDS, **LE_KW)
    g.init_noise(1, rho_mean=(0.7, 0.7, 0.7), amp=(0.01, 0.01, 0.0), dipoles=False)
    g.bc_init()
    return g


def le_probe(precision_name):
    """Child run for rocprofv3 passes: the Lees-Edwards 256x64x64 binary lattice
    (shear velocity 0.05, g_br 0.08) in the given form on the two-pass plan
    (which the tuner picks on this lattice), 10 steps, 5 profiled steps, then
    a device copy of known size (the counter calibration)."""
    import torch  # noqa: F401  (first: see copy_calibration)
    from lb3d_amd.gpu import LB3D_FP32_DEV, LB3D_FP64

    g = le_lattice(LB3D_FP32_DEV if precision_name == "fp32" else LB3D_FP64)
    g.set_plan(0)
    g.time_loop(10)
    g.synchronize()
    print(g.profile_steps(5), flush=True)
    g.close()
    copy_calibration()


def confirm_xcd_chunk(g, chosen, steps=50):
    """The tuner's trial steps all read the same populations, which on a
    lattice whose populations fit the memory-side cache stay there; real steps
    alternate the two buffers.  For a two-pass choice, keep the XCD chunk that
    is fastest over `steps` real steps (the state advances; the bench's timing
    runs on it afterwards)."""
    w = chosen.split()
    if len(w) < 2 or w[1] != "two-pass":
        return chosen
    xg0 = int([t for t in w if t.startswith("xg=")][0][3:])
    best = None
    for xg in sorted({xg0, 0, 4, 8, 16, 32, 64}):
        g.set_plan(0, xcd_group=xg)
        g.time_loop(4)
        ms = g.time_loop_timed(steps) / steps
        if best is None or ms < best[0]:
            best = (ms, xg)
    g.set_plan(0, xcd_group=best[1])
    out = f"chosen: two-pass 0x0 lz=0 - xg={best[1]} {best[0]:.4f} ms"
    print(f"[bench] XCD chunk confirmed over {steps} steps: {out}", file=sys.stderr, flush=True)
    return out


def run_le(steps):
    """The Lees-Edwards lattice in both forms on the tuner's plan: `steps`
    timed steps (HIP events), the per-kernel times, the step's fraction of
    the HBM peak at the form's algorithmic bytes per site."""
    from lb3d_amd.gpu import LB3D_FP32_DEV, LB3D_FP64

    out = {}
    ns = int(np.prod(LE_DIMS))
    for tag, prec, rb in (("fp32", LB3D_FP32_DEV, 4), ("fp64", LB3D_FP64, 8)):
        g = le_lattice(prec)
        rep = g.autotune(trial_steps=3).strip().splitlines()
        chosen = confirm_xcd_chunk(g, rep[-1])
        one_pass = [ln for ln in rep[:-1] if ln.startswith("fused")]
        g.time_loop(10)
        g.synchronize()
        ms = g.time_loop_timed(steps) / steps
        kern, _ = g.profile_steps(5)
        g.close()
        B = bytes_per_site(2, False, rb)
        out[tag] = dict(value=round(ns / (ms * 1e-3) / 1e6, 1), unit="MLUPS", ms_per_step=round(ms, 5),
                        bytes_per_site=B, step_frac=round(B * ns / (ms * 1e-3) / (HBM_PEAK_GBS * 1e9), 4),
                        plan=chosen, kernels_ms={k: round(v, 5) for k, v in kern.items()},
                        best_one_pass_trial=min(one_pass, key=lambda ln: float(ln.split()[-2])) if one_pass else None)
    out["workload"] = ("binary Shan-Chen D3Q19 256x64x64, Lees-Edwards planes at x (shear velocity 0.05), "
                       "g_br=0.08, tau=1, bdist=2, rho=0.7+-0.01 (seed 1)")
    out["steps"] = steps
    out["timing"] = "HIP events on the library stream around the timed steps / steps; kernels over 5 steps"
    return out


ENS_COMMON = dict(ncomp=3, g_bs=-0.006, g_ss=-0.003, tau_d=1.0, beta=10.0, d_0=1.0, bdist=2)
ENS_BYTES = 480  # bytes per site update: 3 x 19 FP32 populations and the 3-vector dipole, read and written


def ensemble_batch(precision, nlat=48, tag="fp64"):
    """The ensemble's lattices with their seeded initial state (perturbed as
    the comparison run `tag` is)."""
    from lb3d_amd.gpu import Batch

    edges, g_br = ensemble_spec(nlat)
    b = Batch([(e, e, e) for e in edges], per_lattice=[dict(g_br=g) for g in g_br], precision=precision,
              **ENS_COMMON)
    m, a = perturbed(tag, (0.01, 0.01, 0.0))
    for k in range(nlat):
        b.init_noise(k, 1000 + k, rho_mean=(m, m, m), amp=a, dipoles=True)
    return b, edges


def batch_plan(chosen):
    """The batch tuner's 'chosen: ...' line as arguments of apply_batch_plan."""
    w = chosen.split()
    kv = {a.split("=")[0]: int(a.split("=")[1]) for a in w if "=" in a}
    if w[1] == "persistent":
        return dict(kind="persistent", nwg=kv["nwg"], xg=kv["xg"], group_mib=int(chosen.split("(")[1].split()[0]))
    if w[1] == "collide-moments":
        return dict(kind="collide-moments", cols=kv["cols"], threads=kv["threads"], lz=kv["lz"], xg=kv["xg"])
    return dict(kind="sequential")


def apply_batch_plan(b, plan):
    if plan["kind"] == "persistent":
        b.set_plan(0, plan["nwg"], plan["xg"])
        b.set_groups(plan["group_mib"] << 20)
    elif plan["kind"] == "collide-moments":
        b.set_collide_moments(plan["cols"], plan["lz"], plan["xg"], tile_threads=plan["threads"])
    else:
        b.set_plan(1)


PROBE_STEPS = 4


def ensemble_probe(plan):
    """Child run under rocprofv3: PROBE_STEPS steps of the FP32 ensemble on
    the given plan, then a device copy of known size (the counter
    calibration)."""
    import torch  # noqa: F401  (first: see copy_calibration)
    from lb3d_amd.gpu import LB3D_FP32_DEV

    b, _ = ensemble_batch(LB3D_FP32_DEV)
    apply_batch_plan(b, plan)
    b.time_loop(PROBE_STEPS)
    b.synchronize()
    b.close()
    copy_calibration()


def copy_calibration():
    # torch must be imported before lb3d_amd initialises the HIP runtime (the
    # probes import it first); a late first import finds no GPU
    import torch

    a = torch.ones(COPY_BYTES // 4, dtype=torch.float32, device="cuda:0")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()


def counter_sums(outdir, counter):
    """Sum, count and largest value of `counter` over the dispatches of each
    kernel, from a rocprofv3 database."""
    import glob
    import sqlite3

    dbs = glob.glob(os.path.join(outdir, "**", "*.db"), recursive=True)
    if not dbs:
        return {}
    con = sqlite3.connect(dbs[0])
    acc = {}
    for name, val in con.execute("select kernel_name, value from counters_collection where counter_name=?",
                                 (counter,)):
        t = acc.setdefault(name, [0.0, 0, 0.0])
        t[0] += val
        t[1] += 1
        t[2] = max(t[2], val)
    con.close()
    return acc


def run_pmc(counter, probe_args, timeout=180):
    """One rocprofv3 --pmc pass of `counter` over `bench.py <probe_args>`."""
    import shutil
    import signal
    import subprocess
    import tempfile

    out = tempfile.mkdtemp(prefix="lb3d_pmc_")
    cmd = ["rocprofv3", "--pmc", counter, "-d", out, "-o", "pmc", "--", sys.executable,
           os.path.abspath(__file__)] + probe_args
    print(f"[bench] traffic: {' '.join(cmd)}", file=sys.stderr, flush=True)
    p = subprocess.Popen(cmd, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL, start_new_session=True)
    try:
        rc = p.wait(timeout=timeout)
    except subprocess.TimeoutExpired:
        os.killpg(p.pid, signal.SIGKILL)
        p.wait()
        shutil.rmtree(out, ignore_errors=True)
        return None, f"rocprofv3 --pmc {counter} timed out"
    sums = counter_sums(out, counter)
    shutil.rmtree(out, ignore_errors=True)
    if not sums:
        return None, f"rocprofv3 --pmc {counter}: rc {rc}, no counters"
    return sums, None


def measure_ensemble_traffic(plan):
    """HBM bytes per step of the ensemble on `plan`, per kernel and in total:
    the kernels of PROBE_STEPS steps under rocprofv3 --pmc FETCH_SIZE, then
    WRITE_SIZE (KiB per dispatch), each scaled by what the counter reads for a
    device copy of COPY_BYTES each way."""
    import shutil

    if not shutil.which("rocprofv3"):
        return None, "rocprofv3 not found"
    # the kernels of a step (the collide-moments plan's moments pass runs
    # only to prime the carried moments)
    step_kernels = (("k_batch_dipole", "k_batch_cm", "k_batch_edge_moments") if plan["kind"] == "collide-moments"
                    else ("k_batch_moments", "k_batch_dipole", "k_batch_collide", "k_moments", "k_dipole",
                          "k_collide"))
    got = {}
    for counter in ("FETCH_SIZE", "WRITE_SIZE"):
        sums, why = run_pmc(counter, ["--ensemble-probe", json.dumps(plan)])
        if sums is None:
            return None, why
        # the calibration copies are the largest copy dispatches (the batch's
        # table and work-list uploads are copies too)
        copy = [v[2] for k, v in sums.items() if "lb3d" not in k and "copy" in k.lower()]
        if not copy:
            return None, f"rocprofv3 --pmc {counter}: no copy kernel among {sorted(sums)[:6]}"
        scale = COPY_BYTES / (copy[0] * 1024.0)
        per = {}
        for k, (tot, n, _) in sums.items():
            short = k.split("(")[0].split("<")[0].replace("void lb3d::", "").strip()
            if short in step_kernels:
                per[short] = per.get(short, 0.0) + tot * 1024.0 * scale / PROBE_STEPS
        got[counter] = (per, scale)
    fetch, fs = got["FETCH_SIZE"]
    write, ws = got["WRITE_SIZE"]
    kernels = {k: dict(read=round(fetch.get(k, 0.0)), write=round(write.get(k, 0.0))) for k in set(fetch) | set(write)}
    total = sum(v["read"] + v["write"] for v in kernels.values())
    return dict(bytes_per_step=total, kernels=kernels, fetch_scale=round(fs, 4), write_scale=round(ws, 4)), None


def counter_means(outdir, counter):
    """Mean value per dispatch of `counter`, by kernel name, from a rocprofv3 database."""
    import glob
    import sqlite3

    dbs = glob.glob(os.path.join(outdir, "**", "*.db"), recursive=True)
    if not dbs:
        return {}
    con = sqlite3.connect(dbs[0])
    acc = {}
    for name, val in con.execute("select kernel_name, value from counters_collection where counter_name=?",
                                 (counter,)):
        acc.setdefault(name, []).append(val)
    con.close()
    return {k: sum(v) / len(v) for k, v in acc.items()}


def measure_traffic(n, plan, kernel_key, timeout=150):
    """HBM bytes per launch of the dominant kernel from two rocprofv3 --pmc
    runs (FETCH_SIZE, then WRITE_SIZE, in KiB per dispatch), each scaled by
    what the counter reads for a device copy of COPY_BYTES each way."""
    import shutil
    import signal
    import subprocess
    import tempfile

    if not shutil.which("rocprofv3"):
        return None, "rocprofv3 not found"
    got = {}
    for counter in ("FETCH_SIZE", "WRITE_SIZE"):
        out = tempfile.mkdtemp(prefix="lb3d_pmc_")
        cmd = ["rocprofv3", "--pmc", counter, "-d", out, "-o", "pmc", "--", sys.executable,
               os.path.abspath(__file__), "--size", str(n), "--traffic-probe", json.dumps(plan)]
        print(f"[bench] traffic: {' '.join(cmd)}", file=sys.stderr, flush=True)
        p = subprocess.Popen(cmd, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL, start_new_session=True)
        try:
            rc = p.wait(timeout=timeout)
        except subprocess.TimeoutExpired:
            os.killpg(p.pid, signal.SIGKILL)
            p.wait()
            return None, f"rocprofv3 --pmc {counter} timed out"
        vals = counter_means(out, counter)
        shutil.rmtree(out, ignore_errors=True)
        kern = [v for k, v in vals.items() if kernel_key in k]
        copy = [v for k, v in vals.items() if "lb3d" not in k and "copy" in k.lower()]
        if not kern or not copy:
            return None, f"rocprofv3 --pmc {counter}: rc {rc}, kernels {sorted(vals)[:6]}"
        got[counter] = (kern[0], copy[0])
    kf, cf = got["FETCH_SIZE"]
    kw, cw = got["WRITE_SIZE"]
    fetch_scale, write_scale = COPY_BYTES / (cf * 1024.0), COPY_BYTES / (cw * 1024.0)
    read, write = kf * 1024.0 * fetch_scale, kw * 1024.0 * write_scale
    return dict(bytes_per_launch=read + write, read=read, write=write, fetch_scale=round(fetch_scale, 4),
                write_scale=round(write_scale, 4), raw_kib=dict(fetch=kf, write=kw)), None


def hbm_ceiling(nbytes=2 << 30, reps=10):
    """The card's practical HBM rates from lb3d_amd/liblb3d_bwprobe.so: best
    of a few grid sizes of a 16-byte-per-lane copy (read + write bytes), a
    read-only and a write-only sweep over nbytes per buffer, in GB/s."""
    import ctypes

    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lb3d_amd", "liblb3d_bwprobe.so")
    lib = ctypes.CDLL(path)
    lib.lb3d_bw_probe.restype = ctypes.c_int
    lib.lb3d_bw_probe.argtypes = [ctypes.c_size_t, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]
    out = (ctypes.c_double * 3)()
    rc = lib.lb3d_bw_probe(nbytes, reps, out)
    if rc:
        raise RuntimeError(f"lb3d_bw_probe failed with HIP error {rc}")
    res = dict(copy=round(out[0], 1), read=round(out[1], 1), write=round(out[2], 1), unit="GB/s",
               probe=f"float4 grid-stride sweeps over {nbytes >> 20} MiB buffers, best of 4 grid sizes")
    lib.lb3d_bw_probe_planes.restype = ctypes.c_int
    lib.lb3d_bw_probe_planes.argtypes = [ctypes.c_size_t, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]
    for planes in (38, 57):
        v = ctypes.c_double()
        rc = lib.lb3d_bw_probe_planes(nbytes, planes, reps, ctypes.byref(v))
        if rc:
            raise RuntimeError(f"lb3d_bw_probe_planes failed with HIP error {rc}")
        res[f"plane_copy_{planes}"] = round(v.value, 1)
    res["plane_copy_probe"] = ("the population layout: one thread per site copies its 38 (binary FP32) or 57 "
                               "(ternary FP32) values between planes ns apart, read + write bytes")
    return res


def ensemble_spec(nlat=48, seed=1):
    """The batched ternary ensemble: cube edges uniform in {16, 24, ..., 96},
    g_br uniform in [0.06, 0.12], noise seed 1000+k (numpy default_rng(seed))."""
    rng = np.random.default_rng(seed)
    edges = 16 + 8 * rng.integers(0, 11, nlat)
    g_br = 0.06 + 0.06 * rng.random(nlat)
    return [int(e) for e in edges], [float(g) for g in g_br]


def group_count(pop_bytes, group_mib):
    """Groups of the batch's grouped plan (consecutive lattices, each group at
    most group_mib of populations unless one lattice is larger), as the
    library forms them."""
    n, acc, cap = 1, 0, group_mib << 20
    for by in pop_bytes:
        if acc > 0 and acc + by > cap:
            n, acc = n + 1, 0
        acc += by
    return n


def confirm_groups(b, chosen, pop_bytes, steps=30):
    """The tuner times its candidates on trial steps that all read the same
    state; the lattice groups' gain from the memory-side cache depends on the
    real alternation of the two population buffers.  Among the tuner's choice
    and XCD chunks of 16 and 64 with groups of 256, 288 and 320 MiB, keep the
    fastest over `steps` real steps (the state advances; the bench's timing
    runs on it afterwards)."""
    plan = batch_plan(chosen)
    if plan["kind"] != "persistent" or not plan["group_mib"]:
        return chosen
    best = None
    for xg in sorted({plan["xg"], 16, 64}):
        for mib in sorted({plan["group_mib"], 256, 288, 320}):
            b.set_plan(0, plan["nwg"], xg)
            b.set_groups(mib << 20)
            b.time_loop(4)
            ms = b.time_loop_timed(steps) / steps
            if best is None or ms < best[0]:
                best = (ms, xg, mib)
    ms, xg, mib = best
    b.set_plan(0, plan["nwg"], xg)
    b.set_groups(mib << 20)
    out = (f"chosen: persistent nwg={plan['nwg']} xg={xg} groups={group_count(pop_bytes, mib)} ({mib} MiB) "
           f"{ms:.4f} ms")
    print(f"[bench] ensemble groups confirmed over {steps} steps: {out}", file=sys.stderr, flush=True)
    return out


def run_ensemble(steps, warmup, seq_steps, nlat=48):
    """The FP32 batched ensemble on the tuner's plan (`steps` timed steps after
    `warmup`), against the lattices run one after another on the plain FP64
    path."""
    from lb3d_amd.gpu import LB3D_FP32_DEV, LB3D_FP64

    b, edges = ensemble_batch(LB3D_FP32_DEV, nlat)
    sites = sum(e ** 3 for e in edges)
    chosen = b.autotune(10).strip().splitlines()[-1]
    print(f"[bench] ensemble autotune: {chosen}", file=sys.stderr, flush=True)
    chosen = confirm_groups(b, chosen, [4 * 3 * 19 * e ** 3 for e in edges])
    b.time_loop(warmup)
    b.synchronize()
    t0 = time.perf_counter()
    ev_ms = b.time_loop_timed(steps)
    dt = time.perf_counter() - t0
    ph, mps = b.profile_steps(3)
    b.close()
    fp32 = dict(mlups=sites * steps / dt / 1e6, ms_per_step=1e3 * dt / steps, event_ms_per_step=ev_ms / steps,
                phases_ms={k: round(v, 4) for k, v in ph.items()}, profiled_ms_per_step=round(mps, 4), plan=chosen)
    print(f"[bench] ensemble fp32_batched: {fp32['mlups']:.1f} MLUPS, {fp32['event_ms_per_step']:.4f} ms per step "
          f"(HIP events) {ph}", file=sys.stderr, flush=True)
    q, _ = ensemble_batch(LB3D_FP64, nlat)
    q.set_plan(1)  # lattice after lattice, the single-lattice FP64 kernels
    q.time_loop(2)
    q.synchronize()
    t0 = time.perf_counter()
    q.time_loop(seq_steps)
    q.synchronize()
    dq = time.perf_counter() - t0
    q.close()
    fp64 = dict(value=round(sites * seq_steps / dq / 1e6, 1), unit="MLUPS", steps=seq_steps,
                ms_per_step=round(1e3 * dq / seq_steps, 4), plan="sequential (one lattice after another)")
    print(f"[bench] ensemble fp64_sequential: {fp64['value']} MLUPS", file=sys.stderr, flush=True)
    return dict(lattices=nlat, sites=sites, edges=edges, fp32=fp32, fp64_sequential=fp64)


def ensemble_accuracy(steps, checkpoints, nlat=48):
    """The FP32 batched ensemble against the FP64 batched ensemble from the
    same initial state, advanced together to each checkpoint: density errors
    of every lattice (the worst lattice reported) beside the FP64 floor
    (2^-51) and an FP64 control perturbed at FP32 resolution (2^-24)."""
    from lb3d_amd.gpu import LB3D_FP32_DEV, LB3D_FP64

    G = np.array([2.0] * 6 + [1.0] * 13)
    runs = {}
    for tag, prec in (("fp64", LB3D_FP64), ("fp32", LB3D_FP32_DEV), ("floor", LB3D_FP64), ("control", LB3D_FP64)):
        runs[tag], edges = ensemble_batch(prec, nlat, tag)

    def densities(b):
        # from the populations (the collision conserves them per site)
        return [np.einsum("i,cizyx->czyx", G, b.populations(k)) for k in range(nlat)]

    def worst(rho, ref):
        linf = [float(np.abs(a - r).max() / np.abs(r).max()) for a, r in zip(rho, ref)]
        return max(linf), linf

    traj, done, final = [], 0, {}
    for cp in list(checkpoints) + [steps]:
        if cp <= done or cp > steps:
            continue
        for b in runs.values():
            b.time_loop(cp - done)
            b.synchronize()
        done = cp
        rho = {tag: densities(b) for tag, b in runs.items()}
        traj.append(dict(step=cp, **{t: worst(rho[t], rho["fp64"])[0] for t in ("fp32", "floor", "control")}))
        final = rho
    for b in runs.values():
        b.close()
    out = {}
    for tag in ("fp32", "floor", "control"):
        a_, r_ = final[tag], final["fp64"]
        _, linf = worst(a_, r_)
        l2 = [float(np.linalg.norm(a - r) / np.linalg.norm(r)) for a, r in zip(a_, r_)]
        # statistics that survive the flow's amplification of rounding: the
        # spread of each density field and each component's mass
        spread = [float(abs(a[0].std() - r[0].std()) / r[0].std()) for a, r in zip(a_, r_)]
        mass = [float(np.max(np.abs(a.sum(axis=(1, 2, 3)) - r.sum(axis=(1, 2, 3))) / r.sum(axis=(1, 2, 3))))
                for a, r in zip(a_, r_)]
        out[tag] = dict(density_linf_rel=max(linf), density_l2_rel=max(l2),
                        worst_lattice_edge=edges[int(np.argmax(linf))],
                        lattices_within_1e5_linf=int(sum(v <= 1e-5 for v in linf)),
                        density_std_rel=max(spread), mass_rel=max(mass))
    res = dict(steps=steps, **out["fp32"])
    res["fp64_perturbed_floor"] = dict(out["floor"], perturbation=PERTURB[0][2])
    res["fp64_control_fp32_resolution"] = dict(out["control"], perturbation=PERTURB[1][2])
    res["density_linf_over_floor"] = res["density_linf_rel"] / max(out["floor"]["density_linf_rel"], 1e-300)
    res["density_linf_over_control"] = res["density_linf_rel"] / max(out["control"]["density_linf_rel"], 1e-300)
    res["trajectory_density_linf"] = traj
    res["first_step_density_linf_over_1e-5"] = {t: first_over(traj, t) for t in ("fp32", "control", "floor")}
    print(f"[bench] ensemble accuracy after {steps} steps: density Linf {res['density_linf_rel']:.3e} "
          f"(control {out['control']['density_linf_rel']:.3e}, floor {out['floor']['density_linf_rel']:.3e}); "
          f"trajectory {traj}", file=sys.stderr, flush=True)
    return res


def host_threads():
    """Threads for the multi-core CPU baseline: OMP_NUM_THREADS if set, else
    the CPUs this process may run on."""
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        return int(env)
    return len(os.sched_getaffinity(0))


def cpu_model():
    cpu = open("/proc/cpuinfo").read() if os.path.exists("/proc/cpuinfo") else ""
    return next((ln.split(":", 1)[1].strip() for ln in cpu.splitlines() if ln.startswith("model name")), "")


def cpu_baseline(sample_lattices, sample_steps, single_n, single_steps):
    """The CPU path on a bounded sample of the ensemble (its first
    `sample_lattices` lattices with edges up to 56, `sample_steps` steps
    each, on all host threads; then the lattice of edge single_n on one core):
    the reference LB3D itself (built by oracle/ref_build.sh, `mpirun -n
    <threads>`, the step-loop time of its own timer) when it is built here,
    and the oracle (the FP64 C restatement, OpenMP over sites) beside it."""
    from oracle import refrun
    from oracle.pyoracle import Oracle, set_threads

    edges, g_br = ensemble_spec(48)
    kw = dict(ENS_COMMON)
    th = host_threads()
    picks = [(k, e) for k, e in enumerate(edges) if e <= 56][:sample_lattices]
    k1 = next(k for k, e in enumerate(edges) if e == single_n) if single_n in edges else 0
    names = ", ".join(str(e) for _, e in picks)

    def port(picks, steps, threads):
        set_threads(threads)
        n_upd, t = 0, 0.0
        for k, e in picks:
            o = Oracle(e, e, e, g_br=g_br[k], **kw)
            o.init_noise(1000 + k, rho_mean=(0.7, 0.7, 0.7), amp=(0.01, 0.01, 0.0), dipoles=True)
            o.bc_init()
            t0 = time.perf_counter()
            o.run(steps)
            t += time.perf_counter() - t0
            n_upd += e ** 3 * steps
            o.close()
        return n_upd / t / 1e6, t

    def reference(picks, steps, ranks):
        n_upd, t = 0, 0.0
        rkw = {k: v for k, v in kw.items() if k in ("g_bs", "g_ss", "tau_d", "beta", "d_0", "bdist")}
        for k, e in picks:
            secs, _ = refrun.run(f"ens{k}", (e, e, e), steps, True, ranks=ranks, g_br=g_br[k], **rkw)
            t += secs
            n_upd += e ** 3 * steps
        return n_upd / t / 1e6, t

    v, dt = port(picks, sample_steps, th)
    v1, dt1 = port([(k1, single_n)], single_steps, 1)
    port_line = dict(value=round(v, 4), unit="MLUPS", cores=th, kind="port",
                     sample=f"oracle (FP64 C restatement of the reference loop, OpenMP over sites, {th} threads): "
                            f"the ensemble's first {len(picks)} lattices with edges <= 56 ({names}), "
                            f"{sample_steps} steps each, same physics and initial state as the workload",
                     seconds=round(dt, 3),
                     single_core=dict(value=round(v1, 4), unit="MLUPS", cores=1,
                                      sample=f"ternary {single_n}^3 lattice of the ensemble, {single_steps} steps",
                                      seconds=round(dt1, 3)))
    ok, why = refrun.available(True)
    if ok:
        try:
            rv, rdt = reference(picks, sample_steps, th)
            rv1, rdt1 = reference([(k1, single_n)], single_steps, 1)
        except Exception as e:  # the baseline is reported, never the product: say why it is missing
            ok, why = False, f"reference run failed: {e}"[:400]
    if not ok:
        port_line.update(cpu_model=cpu_model(), reference_unavailable=why)
        return port_line
    return dict(value=round(rv, 4), unit="MLUPS", cores=th, kind="reference",
                sample=f"the reference LB3D (Fortran + MPI, built from its sources with amdflang and MPICH), "
                       f"mpirun -n {th}, ternary build: the ensemble's first {len(picks)} lattices with edges "
                       f"<= 56 ({names}), {sample_steps} steps each, the same physics from the reference's "
                       "INIT_RAND state (mean density 0.7); time = its own step-loop timer "
                       "(code/lb3d_time_loop.F90:56-95)",
                seconds=round(rdt, 3), cpu_model=cpu_model(),
                single_core=dict(value=round(rv1, 4), unit="MLUPS", cores=1,
                                 sample=f"ternary {single_n}^3 lattice of the ensemble, {single_steps} steps, "
                                        "mpirun -n 1",
                                 seconds=round(rdt1, 3)),
                port=port_line)


PHASE_KERNELS = {"persistent": dict(moments="k_batch_moments", dipole="k_batch_dipole", collide="k_batch_collide"),
                 "collide-moments": dict(moments="k_batch_edge_moments", dipole="k_batch_dipole",
                                         collide="k_batch_cm"),
                 "sequential": dict(moments="k_moments", dipole="k_dipole", collide="k_collide")}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=300, help="timed steps of the ensemble (the headline)")
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--ensemble-fp64-steps", type=int, default=20)
    ap.add_argument("--size", type=int, default=256, help="edge of the binary spinodal lattice")
    ap.add_argument("--binary-steps", type=int, default=200)
    ap.add_argument("--profile-steps", type=int, default=10)
    ap.add_argument("--le-steps", type=int, default=500, help="timed steps of the Lees-Edwards lattice")
    ap.add_argument("--accuracy-steps", type=int, default=1000,
                    help="FP32 against FP64 after this many steps, binary lattice and ensemble (0: skip)")
    ap.add_argument("--no-configs", action="store_true",
                    help="skip the FP32-vs-FP64 accuracy runs of the ternary and Lees-Edwards configurations")
    ap.add_argument("--config-steps-scale", type=float, default=1.0)
    ap.add_argument("--cpu-sample-lattices", type=int, default=8)
    ap.add_argument("--cpu-sample-steps", type=int, default=30)
    ap.add_argument("--cpu-single-size", type=int, default=32)
    ap.add_argument("--cpu-single-steps", type=int, default=10)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-traffic", action="store_true",
                    help="skip the rocprofv3 --pmc child runs that measure HBM traffic")
    ap.add_argument("--traffic-probe", default=None, help=argparse.SUPPRESS)
    ap.add_argument("--ternary-probe", type=int, default=0, help=argparse.SUPPRESS)
    ap.add_argument("--ensemble-probe", default=None, help=argparse.SUPPRESS)
    ap.add_argument("--le-probe", default=None, help=argparse.SUPPRESS)
    ap.add_argument("--fp64-probe", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a device-to-device copy of MIB MiB (a kernel of known "
                         "traffic for scaling the FETCH_SIZE/WRITE_SIZE counters) and exit")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        n = args.copy_calibration * (1 << 20) // 4
        a = torch.ones(n, dtype=torch.float32, device="cuda:0")
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes_each_way": n * 4, "launches": 5}))
        return
    if args.traffic_probe:
        traffic_probe(args.size, json.loads(args.traffic_probe))
        return
    if args.ternary_probe:
        ternary_probe(args.ternary_probe)
        return
    if args.ensemble_probe is not None:
        ensemble_probe(json.loads(args.ensemble_probe))
        return
    if args.le_probe is not None:
        le_probe(args.le_probe)
        return
    if args.fp64_probe:
        fp64_probe()
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(the path is built and measured on a single MI355X)", file=sys.stderr)
        sys.exit(2)

    from lb3d_amd.gpu import LB3D_BC_LEES_EDWARDS, LB3D_FP32_DEV, LB3D_FP64

    # ---- the headline: the FP32-batched ternary ensemble -----------------
    print(f"[bench] ensemble: warmup {args.warmup}, steps {args.steps}", file=sys.stderr, flush=True)
    ens = run_ensemble(args.steps, args.warmup, args.ensemble_fp64_steps)
    r = ens["fp32"]
    sites, B = ens["sites"], ENS_BYTES
    plan = batch_plan(r["plan"])
    ev = r["event_ms_per_step"]
    achieved = B * sites / (ev * 1e-3) / 1e9
    names = PHASE_KERNELS[plan["kind"]]
    dom = max(r["phases_ms"], key=r["phases_ms"].get)
    roofline = dict(bound="hbm", achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                    frac=round(achieved / HBM_PEAK_GBS, 4), traffic=None,
                    kernel=(f"one graph-replayed step of the ensemble on the tuner's plan ({r['plan']}): "
                            f"{', '.join(names.values())}"),
                    algorithmic_bytes_per_launch=B * sites,
                    algorithmic_bytes_note=f"{B} B per site update (3 x 19 FP32 populations and the 3-vector "
                                           f"dipole field, read and written) x {sites} sites per step",
                    mean_launch_ms=round(ev, 5),
                    launch_timing="HIP events on the batch's stream around the timed steps / steps",
                    dominant_kernel=dict(name=names[dom], ms_per_step=r["phases_ms"][dom],
                                         timing="HIP events around each launch over 3 steps"))
    if not args.no_traffic:
        tr, why = measure_ensemble_traffic(plan)
        if tr:
            roofline["traffic"] = round(tr["bytes_per_step"])
            roofline["traffic_detail"] = tr
            dk = tr["kernels"].get(names[dom])
            if dk:
                roofline["dominant_kernel"]["traffic_bytes_per_step"] = dk["read"] + dk["write"]
                roofline["dominant_kernel"]["traffic_rate_gbs"] = round(
                    (dk["read"] + dk["write"]) / (r["phases_ms"][dom] * 1e-3) / 1e9, 1)
            print(f"[bench] ensemble traffic {tr['bytes_per_step'] / 1e9:.3f} GB per step "
                  f"(algorithmic {B * sites / 1e9:.3f}) {tr['kernels']}", file=sys.stderr, flush=True)
        else:
            roofline["traffic_note"] = why
            print(f"[bench] ensemble traffic not measured: {why}", file=sys.stderr, flush=True)
    print("[bench] HBM ceiling probe", file=sys.stderr, flush=True)
    ceiling = hbm_ceiling()
    ceiling["step_algorithmic_frac_of_copy"] = round(achieved / ceiling["copy"], 4)
    if roofline["traffic"]:
        rate = roofline["traffic"] / (ev * 1e-3) / 1e9
        ceiling["step_traffic_frac_of_copy"] = round(rate / ceiling["copy"], 4)
        ceiling["step_traffic_frac_of_plane_copy_57"] = round(rate / ceiling["plane_copy_57"], 4)

    # ---- the binary spinodal lattice: FP32 deviation and FP64 plain ------
    n, ncomp = args.size, 2
    print(f"[bench] binary {n}^3: FP64 plain and FP32 deviation, {args.binary_steps} steps", file=sys.stderr,
          flush=True)
    r64 = run_form(n, ncomp, LB3D_FP64, args.binary_steps, 10, args.profile_steps)
    r32 = run_form(n, ncomp, LB3D_FP32_DEV, args.binary_steps, 10, args.profile_steps)
    ns = n ** 3
    B32, B64 = bytes_per_site(ncomp, False, 4), bytes_per_site(ncomp, False, 8)
    kname = "fused" if "fused" in r32["kernels"] else "collide"
    if kname == "fused":
        # the one-pass plan launches exactly one kernel per periodic step:
        # its mean duration is the HIP-event time of the timed region / steps
        kc, how = r32["event_ms_per_step"], "HIP events on the library stream around the timed steps / steps"
    else:
        kc, how = r32["kernels"][kname], "HIP events on the library stream around each launch"
    bach = B32 * ns / (kc * 1e-3) / 1e9
    broof = dict(bound="hbm", achieved=round(bach, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                 frac=round(bach / HBM_PEAK_GBS, 4), traffic=None,
                 kernel=f"k_{'fused2' if kname == 'fused' else 'collide'} (FP32 deviation), plan {r32['plan']}",
                 algorithmic_bytes_per_launch=B32 * ns, mean_launch_ms=round(kc, 5), launch_timing=how)
    if not args.no_traffic:
        key = "k_fused2" if kname == "fused" else "k_collide"
        btr, why = measure_traffic(n, parse_plan(r32["plan"]), key)
        if btr:
            broof["traffic"] = round(btr["bytes_per_launch"])
            broof["traffic_detail"] = btr
        else:
            broof["traffic_note"] = why
    binary = dict(value=round(r32["mlups"], 1), unit="MLUPS", ms_per_step=round(r32["ms_per_step"], 4),
                  steps=args.binary_steps, lattice=[n, n, n], bytes_per_site=B32,
                  workload=f"binary Shan-Chen spinodal D3Q19 {n}^3 periodic, g_br=0.08, tau=1, bdist=2, "
                           "rho=0.7+-0.01 (seed 1), FP32 deviation storage",
                  roofline=broof, kernels_ms={k: round(v, 5) for k, v in r32["kernels"].items()},
                  fp64_plain=dict(value=round(r64["mlups"], 1), unit="MLUPS", ms_per_step=round(r64["ms_per_step"], 4),
                                  step_frac=round(r64["mlups"] * 1e6 * B64 / (HBM_PEAK_GBS * 1e9), 4),
                                  kernels_ms={k: round(v, 5) for k, v in r64["kernels"].items()}, plan=r64["plan"]),
                  speedup_fp32_vs_fp64=round(r32["mlups"] / r64["mlups"], 3))
    print(f"[bench] binary {n}^3: FP32 {binary['value']} MLUPS, FP64 {binary['fp64_plain']['value']} MLUPS",
          file=sys.stderr, flush=True)

    print("[bench] Lees-Edwards 256x64x64: FP32 deviation and FP64 plain", file=sys.stderr, flush=True)
    le = run_le(args.le_steps)
    print(f"[bench] Lees-Edwards: FP32 {le['fp32']['value']} MLUPS ({le['fp32']['step_frac']}), "
          f"FP64 {le['fp64']['value']} MLUPS ({le['fp64']['step_frac']})", file=sys.stderr, flush=True)

    # ---- FP32 against FP64 per configuration ------------------------------
    accuracy = {}
    if args.accuracy_steps > 0:
        st = args.accuracy_steps
        cps = [c for c in (10, 30, 100, 300, 1000, 2000, 5000) if c < st]
        accuracy["binary_256"] = config_accuracy(
            f"binary {n}^3", (n, n, n), st, dict(ncomp=2, g_br=0.08, bdist=2), (0.01, 0.01, 0.0), cps)
        accuracy["batched_ternary_ensemble"] = ensemble_accuracy(st, [c for c in (30, 100, 300) if c < st])
    if not args.no_configs:
        sc = args.config_steps_scale
        s2, s3 = max(1, int(2000 * sc)), max(1, int(5000 * sc))
        accuracy["ternary_128"] = config_accuracy(
            "ternary 128^3", (128, 128, 128), s2,
            dict(ncomp=3, g_br=0.08, g_bs=-0.006, g_ss=-0.003, tau_d=1.0, beta=10.0, d_0=1.0, bdist=2),
            (0.01, 0.01, 0.0), [c for c in (10, 30, 100, 300, 1000) if c < s2])
        accuracy["lees_edwards_256x64x64"] = config_accuracy(
            "Lees-Edwards 256x64x64", (256, 64, 64), s3,
            dict(ncomp=2, g_br=0.08, bdist=2, boundary=LB3D_BC_LEES_EDWARDS, shear_u=0.05),
            (0.01, 0.01, 0.0), [c for c in (10, 100, 300, 1000, 3000) if c < s3])
    cpu = None
    if not args.no_cpu_baseline:
        print(f"[bench] CPU baseline: the reference (where built) and the oracle on {args.cpu_sample_lattices} "
              f"ensemble lattices x "
              f"{args.cpu_sample_steps} steps (all threads), {args.cpu_single_size}^3 x {args.cpu_single_steps} "
              "(one core)", file=sys.stderr, flush=True)
        cpu = cpu_baseline(args.cpu_sample_lattices, args.cpu_sample_steps, args.cpu_single_size,
                           args.cpu_single_steps)
    fp64s = ens["fp64_sequential"]
    line = {
        "metric": METRIC,
        "value": round(r["mlups"], 1),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(r["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "batched ternary ensemble: 48 periodic D3Q19 lattices, cube edges uniform in "
                               "{16,...,96} (seed 1), g_br uniform in [0.06,0.12], g_bs=-0.006, g_ss=-0.003, "
                               "tau=tau_d=1, beta=10, bdist=2, rho=0.7+-0.01, dipoles on the sphere (seeds 1000+k), "
                               "FP32 deviation storage, one batch, graph-replayed steps, tuned plan",
                   "lattices": ens["lattices"], "sites": sites, "edges": ens["edges"], "components": 3,
                   "bytes_per_site": B, "plan": r["plan"]},
        "roofline": roofline,
        "hbm_ceiling": ceiling,
        "cpu_baseline": cpu,
        "fp64_plain": dict(fp64s, workload="the same 48 lattices run one after another on the plain FP64 path"),
        "speedup_fp32_batched_vs_fp64_plain": round(r["mlups"] / fp64s["value"], 3),
        "phases_ms": r["phases_ms"],
        "phases_note": ("HIP events around each launch over 3 steps launched one by one after the timed loop "
                        "(not graph-replayed); roofline.mean_launch_ms is the timed loop's per-step time"),
        "binary_256": binary,
        "lees_edwards_256x64x64": le,
        "accuracy_fp32_vs_fp64": accuracy,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
