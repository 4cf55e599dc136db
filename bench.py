#!/usr/bin/env python3
"""bench.py -- MLUPS and HBM-roofline fraction of the LB3D D3Q19 timestep on
one MI355X, FP32 deviation storage against the plain FP64 path.

Workload (the spinodal-decomposition configuration): binary immiscible
Shan-Chen, D3Q19 BGK, 256^3 periodic, tau = 1, g_br = 0.08, bdist 2, initial
densities 0.7 + U(-0.01, 0.01) per component and site (seed 1), zero velocity.
Both forms start from the same device-generated state; the timed region is
`--steps` whole timesteps (graph replay) with the lattice resident in HBM,
bracketed by stream synchronisation.  `value` is the FP32-deviation MLUPS;
the FP64-plain MLUPS, the speedup and the density error between the two are
reported beside it.  Per-kernel durations for the roofline come from HIP
events on the library's own stream.  The CPU baseline is the oracle (the FP64
restatement of the reference loop, one core) timed on a bounded sample.

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


def run_form(n, ncomp, precision, steps, warmup, profile_steps, rho_shift=0.0):
    from lb3d_amd.gpu import Lattice

    g = Lattice(n, n, n, ncomp=ncomp, precision=precision, g_br=0.08, bdist=2)
    m = 0.7 + rho_shift
    g.init_noise(1, rho_mean=(m, m, m), amp=(0.01, 0.01, 0.0), dipoles=False)
    g.bc_init()
    report = g.autotune(trial_steps=3)
    chosen = report.strip().splitlines()[-1]
    print(f"[bench] autotune: {chosen}", file=sys.stderr, flush=True)
    g.time_loop(warmup)
    g.synchronize()
    t0 = time.perf_counter()
    ev_ms = g.time_loop_timed(steps)
    dt = time.perf_counter() - t0
    rho = g.density()
    u = velocity(rho, g.momentum())
    kern, ms_step_prof = g.profile_steps(profile_steps)
    m, _, counts = g.sanity_check()
    g.close()
    return dict(seconds=dt, ms_per_step=1e3 * dt / steps, mlups=n ** 3 * steps / dt / 1e6,
                event_ms_per_step=ev_ms / steps, kernels=kern, plan=chosen, prof_ms_per_step=ms_step_prof,
                rho=rho, u=u, mass=m.tolist(), counts=counts.tolist())


def config_accuracy(name, dims, steps, kw, amp):
    """FP32 deviation form against FP64 on one configuration after `steps`
    steps: density and velocity errors, the FP64 floor (FP64 against FP64 from
    an initial mean density 2^-51 higher), and for Lees-Edwards the x-halo
    planes and the mean shear stress."""
    from lb3d_amd.gpu import LB3D_BC_LEES_EDWARDS, LB3D_FP32_DEV, LB3D_FP64, Lattice

    le = kw.get("boundary") == LB3D_BC_LEES_EDWARDS
    res = {}
    for tag, prec, shift in (("fp64", LB3D_FP64, 0.0), ("fp32", LB3D_FP32_DEV, 0.0),
                             ("floor", LB3D_FP64, 2.0 ** -51)):
        g = Lattice(*dims, precision=prec, **kw)
        m = 0.7 + shift
        g.init_noise(1, rho_mean=(m, m, m), amp=amp, dipoles=True)
        g.bc_init()
        g.autotune(trial_steps=2)
        t0 = time.perf_counter()
        g.time_loop(steps)
        g.synchronize()
        dt = time.perf_counter() - t0
        rho = g.density()
        r = dict(rho=rho, u=velocity(rho, g.momentum()), mlups=float(np.prod(dims)) * steps / dt / 1e6,
                 mass=g.sanity_check()[0][:kw["ncomp"]])
        if le:
            r["halo"] = g.le_halo()
            r["stress"] = g.shear_stress()[2]
        res[tag] = r
        g.close()
    a, b, f = res["fp32"], res["fp64"], res["floor"]
    out = dict(steps=steps, lattice=list(dims),
               mlups=dict(fp64=round(b["mlups"], 1), fp32=round(a["mlups"], 1)))
    out["density_linf_rel"], out["density_l2_rel"] = errs(a["rho"], b["rho"])
    out["velocity_linf_rel"], out["velocity_l2_rel"] = errs(a["u"], b["u"])
    fl = errs(f["rho"], b["rho"])
    out["fp64_perturbed_floor"] = dict(density_linf_rel=fl[0], density_l2_rel=fl[1],
                                       velocity_linf_rel=errs(f["u"], b["u"])[0])
    # pointwise errors of an unstable flow grow from the rounding of each
    # step; the ratio to the floor run (a 2^-51 perturbation) says how much of
    # that growth is the flow's own, and the spread of the density field is a
    # statistic that survives it
    out["density_linf_over_floor"] = out["density_linf_rel"] / max(fl[0], 1e-300)
    out["density_std_rel"] = float(abs(a["rho"][0].std() - b["rho"][0].std()) / b["rho"][0].std())
    out["mass_rel"] = float(np.max(np.abs(a["mass"] - b["mass"]) / np.abs(b["mass"])))
    if le:
        out["le_halo_linf_rel"] = errs(a["halo"], b["halo"])[0]
        out["mean_shear_stress"] = dict(fp64=b["stress"], fp32=a["stress"],
                                        rel_diff=abs(a["stress"] - b["stress"]) / abs(b["stress"]))
    print(f"[bench] {name}: density Linf {out['density_linf_rel']:.3e} (floor {fl[0]:.3e})",
          file=sys.stderr, flush=True)
    return out


def traffic_probe(n, plan):
    """Child run under rocprofv3 --pmc: a few steps of the FP32 lattice with the
    given plan, then a device copy of known size (the counter calibration)."""
    import torch
    from lb3d_amd.gpu import LB3D_FP32_DEV, Lattice

    g = Lattice(n, n, n, ncomp=2, precision=LB3D_FP32_DEV, g_br=0.08, bdist=2)
    g.init_noise(1, rho_mean=(0.7, 0.7, 0.7), amp=(0.01, 0.01, 0.0), dipoles=False)
    g.bc_init()
    g.set_plan(plan["mode"], plan["tx"], plan["ty"], max(plan["lz"], 1), xcd_group=plan["xg"])
    g.time_loop(4)
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


def ensemble_probe(groups_mib):
    """Child run for rocprofv3 passes: the 48-lattice FP32 ternary ensemble on
    the persistent plan (one work item per workgroup, XCD chunks of 16 items)
    in groups of groups_mib MiB of populations (0: one group), 3 steps."""
    from lb3d_amd.gpu import LB3D_FP32_DEV, Batch

    edges, g_br = ensemble_spec(48)
    b = Batch([(e, e, e) for e in edges], per_lattice=[dict(g_br=g) for g in g_br], precision=LB3D_FP32_DEV,
              ncomp=3, g_bs=-0.006, g_ss=-0.003, tau_d=1.0, beta=10.0, d_0=1.0, bdist=2)
    for k in range(48):
        b.init_noise(k, 1000 + k, rho_mean=(0.7, 0.7, 0.7), amp=(0.01, 0.01, 0.0), dipoles=True)
    b.set_plan(0, nwg=1 << 30, xcd_group=16)
    b.set_groups(groups_mib << 20)
    b.time_loop(3)
    b.synchronize()
    b.close()


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
    return dict(copy=round(out[0], 1), read=round(out[1], 1), write=round(out[2], 1), unit="GB/s",
                probe=f"float4 grid-stride sweeps over {nbytes >> 20} MiB buffers, best of 4 grid sizes")


def ensemble_spec(nlat=48, seed=1):
    """The batched ternary ensemble: cube edges uniform in {16, 24, ..., 96},
    g_br uniform in [0.06, 0.12], noise seed 1000+k (numpy default_rng(seed))."""
    rng = np.random.default_rng(seed)
    edges = 16 + 8 * rng.integers(0, 11, nlat)
    g_br = 0.06 + 0.06 * rng.random(nlat)
    return [int(e) for e in edges], [float(g) for g in g_br]


def run_ensemble(steps, warmup, seq_steps, nlat=48):
    """FP32 batched persistent launches against the lattices run one after
    another on the plain FP64 path."""
    from lb3d_amd.gpu import LB3D_FP32_DEV, LB3D_FP64, Batch

    edges, g_br = ensemble_spec(nlat)
    dims = [(e, e, e) for e in edges]
    per = [dict(g_br=g) for g in g_br]
    common = dict(ncomp=3, g_bs=-0.006, g_ss=-0.003, tau_d=1.0, beta=10.0, d_0=1.0, bdist=2)
    sites = sum(e ** 3 for e in edges)
    out = {}
    for name, prec, nsteps in (("fp32_batched", LB3D_FP32_DEV, steps), ("fp64_sequential", LB3D_FP64, seq_steps)):
        b = Batch(dims, per_lattice=per, precision=prec, **common)
        for k in range(nlat):
            b.init_noise(k, 1000 + k, rho_mean=(0.7, 0.7, 0.7), amp=(0.01, 0.01, 0.0), dipoles=True)
        if prec == LB3D_FP64:
            b.set_plan(1)   # lattice after lattice, the single-lattice FP64 kernels
            chosen = "sequential"
        else:
            chosen = b.autotune(3).strip().splitlines()[-1]
        b.time_loop(warmup)
        b.synchronize()
        t0 = time.perf_counter()
        b.time_loop(nsteps)
        b.synchronize()
        dt = time.perf_counter() - t0
        ph, mps = b.profile_steps(3)
        out[name] = dict(mlups=sites * nsteps / dt / 1e6, ms_per_step=1e3 * dt / nsteps, plan=chosen,
                         phases_ms=ph, steps=nsteps)
        print(f"[bench] ensemble {name}: {out[name]['mlups']:.1f} MLUPS {chosen} {ph}", file=sys.stderr, flush=True)
        b.close()
    B = bytes_per_site(3, True, 4)
    r = out["fp32_batched"]
    return dict(lattices=nlat, sites=sites, edges=edges, value=round(r["mlups"], 1), unit="MLUPS",
                ms_per_step=round(r["ms_per_step"], 4), steps=r["steps"], plan=r["plan"],
                step_frac=round(r["mlups"] * 1e6 * B / (HBM_PEAK_GBS * 1e9), 4), bytes_per_site=B,
                phases_ms={k: round(v, 4) for k, v in r["phases_ms"].items()},
                fp64_sequential=dict(value=round(out["fp64_sequential"]["mlups"], 1), unit="MLUPS",
                                     steps=out["fp64_sequential"]["steps"]),
                speedup_vs_fp64_sequential=round(r["mlups"] / out["fp64_sequential"]["mlups"], 3))


def ensemble_accuracy(steps, nlat=48):
    """The FP32 batched ensemble against the FP64 batched ensemble after `steps`
    steps from the same initial state: density errors of every lattice (the
    worst lattice reported), beside the FP64 floor (FP64 against FP64 started
    from a mean density 2^-51 higher)."""
    from lb3d_amd.gpu import LB3D_FP32_DEV, LB3D_FP64, Batch

    G = np.array([2.0] * 6 + [1.0] * 13)
    edges, g_br = ensemble_spec(nlat)
    dims = [(e, e, e) for e in edges]
    per = [dict(g_br=g) for g in g_br]
    common = dict(ncomp=3, g_bs=-0.006, g_ss=-0.003, tau_d=1.0, beta=10.0, d_0=1.0, bdist=2)
    rho = {}
    for tag, prec, shift in (("fp64", LB3D_FP64, 0.0), ("fp32", LB3D_FP32_DEV, 0.0), ("floor", LB3D_FP64, 2.0 ** -51)):
        b = Batch(dims, per_lattice=per, precision=prec, **common)
        m = 0.7 + shift
        for k in range(nlat):
            b.init_noise(k, 1000 + k, rho_mean=(m, m, m), amp=(0.01, 0.01, 0.0), dipoles=True)
        b.time_loop(steps)
        b.synchronize()
        # densities from the populations (the collision conserves them per site)
        rho[tag] = [np.einsum("i,cizyx->czyx", G, b.populations(k)) for k in range(nlat)]
        b.close()
    out = {}
    for tag in ("fp32", "floor"):
        linf = [float(np.abs(a - r).max() / np.abs(r).max()) for a, r in zip(rho[tag], rho["fp64"])]
        l2 = [float(np.linalg.norm(a - r) / np.linalg.norm(r)) for a, r in zip(rho[tag], rho["fp64"])]
        # statistics that survive the flow's amplification of rounding: the
        # spread of each density field and each component's mass
        spread = [float(abs(a[0].std() - r[0].std()) / r[0].std()) for a, r in zip(rho[tag], rho["fp64"])]
        mass = [float(np.max(np.abs(a.sum(axis=(1, 2, 3)) - r.sum(axis=(1, 2, 3))) / r.sum(axis=(1, 2, 3))))
                for a, r in zip(rho[tag], rho["fp64"])]
        out[tag] = dict(density_linf_rel=max(linf), density_l2_rel=max(l2),
                        worst_lattice_edge=edges[int(np.argmax(linf))],
                        lattices_within_1e5_linf=int(sum(v <= 1e-5 for v in linf)),
                        density_std_rel=max(spread), mass_rel=max(mass))
    res = dict(steps=steps, **out["fp32"], fp64_perturbed_floor=out["floor"])
    res["density_linf_over_floor"] = res["density_linf_rel"] / max(out["floor"]["density_linf_rel"], 1e-300)
    print(f"[bench] ensemble accuracy after {steps} steps: density Linf {res['density_linf_rel']:.3e} "
          f"L2 {res['density_l2_rel']:.3e} (floor {out['floor']['density_linf_rel']:.3e})", file=sys.stderr, flush=True)
    return res


def host_threads():
    """Threads for the multi-core CPU baseline: OMP_NUM_THREADS if set, else
    the CPUs this process may run on."""
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        return int(env)
    return len(os.sched_getaffinity(0))


def cpu_baseline(sample_n, sample_steps, single_n, single_steps):
    """The oracle (FP64 C restatement of the reference loop) on a bounded
    sample of the workload: all host threads, then one core."""
    from oracle.pyoracle import Oracle, set_threads

    def timed(nn, steps, threads):
        set_threads(threads)
        o = Oracle(nn, nn, nn, ncomp=2, g_br=0.08, bdist=2)
        o.init_noise(1, rho_mean=(0.7, 0.7, 0.7), amp=(0.01, 0.01, 0.0), dipoles=False)
        o.bc_init()
        t0 = time.perf_counter()
        o.run(steps)
        dt = time.perf_counter() - t0
        o.close()
        return nn ** 3 * steps / dt / 1e6, dt

    th = host_threads()
    v, dt = timed(sample_n, sample_steps, th)
    v1, dt1 = timed(single_n, single_steps, 1)
    cpu = open("/proc/cpuinfo").read() if os.path.exists("/proc/cpuinfo") else ""
    model = next((ln.split(":", 1)[1].strip() for ln in cpu.splitlines() if ln.startswith("model name")), "")
    return dict(value=round(v, 4), unit="MLUPS", cores=th, kind="port",
                sample=f"oracle (FP64 C restatement of the reference loop, OpenMP over sites, {th} threads), "
                       f"binary {sample_n}^3 periodic, {sample_steps} steps, same physics as the workload",
                seconds=round(dt, 3), cpu_model=model,
                single_core=dict(value=round(v1, 4), unit="MLUPS", cores=1,
                                 sample=f"binary {single_n}^3, {single_steps} steps", seconds=round(dt1, 3)))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--size", type=int, default=256)
    ap.add_argument("--profile-steps", type=int, default=10)
    ap.add_argument("--cpu-sample-size", type=int, default=96)
    ap.add_argument("--cpu-sample-steps", type=int, default=20)
    ap.add_argument("--cpu-single-size", type=int, default=64)
    ap.add_argument("--cpu-single-steps", type=int, default=12)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--ensemble-steps", type=int, default=300)
    ap.add_argument("--ensemble-fp64-steps", type=int, default=20)
    ap.add_argument("--no-ensemble", action="store_true")
    ap.add_argument("--ensemble-accuracy-steps", type=int, default=1000,
                    help="FP32 against FP64 ensemble after this many steps (0: skip)")
    ap.add_argument("--no-floor", action="store_true")
    ap.add_argument("--no-configs", action="store_true",
                    help="skip the FP32-vs-FP64 accuracy runs of the ternary and Lees-Edwards configurations")
    ap.add_argument("--config-steps-scale", type=float, default=1.0)
    ap.add_argument("--no-traffic", action="store_true",
                    help="skip the two rocprofv3 --pmc child runs that measure the dominant kernel's HBM traffic")
    ap.add_argument("--traffic-probe", default=None, help=argparse.SUPPRESS)
    ap.add_argument("--ternary-probe", type=int, default=0, help=argparse.SUPPRESS)
    ap.add_argument("--ensemble-probe", type=int, default=None, help=argparse.SUPPRESS)
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
        ensemble_probe(args.ensemble_probe)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(the path is built and measured on a single MI355X)", file=sys.stderr)
        sys.exit(2)

    from lb3d_amd.gpu import LB3D_BC_LEES_EDWARDS, LB3D_FP32_DEV, LB3D_FP64

    n, ncomp = args.size, 2
    print(f"[bench] FP64 plain {n}^3 binary: warmup {args.warmup}, steps {args.steps}", file=sys.stderr, flush=True)
    r64 = run_form(n, ncomp, LB3D_FP64, args.steps, args.warmup, args.profile_steps)
    print(f"[bench] FP64 {r64['mlups']:.1f} MLUPS {r64['kernels']}", file=sys.stderr, flush=True)
    print(f"[bench] FP32 deviation {n}^3 binary", file=sys.stderr, flush=True)
    r32 = run_form(n, ncomp, LB3D_FP32_DEV, args.steps, args.warmup, args.profile_steps)
    print(f"[bench] FP32 {r32['mlups']:.1f} MLUPS {r32['kernels']}", file=sys.stderr, flush=True)

    linf, l2 = errs(r32["rho"], r64["rho"])
    vlinf, vl2 = errs(r32["u"], r64["u"])
    acc = dict(steps=args.warmup + args.steps, density_linf_rel=linf, density_l2_rel=l2,
               velocity_linf_rel=vlinf, velocity_l2_rel=vl2,
               density_std_rel=float(abs(r32["rho"][0].std() - r64["rho"][0].std()) / r64["rho"][0].std()),
               mass_rel=float(max(abs(a - b) / abs(b) for a, b in zip(r32["mass"][:ncomp], r64["mass"][:ncomp]))))
    del r32["rho"], r32["u"]
    if not args.no_floor:
        # intrinsic floor of the unstable flow: FP64 against FP64 whose initial
        # mean density differs by 2^-51 (about 2 ulp)
        print("[bench] FP64 floor run (perturbed initial state)", file=sys.stderr, flush=True)
        rf = run_form(n, ncomp, LB3D_FP64, args.steps, args.warmup, 1, rho_shift=2.0 ** -51)
        flinf, fl2 = errs(rf["rho"], r64["rho"])
        acc["fp64_perturbed_floor"] = dict(density_linf_rel=flinf, density_l2_rel=fl2,
                                           velocity_linf_rel=errs(rf["u"], r64["u"])[0],
                                           perturbation="initial mean density + 2^-51")
        acc["density_linf_over_floor"] = linf / max(flinf, 1e-300)
        del rf
    del r64["rho"], r64["u"]

    ns = n ** 3
    B32 = bytes_per_site(ncomp, False, 4)
    B64 = bytes_per_site(ncomp, False, 8)
    kname = "fused" if "fused" in r32["kernels"] else "collide"
    if kname == "fused":
        # the one-pass plan launches exactly one kernel per periodic step:
        # its mean duration is the HIP-event time of the timed region / steps
        kc, how = r32["event_ms_per_step"], "HIP events on the library stream around the timed steps / steps"
    else:
        kc, how = r32["kernels"][kname], "HIP events on the library stream around each launch"
    achieved = B32 * ns / (kc * 1e-3) / 1e9
    roofline = dict(bound="hbm", achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                    frac=round(achieved / HBM_PEAK_GBS, 4), traffic=None,
                    kernel=f"k_{'fused2' if kname == 'fused' else 'collide'} (FP32 deviation), plan {r32['plan']}",
                    algorithmic_bytes_per_launch=B32 * ns, mean_launch_ms=round(kc, 5), launch_timing=how,
                    step_frac=round(r32["mlups"] * 1e6 * B32 / (HBM_PEAK_GBS * 1e9), 4))
    if not args.no_traffic:
        key = "k_fused2" if kname == "fused" else "k_collide"
        tr, why = measure_traffic(n, parse_plan(r32["plan"]), key)
        if tr:
            roofline["traffic"] = round(tr["bytes_per_launch"])
            roofline["traffic_detail"] = tr
            print(f"[bench] traffic {tr['bytes_per_launch'] / 1e9:.3f} GB per launch "
                  f"(algorithmic {B32 * ns / 1e9:.3f})", file=sys.stderr, flush=True)
        else:
            roofline["traffic_note"] = why
            print(f"[bench] traffic not measured: {why}", file=sys.stderr, flush=True)
    print("[bench] HBM ceiling probe", file=sys.stderr, flush=True)
    ceiling = hbm_ceiling()
    # the dominant kernel against what a plain copy reaches on this card
    ceiling["kernel_algorithmic_frac_of_copy"] = round(achieved / ceiling["copy"], 4)
    if roofline["traffic"]:
        ceiling["kernel_traffic_frac_of_copy"] = round(roofline["traffic"] / (kc * 1e-3) / 1e9 / ceiling["copy"], 4)
    print(f"[bench] ceiling {ceiling}", file=sys.stderr, flush=True)
    fp64 = dict(value=round(r64["mlups"], 1), unit="MLUPS", ms_per_step=round(r64["ms_per_step"], 4),
                step_frac=round(r64["mlups"] * 1e6 * B64 / (HBM_PEAK_GBS * 1e9), 4),
                kernels_ms={k: round(v, 5) for k, v in r64["kernels"].items()}, plan=r64["plan"])
    ens = None
    if not args.no_ensemble:
        ens = run_ensemble(args.ensemble_steps, 5, args.ensemble_fp64_steps)
        # the step's algorithmic rate against what a plain copy reaches on this card
        ens["step_frac_of_copy"] = round(ens["value"] * 1e6 * ens["bytes_per_site"] / 1e9 / ceiling["copy"], 4)
        if args.ensemble_accuracy_steps > 0:
            ens["accuracy_fp32_vs_fp64"] = ensemble_accuracy(args.ensemble_accuracy_steps)
    configs = None
    if not args.no_configs:
        sc = args.config_steps_scale
        configs = {
            "ternary_128": config_accuracy(
                "ternary 128^3", (128, 128, 128), max(1, int(2000 * sc)),
                dict(ncomp=3, g_br=0.08, g_bs=-0.006, g_ss=-0.003, tau_d=1.0, beta=10.0, d_0=1.0, bdist=2),
                (0.01, 0.01, 0.0)),
            "lees_edwards_256x64x64": config_accuracy(
                "Lees-Edwards 256x64x64", (256, 64, 64), max(1, int(5000 * sc)),
                dict(ncomp=2, g_br=0.08, bdist=2, boundary=LB3D_BC_LEES_EDWARDS, shear_u=0.05),
                (0.01, 0.01, 0.0)),
        }
    cpu = None
    if not args.no_cpu_baseline:
        print(f"[bench] CPU baseline oracle {args.cpu_sample_size}^3 x{args.cpu_sample_steps} (all threads), "
              f"{args.cpu_single_size}^3 x{args.cpu_single_steps} (one core)", file=sys.stderr, flush=True)
        cpu = cpu_baseline(args.cpu_sample_size, args.cpu_sample_steps, args.cpu_single_size,
                           args.cpu_single_steps)
    line = {
        "metric": METRIC,
        "value": round(r32["mlups"], 1),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(r32["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"binary Shan-Chen spinodal D3Q19 {n}^3 periodic, g_br=0.08, tau=1, bdist=2, "
                               "rho=0.7+-0.01 (seed 1), FP32 deviation storage",
                   "lattice": [n, n, n], "components": ncomp, "bytes_per_site": B32},
        "roofline": roofline,
        "hbm_ceiling": ceiling,
        "cpu_baseline": cpu,
        "fp64_plain": fp64,
        "speedup_fp32_vs_fp64": round(r32["mlups"] / r64["mlups"], 3),
        "accuracy_fp32_vs_fp64": acc,
        "kernels_ms": {k: round(v, 5) for k, v in r32["kernels"].items()},
        "kernels_ms_note": (f"HIP events around each launch over {args.profile_steps} steps launched one by one "
                            "after the timed loop (not graph-replayed, with an event pair per launch); "
                            "roofline.mean_launch_ms is the timed loop's"),
        "batched_ternary_ensemble": ens,
        "config_accuracy": configs,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
