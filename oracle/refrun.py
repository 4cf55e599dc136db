"""Runs the reference LB3D built by oracle/ref_build.sh (oracle/_ref/).

Test and baseline infrastructure only: the FP64 fixture script
(tests/golden/make_ref_fp64.py) and the CPU-baseline leg of bench.py use it;
the product never does.  It writes the reference's namelist input
(code/lb3d_config.F90:319-335 lists the groups), its 'ori' checkpoints
(code/lb3d_io_checkpoint.F90:200-250: one unformatted record per site, x
outer .. z inner, n_r[19] n_b[19] [n_s[19] da[3] db[3]] rock_state, with
4-byte record markers), starts `mpirun -n R lb3d_<flavour> -f <input>` in a
scratch directory, and reads back the checkpoint it wrote and the time of its
step loop (the "Total time" line of code/lb3d_timer.F90).
"""
import glob
import os
import shutil
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF_DIR = os.path.join(HERE, "_ref")
MPI_PREFIX = os.environ.get("LB3D_MPI_PREFIX", "/opt/conda")


def binary_path(ternary):
    return os.path.join(REF_DIR, "lb3d_ternary" if ternary else "lb3d_binary")


def mpirun_path():
    for p in (os.path.join(MPI_PREFIX, "bin", "mpirun"), shutil.which("mpirun")):
        if p and os.path.exists(p):
            return p
    return None


def available(ternary):
    """(True, '') when the reference build and mpirun are present, else (False, why)."""
    b = binary_path(ternary)
    if not os.path.exists(b):
        return False, f"{b} not built (oracle/ref_build.sh)"
    if mpirun_path() is None:
        return False, "no mpirun"
    return True, ""


def record_length(ternary):
    return 64 if ternary else 39


def namelist(name, dims, steps, ternary, restore=False, n_checkpoint=None, init_rho=0.7, lees_edwards=False,
             g_br=0.08, g_bs=-0.006, g_ss=-0.003, tau_d=1.0, beta=10.0, d_0=1.0, bdist=2, shear_u=0.0,
             tau=(1.0, 1.0, 1.0), amass=(1.0, 1.0, 1.0), seed=1, g_accn=(0.0, 0.0, 0.0), g_accn_min=(0, 0, 0),
             g_accn_max=(0, 0, 0), inv_fluid=0, inv_type=0, inv_flux=(0.0, 0.0, 0.0)):
    """The reference's input file: `steps` timesteps of a dims lattice, either
    restored from ./checkpoint/checkpoint_<name>_t00000000-0000000000_p000000.bin
    ('ori' format, init_cond 7) or started from INIT_RAND (init_cond -1,
    mean density init_rho per component); no dumps; a checkpoint every
    n_checkpoint steps when given.  The physics arguments are those of
    oracle.pyoracle.Oracle: g_accn, g_accn_min and g_accn_max are (x, y, z)
    (the reference's g_accn_x, g_accn_y, g_accn and their box bounds, 1-based,
    0 for the lattice's extent), inv_fluid 12 with inv_type and the fluxes
    inv_flux = (pr, pb, pg) is the Zou-He boundary on the z planes."""
    nx, ny, nz = dims
    f = 2.0 * init_rho  # INIT_RAND draws each population in [0, f g_i^-1 / 19]
    ck = n_checkpoint if n_checkpoint else 10 * steps + 10
    zou_he = inv_fluid == 12
    inv = 5 if lees_edwards else (12 if zou_he else -1)
    return f"""&fixed_input
nx = {nx}
ny = {ny}
nz = {nz}
seed = {seed}
obs_file = 'empty.dat'
boundary_cond = {-1 if (lees_edwards or zou_he) else 0}
/

&variable_input
n_iteration = {steps}
n_sci_start = 0
sci_int = .false.
sci_sur = .false.
sci_od = .false.
sci_wd = .false.
sci_dir = .false.
sci_vel = .false.
sci_flo = .false.
sci_arrows = .false.
sci_pressure = .false.
post = .false.
folder = './output'
cpfolder = '.'
srccpfolder = './checkpoint'
gr_out_file = '{name}'
init_cond = {7 if restore else -1}
fr = {f}
fb = {f}
fg = {f if ternary else 0.0}
inv_fluid = {inv}
inv_type = {inv_type}
pr = {inv_flux[0]}
pb = {inv_flux[1]}
pg = {inv_flux[2]}
beta = {beta}
/

&lbe_input
amass_r = {amass[0]}
amass_b = {amass[1]}
amass_s = {amass[2]}
tau_r = {tau[0]}
tau_b = {tau[1]}
tau_s = {tau[2]}
tau_d = {tau_d}
g_br = {g_br}
g_bs = {g_bs if ternary else 0.0}
g_ss = {g_ss if ternary else 0.0}
d_0 = {d_0}
bdist = {bdist}
perturbation = 0
shear_u = {shear_u}
shear_omega = 0
g_accn = {g_accn[2]}
g_accn_x = {g_accn[0]}
g_accn_y = {g_accn[1]}
g_accn_min = {g_accn_min[2]}
g_accn_min_x = {g_accn_min[0]}
g_accn_min_y = {g_accn_min[1]}
g_accn_max = {g_accn_max[2]}
g_accn_max_x = {g_accn_max[0]}
g_accn_max_y = {g_accn_max[1]}
n_checkpoint = {ck}
restore_string = 't00000000-0000000000'
num_chkp_files = 0
psifunc = 2
dump_format = 'xdr'
checkpoint_safe = .true.
checkpoint_format = 'ori'
write_AVS_fld = .false.
dump_double = .false.
n_sanity_check = {10 * steps + 10}
/
"""


def write_ori(path, rec):
    """An 'ori' checkpoint from records rec[x, y, z, :] (float64)."""
    nx, ny, nz, r = rec.shape
    mark = np.array([8 * r], dtype="<i4").tobytes()
    flat = np.ascontiguousarray(rec, dtype="<f8").reshape(-1, r)
    with open(path, "wb") as fh:
        for row in flat:
            fh.write(mark)
            fh.write(row.tobytes())
            fh.write(mark)


def read_ori(path, dims, r):
    nx, ny, nz = dims
    raw = np.fromfile(path, dtype=np.uint8)
    per = 8 + 8 * r
    assert raw.size == nx * ny * nz * per, (path, raw.size)
    body = raw.reshape(-1, per)[:, 4:4 + 8 * r].copy()
    return body.view("<f8").reshape(nx, ny, nz, r)


def records_from_sites(interior, ternary):
    """Records rec[x, y, z, :] from an lbe_site interior array (z, y, x, site)
    in the layout of oracle.pyoracle.Oracle.N."""
    I = np.transpose(interior, (2, 1, 0, 3))
    if ternary:
        return np.concatenate([I[..., 0:19], I[..., 24:43], I[..., 48:67], I[..., 72:75], I[..., 75:78],
                               I[..., 78:79]], axis=-1)
    return np.concatenate([I[..., 0:19], I[..., 24:43], I[..., 48:49]], axis=-1)


def run(name, dims, steps, ternary, ranks=1, start=None, checkpoint_at=None, timeout=600, **kw):
    """Runs the reference; returns (step-loop seconds, records at checkpoint_at
    or None).  start: records of the initial state (restored), else INIT_RAND."""
    b = binary_path(ternary)
    with tempfile.TemporaryDirectory(prefix="lb3d_ref_") as d:
        os.makedirs(os.path.join(d, "output"))
        os.makedirs(os.path.join(d, "checkpoint"))
        if start is not None:
            write_ori(os.path.join(d, "checkpoint", f"checkpoint_{name}_t00000000-0000000000_p000000.bin"), start)
        with open(os.path.join(d, "run.in"), "w") as fh:
            fh.write(namelist(name, dims, steps, ternary, restore=start is not None, n_checkpoint=checkpoint_at, **kw))
        cmd = f"ulimit -s unlimited 2>/dev/null; exec '{mpirun_path()}' -n {int(ranks)} '{b}' -f run.in"
        p = subprocess.run(["bash", "-c", cmd], cwd=d, capture_output=True, text=True, timeout=timeout)
        if p.returncode != 0:
            raise RuntimeError(f"reference run failed ({p.returncode}): {p.stdout[-2000:]} {p.stderr[-2000:]}")
        secs = None
        for line in p.stdout.splitlines():
            if line.startswith("Total") and "time:" in line:
                secs = float(line.split()[4])  # "Total time: <ave> ave <max> max <min> min"
        out = None
        if checkpoint_at:
            hits = glob.glob(os.path.join(d, "output", f"checkpoint_{name}_t{checkpoint_at:08d}-*_p000000.bin"))
            if len(hits) != 1:
                raise RuntimeError(f"checkpoint of t={checkpoint_at} not found: {os.listdir(os.path.join(d, 'output'))}")
            out = read_ori(hits[0], dims, record_length(ternary))
        return secs, out
