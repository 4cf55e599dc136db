"""Python host for the MI355X LB3D timestep library (ctypes over the C ABI of
include/lb3d_gpu.h).

`Lattice` mirrors the reference's Fortran call surface for one lattice:
  bc_init()            lb3d_bc_init(4)           code/lb3d_bc.F90:48-90
  force_apply(stage)   lb3d_force_apply(stage)   code/lb3d_force.F90:98-121
  bc_apply(stage)      lb3d_bc_apply(stage)      code/lb3d_bc.F90:95-155
  advection()          lb3d_advection()          code/lb3d_advection.F90:40-50
  collision()          lb3d_collision()          code/lb3d_collision.F90:41-47
  time_loop(n)         lb3d_time_loop()          code/lb3d_time_loop.F90:42-97
  sanity_check()       lbe_sanity_check(N, ...)  code/lb3d_sanity.F90:79-224
Errors raise LB3DError (the reference aborts through abend/MPI_Abort).  The
library is required: there is no CPU fallback, and a missing or unloadable
liblb3d_gpu.so raises at import of the binding.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblb3d_gpu.so")

LB3D_FP64 = 0
LB3D_FP32_DEV = 1
LB3D_BC_PERIODIC = 0
LB3D_BC_LEES_EDWARDS = 1
ERRORS = {1: "EINVAL", 2: "EHIP", 3: "ESTATE", 4: "ENOMEM"}

# Every entry point declared in include/lb3d_gpu.h.
EXPORTS = [
    "lb3d_gpu_abi_version", "lb3d_gpu_last_error", "lb3d_gpu_create", "lb3d_gpu_destroy",
    "lb3d_gpu_upload_sites", "lb3d_gpu_download_sites", "lb3d_gpu_init_noise",
    "lb3d_gpu_bc_init", "lb3d_gpu_force_apply", "lb3d_gpu_bc_apply", "lb3d_gpu_advection",
    "lb3d_gpu_collision", "lb3d_gpu_run_steps", "lb3d_gpu_run_steps_timed", "lb3d_gpu_get_nt", "lb3d_gpu_diagnostics",
    "lb3d_gpu_le_state", "lb3d_gpu_get_populations", "lb3d_gpu_profile_steps",
    "lb3d_gpu_get_density", "lb3d_gpu_get_momentum", "lb3d_gpu_le_halo", "lb3d_gpu_sanity",
    "lb3d_gpu_batch_sanity",
    "lb3d_gpu_synchronize", "lb3d_gpu_autotune", "lb3d_gpu_set_plan",
    "lb3d_gpu_batch_create", "lb3d_gpu_batch_destroy", "lb3d_gpu_batch_init_noise",
    "lb3d_gpu_batch_upload_sites", "lb3d_gpu_batch_run_steps", "lb3d_gpu_batch_get_nt",
    "lb3d_gpu_batch_synchronize", "lb3d_gpu_batch_get_populations", "lb3d_gpu_batch_diagnostics",
    "lb3d_gpu_batch_set_plan", "lb3d_gpu_batch_set_groups", "lb3d_gpu_batch_autotune",
    "lb3d_gpu_batch_profile_steps", "lb3d_gpu_batch_set_collide_moments", "lb3d_gpu_step_noalloc",
    "lb3d_gpu_batch_run_steps_timed", "lb3d_gpu_get_arrows",
]


class Sanity(ctypes.Structure):
    """lb3d_sanity of include/lb3d_gpu.h."""
    _fields_ = [("mass", ctypes.c_double * 3), ("momentum", ctypes.c_double * 3),
                ("vel_sum", (ctypes.c_double * 3) * 3), ("max_vel", (ctypes.c_double * 3) * 3),
                ("max_pos", ((ctypes.c_int32 * 3) * 3) * 3), ("pad0", ctypes.c_int32),
                ("fluid_sites", ctypes.c_int64), ("nan_sites", ctypes.c_int64),
                ("negative_sites", ctypes.c_int64), ("fast_sites", ctypes.c_int64)]

    def as_dict(self):
        return dict(mass=np.array(self.mass[:]), momentum=np.array(self.momentum[:]),
                    vel_sum=np.array([r[:] for r in self.vel_sum]),
                    max_vel=np.array([r[:] for r in self.max_vel]),
                    max_pos=np.array([[p[:] for p in r] for r in self.max_pos]),
                    fluid_sites=self.fluid_sites, nan_sites=self.nan_sites,
                    negative_sites=self.negative_sites, fast_sites=self.fast_sites)


class LB3DError(RuntimeError):
    def __init__(self, code, where, msg):
        super().__init__(f"{where}: {ERRORS.get(code, code)}: {msg}")
        self.code = code


class Params(ctypes.Structure):
    """lb3d_params of include/lb3d_gpu.h."""
    _fields_ = [
        ("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("nz", ctypes.c_int32),
        ("ncomp", ctypes.c_int32), ("bdist", ctypes.c_int32), ("precision", ctypes.c_int32),
        ("boundary", ctypes.c_int32), ("restore", ctypes.c_int32),
        ("amass", ctypes.c_double * 3), ("tau", ctypes.c_double * 3),
        ("tau_d", ctypes.c_double), ("beta", ctypes.c_double), ("d_0", ctypes.c_double),
        ("g_br", ctypes.c_double), ("g_bs", ctypes.c_double), ("g_ss", ctypes.c_double),
        ("shear_u", ctypes.c_double), ("shear_omega", ctypes.c_double),
        ("shear_sum", ctypes.c_double),
        ("g_accn", ctypes.c_double * 3),
        ("g_accn_min", ctypes.c_int32 * 3), ("g_accn_max", ctypes.c_int32 * 3),
        ("rho0", ctypes.c_double * 3),
        ("g_wr", ctypes.c_double), ("g_wb", ctypes.c_double),
        ("tau_wr", ctypes.c_double), ("tau_wb", ctypes.c_double),
        ("rock_colour_r", ctypes.c_double), ("rock_colour_b", ctypes.c_double),
        ("inv_fluid", ctypes.c_int32), ("inv_type", ctypes.c_int32),
        ("inv_flux", ctypes.c_double * 3),
    ]


_lib = None


def lib():
    """The loaded library; raises if it is missing (no fallback path)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with python -m lb3d_amd.build "
                               "(hipcc, gfx950); there is no CPU fallback")
        L = ctypes.CDLL(LIB_PATH)
        vp, ip, dp = ctypes.c_void_p, ctypes.c_int32, ctypes.POINTER(ctypes.c_double)
        L.lb3d_gpu_last_error.restype = ctypes.c_char_p
        L.lb3d_gpu_create.argtypes = [ctypes.POINTER(Params), ctypes.POINTER(vp)]
        L.lb3d_gpu_destroy.argtypes = [vp]
        L.lb3d_gpu_upload_sites.argtypes = [vp, dp, ip, ip, ip]
        L.lb3d_gpu_download_sites.argtypes = [vp, dp, ip, ip]
        L.lb3d_gpu_init_noise.argtypes = [vp, ctypes.c_uint64, dp, dp, ip]
        L.lb3d_gpu_bc_init.argtypes = [vp]
        L.lb3d_gpu_force_apply.argtypes = [vp, ip]
        L.lb3d_gpu_bc_apply.argtypes = [vp, ip, ip]
        L.lb3d_gpu_advection.argtypes = [vp]
        L.lb3d_gpu_collision.argtypes = [vp]
        L.lb3d_gpu_run_steps.argtypes = [vp, ip]
        L.lb3d_gpu_run_steps_timed.argtypes = [vp, ip, dp]
        L.lb3d_gpu_get_nt.argtypes = [vp, ctypes.POINTER(ip)]
        L.lb3d_gpu_diagnostics.argtypes = [vp, dp, dp, ctypes.POINTER(ctypes.c_int64)]
        L.lb3d_gpu_le_state.argtypes = [vp, dp, ctypes.POINTER(ip), dp, dp]
        L.lb3d_gpu_get_populations.argtypes = [vp, dp]
        L.lb3d_gpu_get_density.argtypes = [vp, dp]
        L.lb3d_gpu_get_momentum.argtypes = [vp, dp]
        L.lb3d_gpu_le_halo.argtypes = [vp, dp]
        L.lb3d_gpu_sanity.argtypes = [vp, ctypes.POINTER(Sanity)]
        L.lb3d_gpu_batch_sanity.argtypes = [vp, ip, ctypes.POINTER(Sanity)]
        L.lb3d_gpu_synchronize.argtypes = [vp]
        L.lb3d_gpu_autotune.argtypes = [vp, ip, ctypes.c_char_p, ip]
        L.lb3d_gpu_set_plan.argtypes = [vp, ip, ip, ip, ip, ip]
        L.lb3d_gpu_profile_steps.argtypes = [vp, ip, ip, ctypes.POINTER(ctypes.c_char_p), dp,
                                             ctypes.POINTER(ip), dp]
        L.lb3d_gpu_batch_create.argtypes = [ip, ctypes.POINTER(Params), ctypes.POINTER(vp)]
        L.lb3d_gpu_batch_destroy.argtypes = [vp]
        L.lb3d_gpu_batch_init_noise.argtypes = [vp, ip, ctypes.c_uint64, dp, dp, ip]
        L.lb3d_gpu_batch_upload_sites.argtypes = [vp, ip, dp, ip, ip]
        L.lb3d_gpu_batch_run_steps.argtypes = [vp, ip]
        L.lb3d_gpu_batch_run_steps_timed.argtypes = [vp, ip, dp]
        L.lb3d_gpu_get_arrows.argtypes = [vp, dp]
        L.lb3d_gpu_batch_get_nt.argtypes = [vp, ctypes.POINTER(ip)]
        L.lb3d_gpu_batch_synchronize.argtypes = [vp]
        L.lb3d_gpu_batch_get_populations.argtypes = [vp, ip, dp]
        L.lb3d_gpu_batch_diagnostics.argtypes = [vp, ip, dp, dp, ctypes.POINTER(ctypes.c_int64)]
        L.lb3d_gpu_batch_set_plan.argtypes = [vp, ip, ip, ip]
        L.lb3d_gpu_batch_set_collide_moments.argtypes = [vp, ip, ip, ip, ip]
        L.lb3d_gpu_step_noalloc.argtypes = [ctypes.POINTER(Params), vp, vp, vp, vp, vp]
        L.lb3d_gpu_batch_set_groups.argtypes = [vp, ctypes.c_uint64]
        L.lb3d_gpu_batch_autotune.argtypes = [vp, ip, ctypes.c_char_p, ip]
        L.lb3d_gpu_batch_profile_steps.argtypes = [vp, ip, dp, dp]
        _lib = L
    return _lib


def check(rc, where):
    if rc != 0:
        raise LB3DError(rc, where, lib().lb3d_gpu_last_error().decode())


def make_params(nx, ny, nz, ncomp=2, bdist=2, precision=LB3D_FP64, boundary=LB3D_BC_PERIODIC,
                restore=0, amass=(1.0, 1.0, 1.0), tau=(1.0, 1.0, 1.0), tau_d=1.0, beta=10.0,
                d_0=1.0, g_br=0.08, g_bs=-0.006, g_ss=-0.003, shear_u=0.0, shear_omega=0.0,
                shear_sum=0.0, g_accn=(0.0, 0.0, 0.0), g_accn_min=(0, 0, 0),
                g_accn_max=(0, 0, 0), rho0=(0.7, 0.7, 0.7), g_wr=0.0, g_wb=0.0, tau_wr=1.0,
                tau_wb=1.0, rock_colour_r=0.0, rock_colour_b=0.0, inv_fluid=0, inv_type=0,
                inv_flux=(0.0, 0.0, 0.0)):
    p = Params()
    p.nx, p.ny, p.nz = nx, ny, nz
    p.ncomp, p.bdist, p.precision, p.boundary, p.restore = ncomp, bdist, precision, boundary, restore
    p.amass[:] = list(map(float, amass))
    p.tau[:] = list(map(float, tau))
    p.tau_d, p.beta, p.d_0 = float(tau_d), float(beta), float(d_0)
    p.g_br, p.g_bs, p.g_ss = float(g_br), float(g_bs), float(g_ss)
    p.shear_u, p.shear_omega, p.shear_sum = float(shear_u), float(shear_omega), float(shear_sum)
    p.g_accn[:] = list(map(float, g_accn))
    p.g_accn_min[:] = list(map(int, g_accn_min))
    p.g_accn_max[:] = list(map(int, g_accn_max))
    p.rho0[:] = list(map(float, rho0))
    p.g_wr, p.g_wb, p.tau_wr, p.tau_wb = float(g_wr), float(g_wb), float(tau_wr), float(tau_wb)
    p.rock_colour_r, p.rock_colour_b = float(rock_colour_r), float(rock_colour_b)
    p.inv_fluid, p.inv_type = int(inv_fluid), int(inv_type)
    p.inv_flux[:] = list(map(float, inv_flux))
    return p


class Lattice:
    """One lattice on the GPU behind the reference's call surface."""

    def __init__(self, nx, ny, nz, **kw):
        self.nx, self.ny, self.nz = nx, ny, nz
        self.params = make_params(nx, ny, nz, **kw)
        self.ncomp = self.params.ncomp
        self.site_doubles = 79 if self.ncomp == 3 else 49
        h = ctypes.c_void_p()
        check(lib().lb3d_gpu_create(ctypes.byref(self.params), ctypes.byref(h)), "lb3d_gpu_create")
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            lib().lb3d_gpu_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # --- state ---------------------------------------------------------
    def upload_sites(self, whole_N, halo_extent=1, nt0=0):
        a = np.ascontiguousarray(whole_N, dtype=np.float64)
        check(lib().lb3d_gpu_upload_sites(self._h, a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                          halo_extent, self.site_doubles, nt0), "upload_sites")

    def download_sites(self, halo_extent=1):
        h = halo_extent
        out = np.zeros((self.nz + 2 * h, self.ny + 2 * h, self.nx + 2 * h, self.site_doubles))
        check(lib().lb3d_gpu_download_sites(self._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                            h, self.site_doubles), "download_sites")
        return out

    def init_noise(self, seed, rho_mean=(0.7, 0.7, 0.7), amp=(0.01, 0.01, 0.0), dipoles=True):
        m = (ctypes.c_double * 3)(*rho_mean)
        a = (ctypes.c_double * 3)(*amp)
        check(lib().lb3d_gpu_init_noise(self._h, ctypes.c_uint64(seed), m, a, 1 if dipoles else 0),
              "init_noise")

    def populations(self):
        """Interior populations in FP64, shape (ncomp, 19, nz, ny, nx)."""
        out = np.empty((self.ncomp, 19, self.nz, self.ny, self.nx))
        check(lib().lb3d_gpu_get_populations(self._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))),
              "get_populations")
        return out

    def density(self):
        """Interior densities rho_c in FP64, shape (ncomp, nz, ny, nx)."""
        out = np.empty((self.ncomp, self.nz, self.ny, self.nx))
        check(lib().lb3d_gpu_get_density(self._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))),
              "get_density")
        return out

    def sanity(self):
        """The full sanity check (code/lb3d_sanity.F90:79-384) as a dict."""
        r = Sanity()
        check(lib().lb3d_gpu_sanity(self._h, ctypes.byref(r)), "sanity")
        return r.as_dict()

    def momentum(self):
        """Interior momenta P_c = sum_i g_i n_c,i c_i in FP64, shape (ncomp, 3, nz, ny, nx)."""
        out = np.empty((self.ncomp, 3, self.nz, self.ny, self.nx))
        check(lib().lb3d_gpu_get_momentum(self._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))),
              "get_momentum")
        return out

    def arrows(self):
        """The reference's arr observable, shape (nz, ny, nx, 3) in FP64."""
        out = np.empty((self.nz, self.ny, self.nx, 3))
        check(lib().lb3d_gpu_get_arrows(self._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))), "get_arrows")
        return out

    def le_halo(self):
        """Lees-Edwards x-halo planes that the pre-advection pass builds from the
        current populations and offset, FP64, shape (2, 2, 19, nz, ny): plane
        x = 0 then x = nx+1, component, direction, z, y."""
        out = np.empty((2, 2, 19, self.nz, self.ny))
        check(lib().lb3d_gpu_le_halo(self._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))),
              "le_halo")
        return out

    def shear_stress(self, minx=1, maxx=None):
        """Momentum transfer through the planes x = minx and x = maxx (1-based),
        the reference's local_fluid_momentum_transfer and dump_stress
        (code/lb3d_io_stress.F90:38-173, one rank): vectors 9, 10 enter plane
        minx from x = minx-1 and their bounces 13, 14 leave it; 13, 14 enter
        plane maxx from maxx+1.  The x = 0 / nx+1 planes are the pre-advection
        halo (le_halo for Lees-Edwards, periodic images otherwise).  Returns
        (out0, out1, mean) with out halved as dump_stress does and
        mean = (out1 - out0)/(ny*nz)."""
        maxx = self.nx if maxx is None else maxx
        if not (1 <= minx <= self.nx and 1 <= maxx <= self.nx):
            raise ValueError("stress planes must lie in 1..nx")
        pops = self.populations()
        if self.params.boundary == LB3D_BC_LEES_EDWARDS:
            halo = self.le_halo()
        else:
            halo = np.stack([pops[..., -1], pops[..., 0]])

        def plane(x):  # populations of plane x (0..nx+1), shape (ncomp, 19, nz, ny)
            if x == 0:
                return halo[0]
            if x == self.nx + 1:
                return halo[1]
            return pops[..., x - 1]

        cz = (0, 0, 0, 0, 1, -1, 0, 0, 1, -1, 0, 0, 1, -1, 1, -1, 1, -1, 0)
        bounce = (1, 0, 3, 2, 5, 4, 11, 10, 13, 12, 7, 6, 9, 8, 17, 16, 15, 14, 18)
        out = []
        for m, src, vecs in ((minx, minx - 1, (8, 9)), (maxx, maxx + 1, (12, 13))):
            A, B = plane(src), plane(m)
            tot = 0.0
            for v in vecs:  # g = 1 on these diagonal vectors
                r = bounce[v]
                # A at z - cz_v (periodic in z): roll by +cz_v along z
                tot += float((np.roll(A[:, v], cz[v], axis=1) * cz[v]).sum() - (B[:, r] * cz[r]).sum())
            out.append(0.5 * tot)
        return out[0], out[1], (out[1] - out[0]) / (self.ny * self.nz)

    def autotune(self, trial_steps=3):
        """Time every execution plan on this lattice and keep the fastest;
        returns the tuner's report."""
        buf = ctypes.create_string_buffer(32768)
        check(lib().lb3d_gpu_autotune(self._h, int(trial_steps), buf, len(buf)), "autotune")
        return buf.value.decode()

    def set_plan(self, fused, tx=64, ty=4, lz=64, xcd_group=1):
        check(lib().lb3d_gpu_set_plan(self._h, int(fused), int(tx), int(ty), int(lz), int(xcd_group)),
              "set_plan")

    def synchronize(self):
        check(lib().lb3d_gpu_synchronize(self._h), "synchronize")

    @property
    def nt(self):
        v = ctypes.c_int32()
        check(lib().lb3d_gpu_get_nt(self._h, ctypes.byref(v)), "get_nt")
        return v.value

    # --- the reference's call surface ----------------------------------
    def bc_init(self):
        check(lib().lb3d_gpu_bc_init(self._h), "bc_init")

    def force_apply(self, stage):
        check(lib().lb3d_gpu_force_apply(self._h, stage), "force_apply")

    def bc_apply(self, stage, nt=None):
        if nt is None:
            nt = self.nt + 1 if stage == 6 else self.nt
        check(lib().lb3d_gpu_bc_apply(self._h, stage, nt), "bc_apply")

    def advection(self):
        check(lib().lb3d_gpu_advection(self._h), "advection")

    def collision(self):
        check(lib().lb3d_gpu_collision(self._h), "collision")

    def time_step(self):
        """One timestep through the staged surface (code/lb3d_time_loop.F90:58-92)."""
        nt = self.nt + 1
        self.force_apply(6)
        self.bc_apply(6, nt)
        self.advection()
        self.bc_apply(8, nt)
        self.force_apply(8)
        self.collision()

    def time_loop(self, nsteps):
        """nsteps timesteps replayed from a captured HIP graph."""
        check(lib().lb3d_gpu_run_steps(self._h, int(nsteps)), "run_steps")

    def time_loop_timed(self, nsteps):
        """time_loop between HIP events on the library's stream; returns ms."""
        ms = ctypes.c_double()
        check(lib().lb3d_gpu_run_steps_timed(self._h, int(nsteps), ctypes.byref(ms)), "run_steps_timed")
        return ms.value

    def sanity_check(self):
        m = (ctypes.c_double * 3)()
        p = (ctypes.c_double * 3)()
        c = (ctypes.c_int64 * 3)()
        check(lib().lb3d_gpu_diagnostics(self._h, m, p, c), "diagnostics")
        return np.array(m[:]), np.array(p[:]), np.array(c[:])

    def le_state(self):
        s, f, u = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        lo = ctypes.c_int32()
        check(lib().lb3d_gpu_le_state(self._h, ctypes.byref(s), ctypes.byref(lo), ctypes.byref(f),
                                      ctypes.byref(u)), "le_state")
        return s.value, lo.value, f.value, u.value

    def profile_steps(self, nsteps):
        """Mean HIP-event duration per kernel launch over nsteps timesteps."""
        mx = 8
        names = (ctypes.c_char_p * mx)()
        ms = (ctypes.c_double * mx)()
        nk = ctypes.c_int32()
        mps = ctypes.c_double()
        check(lib().lb3d_gpu_profile_steps(self._h, int(nsteps), mx, names, ms, ctypes.byref(nk),
                                           ctypes.byref(mps)), "profile_steps")
        return {names[i].decode(): ms[i] for i in range(nk.value)}, mps.value


class Batch:
    """An ensemble of independent periodic lattices advanced together
    (lb3d_gpu_batch_* of include/lb3d_gpu.h).  `dims` is a list of (nx, ny, nz);
    `per_lattice` optionally a list of keyword dicts (e.g. g_br) per lattice;
    common keywords apply to all."""

    def __init__(self, dims, per_lattice=None, **common):
        self.dims = list(dims)
        per = per_lattice or [{} for _ in self.dims]
        arr = (Params * len(self.dims))()
        for k, (d, extra) in enumerate(zip(self.dims, per)):
            kw = dict(common)
            kw.update(extra)
            arr[k] = make_params(*d, **kw)
        self.ncomp = arr[0].ncomp
        self.site_doubles = 79 if self.ncomp == 3 else 49
        h = ctypes.c_void_p()
        check(lib().lb3d_gpu_batch_create(len(self.dims), arr, ctypes.byref(h)), "batch_create")
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            lib().lb3d_gpu_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def init_noise(self, k, seed, rho_mean=(0.7, 0.7, 0.7), amp=(0.01, 0.01, 0.0), dipoles=True):
        m = (ctypes.c_double * 3)(*rho_mean)
        a = (ctypes.c_double * 3)(*amp)
        check(lib().lb3d_gpu_batch_init_noise(self._h, k, ctypes.c_uint64(seed), m, a, 1 if dipoles else 0),
              "batch_init_noise")

    def upload_sites(self, k, whole_N, halo_extent=1):
        a = np.ascontiguousarray(whole_N, dtype=np.float64)
        check(lib().lb3d_gpu_batch_upload_sites(self._h, k, a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                                halo_extent, self.site_doubles), "batch_upload_sites")

    def time_loop(self, nsteps):
        check(lib().lb3d_gpu_batch_run_steps(self._h, int(nsteps)), "batch_run_steps")

    def time_loop_timed(self, nsteps):
        """time_loop between HIP events on the batch's stream; returns ms."""
        ms = ctypes.c_double()
        check(lib().lb3d_gpu_batch_run_steps_timed(self._h, int(nsteps), ctypes.byref(ms)), "batch_run_steps_timed")
        return ms.value

    def synchronize(self):
        check(lib().lb3d_gpu_batch_synchronize(self._h), "batch_synchronize")

    @property
    def nt(self):
        v = ctypes.c_int32()
        check(lib().lb3d_gpu_batch_get_nt(self._h, ctypes.byref(v)), "batch_get_nt")
        return v.value

    def populations(self, k):
        nx, ny, nz = self.dims[k]
        out = np.empty((self.ncomp, 19, nz, ny, nx))
        check(lib().lb3d_gpu_batch_get_populations(self._h, k, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))),
              "batch_get_populations")
        return out

    def sanity_check(self, k):
        m = (ctypes.c_double * 3)()
        p = (ctypes.c_double * 3)()
        c = (ctypes.c_int64 * 3)()
        check(lib().lb3d_gpu_batch_diagnostics(self._h, k, m, p, c), "batch_diagnostics")
        return np.array(m[:]), np.array(p[:]), np.array(c[:])

    def sanity(self, k):
        """The full sanity check of lattice k as a dict."""
        r = Sanity()
        check(lib().lb3d_gpu_batch_sanity(self._h, int(k), ctypes.byref(r)), "batch_sanity")
        return r.as_dict()

    def set_plan(self, sequential, nwg=2048, xcd_group=0):
        check(lib().lb3d_gpu_batch_set_plan(self._h, int(sequential), int(nwg), int(xcd_group)),
              "batch_set_plan")

    def set_collide_moments(self, tile_cols=62, lz=8, xcd_group=16, tile_threads=256):
        """The collide-moments plan (FP32 ternary): the collision gathers the
        next step's moments, populations are read once per step; tile_cols = 0
        turns it off."""
        check(lib().lb3d_gpu_batch_set_collide_moments(self._h, int(tile_cols), int(tile_threads), int(lz),
                                                       int(xcd_group)), "batch_set_collide_moments")

    def set_groups(self, group_bytes):
        """Run the persistent plan in groups of lattices of about group_bytes
        of populations each (0: one group)."""
        check(lib().lb3d_gpu_batch_set_groups(self._h, int(group_bytes)), "batch_set_groups")

    def autotune(self, trial_steps=3):
        buf = ctypes.create_string_buffer(32768)
        check(lib().lb3d_gpu_batch_autotune(self._h, int(trial_steps), buf, len(buf)), "batch_autotune")
        return buf.value.decode()

    def profile_steps(self, nsteps):
        ph = (ctypes.c_double * 3)()
        mps = ctypes.c_double()
        check(lib().lb3d_gpu_batch_profile_steps(self._h, int(nsteps), ph, ctypes.byref(mps)),
              "batch_profile_steps")
        return dict(moments=ph[0], dipole=ph[1], collide=ph[2]), mps.value
