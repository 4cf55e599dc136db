"""ctypes binding of the CPU oracle (liblb3d_oracle.so).

TEST INFRASTRUCTURE ONLY: the checker that tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg compare the HIP path against.  The product library
never imports this module.

The oracle works on the reference's own array-of-structures lattice `whole_N`
(code/lb3d_lattice.F90:35-83, 237-287): shape (nz+2, ny+2, nx+2, site_doubles)
in C order, i.e. Fortran order with x fastest, one-site halo.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblb3d_oracle.so")


class OrParams(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("nz", ctypes.c_int32),
        ("ncomp", ctypes.c_int32), ("bdist", ctypes.c_int32),
        ("lees_edwards", ctypes.c_int32), ("restore", ctypes.c_int32),
        ("pad0", ctypes.c_int32),
        ("amass", ctypes.c_double * 3), ("tau", ctypes.c_double * 3),
        ("tau_d", ctypes.c_double), ("beta", ctypes.c_double), ("d_0", ctypes.c_double),
        ("g_br", ctypes.c_double), ("g_bs", ctypes.c_double), ("g_ss", ctypes.c_double),
        ("shear_u", ctypes.c_double), ("shear_omega", ctypes.c_double),
        ("shear_sum", ctypes.c_double),
        ("g_accn", ctypes.c_double * 3),
        ("g_accn_min", ctypes.c_int32 * 3), ("g_accn_max", ctypes.c_int32 * 3),
        ("g_wr", ctypes.c_double), ("g_wb", ctypes.c_double),
        ("tau_wr", ctypes.c_double), ("tau_wb", ctypes.c_double),
        ("rock_colour_r", ctypes.c_double), ("rock_colour_b", ctypes.c_double),
        ("inv_fluid", ctypes.c_int32), ("inv_type", ctypes.c_int32),
        ("inv_flux", ctypes.c_double * 3),
    ]


def build():
    """Compile the oracle in place (gcc, a few seconds)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_lib = None


class OrSanity(ctypes.Structure):
    """or_sanity_report of lb3d_oracle.h."""
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


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        vp, dp, ip = ctypes.c_void_p, ctypes.POINTER(ctypes.c_double), ctypes.c_int
        L.or_create.restype = vp
        L.or_create.argtypes = [ctypes.POINTER(OrParams)]
        L.or_destroy.argtypes = [vp]
        L.or_sites.restype = dp
        L.or_sites.argtypes = [vp]
        L.or_site_doubles.argtypes = [vp]
        L.or_nt.argtypes = [vp]
        L.or_set_nt.argtypes = [vp, ip]
        for name in ("or_bc_init", "or_advection", "or_collision", "or_time_step"):
            getattr(L, name).argtypes = [vp]
        L.or_bc_apply.argtypes = [vp, ip]
        L.or_run.argtypes = [vp, ip]
        L.or_le_state.argtypes = [vp, dp, ctypes.POINTER(ctypes.c_int), dp, dp]
        L.or_sanity.argtypes = [vp, dp, dp, ctypes.POINTER(ctypes.c_int64)]
        for name in ("or_dump_od", "or_dump_wd", "or_dump_sur", "or_dump_dir",
                     "or_dump_vel", "or_dump_arr"):
            getattr(L, name).argtypes = [vp, dp]
        L.or_boltz_dist.argtypes = [ip] + [ctypes.c_double] * 6 + [dp]
        L.or_boltz_dist_dd.argtypes = [ip] + [ctypes.c_double] * 9 + [dp]
        L.or_zou_he_site.argtypes = [dp, ctypes.c_double, ip, ip]
        L.or_sc_force.argtypes = [vp, ip, ip, ip, ctypes.POINTER(ctypes.c_double * 9)]
        L.or_init_noise.argtypes = [vp, ctypes.c_uint64, dp, dp, ip]
        L.or_hash_uniform.restype = ctypes.c_double
        L.or_hash_uniform.argtypes = [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint64]
        L.or_set_threads.argtypes = [ip]
        L.or_halo_refresh.argtypes = [vp]
        L.or_stress.argtypes = [vp, ip, ip, dp]
        L.or_sanity_full.argtypes = [vp, ctypes.POINTER(OrSanity)]
        L.or_init_rock.argtypes = [vp, dp, ctypes.c_double]
        L.or_threads.restype = ip
        _lib = L
    return _lib


def set_threads(n):
    """Worker threads of the oracle's site loops (results do not depend on it)."""
    lib().or_set_threads(int(n))


def threads():
    return int(lib().or_threads())


DEFAULTS = dict(
    ncomp=2, bdist=2, lees_edwards=0, restore=0,
    amass=(1.0, 1.0, 1.0), tau=(1.0, 1.0, 1.0), tau_d=1.0, beta=10.0, d_0=1.0,
    g_br=0.08, g_bs=-0.006, g_ss=-0.003, shear_u=0.0, shear_omega=0.0, shear_sum=0.0,
    g_accn=(0.0, 0.0, 0.0), g_accn_min=(0, 0, 0), g_accn_max=None,
    g_wr=0.0, g_wb=0.0, tau_wr=1.0, tau_wb=1.0, rock_colour_r=0.0, rock_colour_b=0.0,
    inv_fluid=-1, inv_type=0, inv_flux=(0.0, 0.0, 0.0),
)


def make_params(nx, ny, nz, **kw):
    d = dict(DEFAULTS)
    d.update(kw)
    p = OrParams()
    p.nx, p.ny, p.nz = nx, ny, nz
    for k in ("ncomp", "bdist", "lees_edwards", "restore", "inv_fluid", "inv_type"):
        setattr(p, k, int(d[k]))
    p.inv_flux[:] = [float(v) for v in d["inv_flux"]]
    for k in ("tau_d", "beta", "d_0", "g_br", "g_bs", "g_ss", "shear_u", "shear_omega",
              "shear_sum", "g_wr", "g_wb", "tau_wr", "tau_wb", "rock_colour_r", "rock_colour_b"):
        setattr(p, k, float(d[k]))
    p.amass[:] = [float(v) for v in d["amass"]]
    p.tau[:] = [float(v) for v in d["tau"]]
    p.g_accn[:] = [float(v) for v in d["g_accn"]]
    p.g_accn_min[:] = [int(v) for v in d["g_accn_min"]]
    gmax = d["g_accn_max"] or (nx, ny, nz)  # code/lb3d_config.F90:931-933
    p.g_accn_max[:] = [int(v) for v in gmax]
    return p


class Oracle:
    """One lattice held by the CPU oracle."""

    def __init__(self, nx, ny, nz, **kw):
        self.nx, self.ny, self.nz = nx, ny, nz
        self.params = make_params(nx, ny, nz, **kw)
        self.ncomp = self.params.ncomp
        self._h = lib().or_create(ctypes.byref(self.params))
        if not self._h:
            raise MemoryError("or_create failed")
        self.sd = lib().or_site_doubles(self._h)

    def close(self):
        if self._h:
            lib().or_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def N(self):
        """whole_N as a numpy view, shape (nz+2, ny+2, nx+2, site_doubles)."""
        ptr = lib().or_sites(self._h)
        n = (self.nz + 2) * (self.ny + 2) * (self.nx + 2) * self.sd
        return np.ctypeslib.as_array(ptr, shape=(n,)).reshape(
            self.nz + 2, self.ny + 2, self.nx + 2, self.sd)

    @property
    def nt(self):
        return lib().or_nt(self._h)

    @nt.setter
    def nt(self, v):
        lib().or_set_nt(self._h, int(v))

    def bc_init(self):
        lib().or_bc_init(self._h)

    def bc_apply(self, stage):
        lib().or_bc_apply(self._h, int(stage))

    def advection(self):
        lib().or_advection(self._h)

    def collision(self):
        lib().or_collision(self._h)

    def step(self):
        lib().or_time_step(self._h)

    def run(self, n):
        lib().or_run(self._h, int(n))

    def le_state(self):
        s, f, u = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        lo = ctypes.c_int()
        lib().or_le_state(self._h, ctypes.byref(s), ctypes.byref(lo), ctypes.byref(f),
                          ctypes.byref(u))
        return s.value, lo.value, f.value, u.value

    def init_rock(self, obs, threshold=0.0):
        """Obstacles from a per-site field obs (nz, ny, nx): rock where obs > threshold
        (lbe_init_rock, code/lb3d_ic_rock.F90:50-189)."""
        a = np.ascontiguousarray(obs, dtype=np.float64)
        assert a.shape == (self.nz, self.ny, self.nx)
        lib().or_init_rock(self._h, a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), float(threshold))

    def halo_refresh(self):
        """Rebuild the halo as the pre-advection pass would, offset not advanced."""
        lib().or_halo_refresh(self._h)

    def stress(self, minx=1, maxx=None):
        """Halved momentum transfer through x = minx and x = maxx
        (code/lb3d_io_stress.F90:38-173) with the halo as it stands; returns
        (out0, out1, mean shear stress)."""
        out = (ctypes.c_double * 2)()
        lib().or_stress(self._h, int(minx), int(self.nx if maxx is None else maxx), out)
        return out[0], out[1], (out[1] - out[0]) / (self.ny * self.nz)

    def sanity_full(self):
        """The full sanity check (code/lb3d_sanity.F90:79-384) as a dict."""
        r = OrSanity()
        lib().or_sanity_full(self._h, ctypes.byref(r))
        return r.as_dict()

    def sanity(self):
        m = (ctypes.c_double * 3)()
        p = (ctypes.c_double * 3)()
        c = (ctypes.c_int64 * 3)()
        lib().or_sanity(self._h, m, p, c)
        return np.array(m[:]), np.array(p[:]), np.array(c[:])

    def dump(self, name):
        n = self.nx * self.ny * self.nz
        vec = name in ("dir", "arr")
        out = np.zeros(n * (3 if vec else 1))
        getattr(lib(), "or_dump_" + name)(self._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        return out.reshape(self.nz, self.ny, self.nx, 3) if vec else out.reshape(self.nz, self.ny, self.nx)

    def sc_force(self, x, y, z):
        f = (ctypes.c_double * 9)()
        lib().or_sc_force(self._h, x, y, z, ctypes.byref(f))
        return np.array(f[:]).reshape(3, 3)

    def init_noise(self, seed, rho_mean=(0.7, 0.7, 0.7), amp=(0.01, 0.01, 0.0), dipoles=True):
        rm = (ctypes.c_double * 3)(*rho_mean)
        am = (ctypes.c_double * 3)(*amp)
        lib().or_init_noise(self._h, ctypes.c_uint64(seed), rm, am, 1 if dipoles else 0)

    # --- state conversion -------------------------------------------------
    def interior(self):
        return self.N[1:-1, 1:-1, 1:-1, :]

    def populations(self):
        """Interior populations, shape (ncomp, 19, nz, ny, nx)."""
        I = self.interior()
        return np.stack([np.moveaxis(I[..., 24 * c:24 * c + 19], -1, 0)
                         for c in range(self.ncomp)])

    def dipole(self):
        """Interior db, shape (3, nz, ny, nx) (ternary only)."""
        return np.moveaxis(self.interior()[..., 75:78], -1, 0)

    def load_checkpoint_array(self, ck):
        """Fill the interior from a reference XDR checkpoint record array,
        shape (nx, ny, nz, 39|64) (code/lb3d_io_checkpoint.F90:790-873)."""
        I = self.interior()
        T = np.transpose(ck, (2, 1, 0, 3))  # -> (z, y, x, rec)
        I[...] = 0.0
        I[..., 0:19] = T[..., 0:19]
        I[..., 24:43] = T[..., 19:38]
        if self.ncomp == 3:
            I[..., 48:67] = T[..., 38:57]
            I[..., 72:75] = T[..., 57:60]
            I[..., 75:78] = T[..., 60:63]
            I[..., 78] = T[..., 63]
        else:
            I[..., 48] = T[..., 38]


def hash_uniform(seed, stream, site):
    return lib().or_hash_uniform(ctypes.c_uint64(seed), ctypes.c_uint32(stream), ctypes.c_uint64(site))
