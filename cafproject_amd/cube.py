"""ctypes binding of libcube_hip.so with the reference's subroutine interface.

The reference drives its hot path through parameterless subroutines that
mutate module state (CUBEnu/work/main/main.f90:102-118):

    call timestep; call update_x; call buffer_grid; call buffer_x
    if (Extended_pp_force) call ext_pp_force
    call particle_mesh; call buffer_v

`Cube` keeps that shape: the state lives on the GPU inside the context, and
each method is one of those subroutines with its inputs (dt, a_mid, sigma_vi,
...) passed explicitly instead of read from module globals.  There is no CPU
fallback: constructing a `Cube` without the built library or without a GPU
raises.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "libcube_hip.so")

ERR_NAMES = {
    1: "TILE_OVERFLOW", 2: "IMAGE_OVERFLOW", 3: "PP_OVERFLOW", 4: "CHECKSUM", 5: "COUNT",
    6: "DISPLACEMENT", 10: "HIP", 11: "ALLOC", 12: "INVALID", 13: "STATE",
}
FIELD = dict(rho_f=0, force_f=1, rho_c=2, force_c=3, kern_f=4, kern_c=5, phi_f=6)


class CubeError(RuntimeError):
    def __init__(self, code, where, msg):
        super().__init__(f"{where}: CUBE_ERR_{ERR_NAMES.get(code, code)} ({msg})")
        self.code = code


class Params(ctypes.Structure):
    _fields_ = [("nc", ctypes.c_int64), ("nnt", ctypes.c_int64), ("ncb", ctypes.c_int64),
                ("np_nc", ctypes.c_int64), ("izipx", ctypes.c_int32), ("izipv", ctypes.c_int32),
                ("has_pid", ctypes.c_int32), ("pp_on", ctypes.c_int32), ("pp_range", ctypes.c_int32),
                ("rsoft", ctypes.c_float), ("image_buffer", ctypes.c_float),
                ("tile_buffer", ctypes.c_float), ("device", ctypes.c_int32),
                ("mem_layout", ctypes.c_int32)]


class VStats(ctypes.Structure):
    _fields_ = [("sum_c", ctypes.c_double), ("sum_r", ctypes.c_double), ("sum_s", ctypes.c_double),
                ("std_vsim", ctypes.c_double), ("std_vsim_c", ctypes.c_double),
                ("std_vsim_res", ctypes.c_double), ("sigma_vi_new", ctypes.c_float),
                ("nlayer", ctypes.c_int32), ("nplocal", ctypes.c_int64), ("lost", ctypes.c_int64),
                ("overhead_tile", ctypes.c_double)]


class PMOut(ctypes.Structure):
    _fields_ = [("f2_max_fine", ctypes.c_float), ("f2_max_coarse", ctypes.c_float),
                ("dt_fine", ctypes.c_float), ("dt_coarse", ctypes.c_float),
                ("dt_vmax", ctypes.c_float), ("vz_max", ctypes.c_float),
                ("vmax", ctypes.c_float * 3), ("reserved", ctypes.c_float),
                ("testrhof", ctypes.c_double), ("testrhoc", ctypes.c_double)]


class ICParams(ctypes.Structure):
    _fields_ = [("seed", ctypes.c_uint64), ("box", ctypes.c_double), ("vfactor", ctypes.c_double),
                ("sigma_vi", ctypes.c_float), ("ntab", ctypes.c_int32), ("lnk0", ctypes.c_double),
                ("dlnk", ctypes.c_double), ("lnp", ctypes.c_void_p)]


def _struct_dict(s):
    out = {}
    for name, _ in s._fields_:
        v = getattr(s, name)
        out[name] = list(v) if hasattr(v, "__len__") else v
    return out


_LIB = None
EXPORTS = ["cube_ctx_create", "cube_ctx_destroy", "cube_strerror", "cube_ctx_info", "cube_init_kernels",
           "cube_set_state", "cube_get_state", "cube_update_x", "cube_buffer_grid", "cube_buffer_x",
           "cube_pp_force", "cube_particle_mesh", "cube_buffer_v", "cube_pm_fine_deposit",
           "cube_pm_fine_force", "cube_pm_fine_kick", "cube_pm_coarse_deposit", "cube_pm_coarse_force",
           "cube_pm_coarse_kick", "cube_get_field", "cube_set_field", "cube_sync", "cube_host_atan",
           "cube_profile", "cube_timers", "cube_ic_zeldovich", "cube_set_state_xv", "cube_power",
           "cube_power_noalloc", "cube_pm_fine_potential"]
TIMERS = ["drift", "vstats", "pp", "fine_deposit", "fine_fft", "fine_green", "fine_invy", "fine_c2r",
          "fine_kick", "coarse_deposit", "coarse_force", "coarse_kick", "checks"]


def lib():
    """Load the HIP library (raises if it has not been built)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIBPATH):
            raise RuntimeError(f"{LIBPATH} missing: build it with `python -m cafproject_amd.build`")
        # torch ships its own HIP and HSA runtimes under the same soname.
        # Loaded first, they also serve this library (one runtime per
        # process); loaded after it, they start a second runtime that finds
        # no device.  So torch, when installed, is imported before the load.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = ctypes.CDLL(LIBPATH)
        P, i64, f32, i32 = ctypes.c_void_p, ctypes.c_int64, ctypes.c_float, ctypes.c_int
        L.cube_ctx_create.argtypes = [ctypes.POINTER(Params), ctypes.POINTER(P)]
        L.cube_ctx_destroy.argtypes = [P]
        L.cube_ctx_destroy.restype = None
        L.cube_strerror.argtypes = [i32]
        L.cube_strerror.restype = ctypes.c_char_p
        L.cube_ctx_info.argtypes = [P, P]
        L.cube_init_kernels.argtypes = [P, P, P]
        L.cube_set_state.argtypes = [P, i64, i64, P, P, P, P, P]
        L.cube_get_state.argtypes = [P, P, P, P, P, P, P]
        L.cube_update_x.argtypes = [P, f32, f32, f32, f32, ctypes.POINTER(VStats)]
        for n in ("cube_buffer_grid", "cube_buffer_x", "cube_buffer_v", "cube_sync"):
            getattr(L, n).argtypes = [P]
        L.cube_pp_force.argtypes = [P, f32, f32, f32, f32, P]
        L.cube_power.argtypes = [P, ctypes.c_int32, ctypes.c_int32, P, P, P]
        L.cube_power_noalloc.argtypes = [P, ctypes.c_int32, ctypes.c_int32, P, P, P]
        L.cube_particle_mesh.argtypes = [P, f32, f32, f32, f32, f32, ctypes.POINTER(PMOut)]
        L.cube_pm_fine_deposit.argtypes = [P, f32, P]
        L.cube_pm_fine_force.argtypes = [P, P]
        L.cube_pm_fine_potential.argtypes = [P, P]
        L.cube_pm_fine_kick.argtypes = [P, f32, f32, f32, f32]
        L.cube_pm_coarse_deposit.argtypes = [P, f32, P]
        L.cube_pm_coarse_force.argtypes = [P, P]
        L.cube_pm_coarse_kick.argtypes = [P, f32, f32, f32, P]
        L.cube_get_field.argtypes = [P, i32, P, ctypes.c_size_t]
        L.cube_set_field.argtypes = [P, i32, P, ctypes.c_size_t]
        L.cube_host_atan.argtypes = [ctypes.c_double]
        L.cube_host_atan.restype = ctypes.c_double
        L.cube_ic_zeldovich.argtypes = [P, ctypes.POINTER(ICParams), P]
        L.cube_set_state_xv.argtypes = [P, i64, P, P, f32, P]
        L.cube_profile.argtypes = [P, i32]
        L.cube_timers.argtypes = [P, P, P, i32]
        _LIB = L
    return _LIB


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def load_potential_table(datadir=None):
    """The fine potential kernel table CUBEnu/kernels/wg1.ascii (16^3 lines
    `i j k value`, i fastest) as 4096 float32 values in (i,j,k) order."""
    d = datadir or os.path.join(HERE, "data")
    t = np.loadtxt(os.path.join(d, "wg1.ascii"))
    idx = t[:, :3].astype(int)
    want = np.stack(np.meshgrid(np.arange(1, 17), np.arange(1, 17), np.arange(1, 17), indexing="ij"),
                    -1)[..., ::-1].reshape(-1, 3)
    if not np.array_equal(idx, want):
        raise ValueError("wg1.ascii: index mismatch")
    return np.ascontiguousarray(t[:, 3], np.float32)


def load_tables(datadir=None):
    """The kernel tables CUBEnu/kernels/wfxyzf.3.ascii (16^3 lines
    `i j k fx fy fz`) and wfxyzc.2.ascii (4^3 lines, format (3i4,3e16.8)),
    as flat float32 arrays in (i,j,k) order with 3 components per entry, the
    layout cube_init_kernels takes (W/kernel_f.f90:12-21,
    W/kernel_c.f90:31-38)."""
    d = datadir or os.path.join(HERE, "data")
    f = np.loadtxt(os.path.join(d, "wfxyzf.3.ascii"))
    idx = f[:, :3].astype(int)
    n = 0
    for k in range(1, 17):
        for j in range(1, 17):
            for i in range(1, 17):
                if tuple(idx[n]) != (i, j, k):
                    raise ValueError("wfxyzf.3.ascii: index mismatch")  # W/kernel_f.f90:19
                n += 1
    fk = np.ascontiguousarray(f[:, 3:6], np.float32).reshape(-1)
    lines = open(os.path.join(d, "wfxyzc.2.ascii")).read().splitlines()
    ck = np.array([[float(s[12 + 16 * q:28 + 16 * q]) for q in range(3)] for s in lines[:64]], np.float32)
    return fk, np.ascontiguousarray(ck.reshape(-1))


class Cube:
    """One image (the whole periodic box) of the CUBE state on one GPU."""

    def __init__(self, nc, nnt, ncb=6, np_nc=2, izipx=2, izipv=2, has_pid=True, pp_on=False,
                 pp_range=1, rsoft=0.3, image_buffer=1.5, tile_buffer=2.5, device=0, mem_layout=0):
        L = lib()
        self.p = Params(nc, nnt, ncb, np_nc, izipx, izipv, int(has_pid), int(pp_on), pp_range, rsoft,
                        image_buffer, tile_buffer, device, mem_layout)
        h = ctypes.c_void_p()
        self._check(L.cube_ctx_create(ctypes.byref(self.p), ctypes.byref(h)), "cube_ctx_create")
        self._h = h
        self.nc, self.nnt, self.ncb, self.izipx, self.izipv = nc, nnt, ncb, izipx, izipv
        self.has_pid, self.pp_on = has_pid, pp_on
        info = self.info()
        self.nt, self.nf = info[0], info[1]
        self.np_image, self.np_image_max, self.np_tile_max = info[3], info[4], info[5]

    @staticmethod
    def _check(code, where):
        if code != 0:
            raise CubeError(code, where, lib().cube_strerror(code).decode())

    def close(self):
        if getattr(self, "_h", None):
            lib().cube_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self):
        o = np.zeros(16, np.int64)
        self._check(lib().cube_ctx_info(self._h, _ptr(o)), "cube_ctx_info")
        return [int(v) for v in o]

    @property
    def nplocal(self):
        return self.info()[8]

    # -- kernel_f / kernel_c (W/initialize.f90:37-38) ---------------------------
    def init_kernels(self, fk=None, ck=None):
        if fk is None:
            fk, ck = load_tables()
        self._check(lib().cube_init_kernels(self._h, _ptr(fk), _ptr(ck)), "init_kernels")

    # -- particle_initialization / checkpoint -----------------------------------
    def set_state(self, xp, vp, pid, rhoc, vfield, npglobal=None):
        xt = np.int8 if self.izipx == 1 else np.int16
        vt = np.int8 if self.izipv == 1 else np.int16
        xp = np.ascontiguousarray(xp, xt)
        vp = np.ascontiguousarray(vp, vt)
        pid = None if pid is None else np.ascontiguousarray(pid, np.int32)
        rhoc = np.ascontiguousarray(rhoc, np.int32)
        vfield = np.ascontiguousarray(vfield, np.float32)
        n = xp.shape[0]
        self._check(lib().cube_set_state(self._h, n, n if npglobal is None else npglobal, _ptr(xp),
                                         _ptr(vp), _ptr(pid), _ptr(rhoc), _ptr(vfield)), "set_state")

    def set_state_xv(self, x, v, sigma_vi=0.0):
        """Count-sort continuous particles (global coarse units) into the
        disjoint state; returns the sigma_vi used."""
        x = np.ascontiguousarray(x, np.float64)
        v = np.ascontiguousarray(v, np.float32)
        s = ctypes.c_float()
        self._check(lib().cube_set_state_xv(self._h, x.shape[0], _ptr(x), _ptr(v), sigma_vi, ctypes.byref(s)),
                    "set_state_xv")
        return s.value

    def make_ics(self, box, z_i, seed, cosmo=None, sigma_vi=0.0, ntab=4096):
        """Zel'dovich ICs on the GPU (cube_ic_zeldovich); returns (a, sigma_vi)."""
        from cafproject_amd import ics
        cosmo = cosmo or ics.Cosmology()
        a = 1.0 / (1.0 + z_i)
        ng = self.nc * self.p.np_nc
        kmin = 2 * np.pi / box * 0.25
        kmax = np.sqrt(3.0) * np.pi * ng / box * 1.05
        lnk = np.linspace(np.log(kmin), np.log(kmax), ntab)
        lnp = np.ascontiguousarray(np.log(np.maximum(cosmo.pk(np.exp(lnk), a), 1e-300)), np.float64)
        ip = ICParams(seed, box, ics.vfactor(a, cosmo.omega_m, cosmo.omega_l), sigma_vi, ntab, float(lnk[0]),
                      float(lnk[1] - lnk[0]), lnp.ctypes.data_as(ctypes.c_void_p))
        s = ctypes.c_float()
        self._check(lib().cube_ic_zeldovich(self._h, ctypes.byref(ip), ctypes.byref(s)), "make_ics")
        return a, s.value

    def get_state(self, particles=True):
        """The disjoint state (cube_get_state); particles=False fetches only
        the cell arrays rhoc and vfield."""
        n = self.nplocal
        xt = np.int8 if self.izipx == 1 else np.int16
        vt = np.int8 if self.izipv == 1 else np.int16
        xp = np.zeros((n, 3), xt) if particles else None
        vp = np.zeros((n, 3), vt) if particles else None
        pid = np.zeros(n, np.int32) if self.has_pid and particles else None
        rhoc = np.zeros(self.nc ** 3, np.int32)
        vf = np.zeros((self.nc ** 3, 3), np.float32)
        nn = ctypes.c_int64()
        self._check(lib().cube_get_state(self._h, ctypes.byref(nn), _ptr(xp), _ptr(vp), _ptr(pid),
                                         _ptr(rhoc), _ptr(vf)), "get_state")
        return dict(xp=xp, vp=vp, pid=pid, rhoc=rhoc, vfield=vf, nplocal=nn.value)

    # -- the step (W/main.f90:102-118) -------------------------------------------
    def update_x(self, dt_old, dt, vz_max, sigma_vi):
        s = VStats()
        self._check(lib().cube_update_x(self._h, dt_old, dt, vz_max, sigma_vi, ctypes.byref(s)), "update_x")
        return _struct_dict(s)

    def buffer_grid(self):
        self._check(lib().cube_buffer_grid(self._h), "buffer_grid")

    def buffer_x(self):
        self._check(lib().cube_buffer_x(self._h), "buffer_x")

    def buffer_v(self):
        self._check(lib().cube_buffer_v(self._h), "buffer_v")

    def ext_pp_force(self, a_mid, dt, sigma_vi, mass_p):
        f2 = ctypes.c_float()
        self._check(lib().cube_pp_force(self._h, a_mid, dt, sigma_vi, mass_p, ctypes.byref(f2)),
                    "ext_pp_force")
        return f2.value

    def particle_mesh(self, a_mid, dt, sigma_vi, sigma_vi_new, mass_p):
        o = PMOut()
        self._check(lib().cube_particle_mesh(self._h, a_mid, dt, sigma_vi, sigma_vi_new, mass_p,
                                             ctypes.byref(o)), "particle_mesh")
        return _struct_dict(o)

    # -- stages of particle_mesh --------------------------------------------------
    def fine_deposit(self, mass_p):
        t = ctypes.c_double()
        self._check(lib().cube_pm_fine_deposit(self._h, mass_p, ctypes.byref(t)), "fine_deposit")
        return t.value

    def fine_potential(self, wg1=None):
        """Fine potential of the current fine density (cube_pm_fine_potential),
        returned as [z][y][x]."""
        if wg1 is None:
            wg1 = load_potential_table()
        self._check(lib().cube_pm_fine_potential(self._h, _ptr(wg1)), "fine_potential")
        return self.get_field("phi_f")

    def fine_force(self):
        f = ctypes.c_float()
        self._check(lib().cube_pm_fine_force(self._h, ctypes.byref(f)), "fine_force")
        return f.value

    def fine_kick(self, a_mid, dt, sigma_vi, sigma_vi_new):
        self._check(lib().cube_pm_fine_kick(self._h, a_mid, dt, sigma_vi, sigma_vi_new), "fine_kick")

    def coarse_deposit(self, mass_p):
        t = ctypes.c_double()
        self._check(lib().cube_pm_coarse_deposit(self._h, mass_p, ctypes.byref(t)), "coarse_deposit")
        return t.value

    def coarse_force(self):
        f = ctypes.c_float()
        self._check(lib().cube_pm_coarse_force(self._h, ctypes.byref(f)), "coarse_force")
        return f.value

    def coarse_kick(self, a_mid, dt, sigma_vi):
        v = np.zeros(3, np.float32)
        self._check(lib().cube_pm_coarse_kick(self._h, a_mid, dt, sigma_vi, _ptr(v)), "coarse_kick")
        return v

    # -- fields ------------------------------------------------------------------
    def field_shape(self, name):
        nf, nc = self.nf, self.nc
        ho = nf // 2 + 1
        return dict(rho_f=(nf, nf, nf), force_f=(nf, nf, nf, 3), rho_c=(nc, nc, nc),
                    force_c=(nc, nc, nc, 3), kern_f=(3, ho, ho, ho),
                    kern_c=(3, nc, nc, nc // 2 + 1), phi_f=(nf, nf, nf))[name]

    def get_field(self, name):
        a = np.zeros(self.field_shape(name), np.float32)
        self._check(lib().cube_get_field(self._h, FIELD[name], _ptr(a), a.nbytes), "get_field")
        return a

    def set_field(self, name, arr):
        a = np.ascontiguousarray(arr, np.float32)
        if a.shape != self.field_shape(name):
            raise ValueError(f"{name}: shape {a.shape} != {self.field_shape(name)}")
        self._check(lib().cube_set_field(self._h, FIELD[name], _ptr(a), a.nbytes), "set_field")

    def profile(self, on=True):
        self._check(lib().cube_profile(self._h, int(on)), "profile")

    def timers(self, reset=False):
        ms = np.zeros(len(TIMERS), np.float64)
        calls = np.zeros(len(TIMERS), np.int64)
        self._check(lib().cube_timers(self._h, _ptr(ms), _ptr(calls), int(reset)), "timers")
        return {k: (float(ms[i]), int(calls[i])) for i, k in enumerate(TIMERS)}

    def sync(self):
        self._check(lib().cube_sync(self._h), "sync")

    def power(self, ng, box, noalloc=False):
        """P(k) of the current state on an ng^3 mesh (cube_power, or
        cube_power_noalloc with noalloc=True): returns (k in h/Mpc, Delta^2,
        modes) per linear bin 1..nint(ng/2 sqrt 3), as
        cafproject_amd.power.state_power does on the host."""
        nbin = int(round(ng // 2 * np.sqrt(3.0)))
        ks, ps = np.zeros(nbin), np.zeros(nbin)
        cnt = np.zeros(nbin, np.int64)
        fn = lib().cube_power_noalloc if noalloc else lib().cube_power
        self._check(fn(self._h, ng, nbin, _ptr(ks), _ptr(ps), _ptr(cnt)), "power")
        with np.errstate(invalid="ignore", divide="ignore"):
            return ks / cnt * 2 * np.pi / box, ps / cnt, cnt
