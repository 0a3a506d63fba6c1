"""CPU parity oracle for the CUBE PM+PP step -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import
this module, and only as the checker / CPU baseline.  The product library in
cafproject_amd/ never imports it.

The particle loops (update_xp, buffers, deposits, kicks, PP) are the C
restatement in cube_oracle.c; this module adds the FFT parts of the reference
with scipy.fft in float32 (standing in for FFTW single precision, which the
reference links; W/cubefft.f90:15-25, W/pencil_fft.f90:31-60) and, for error
statements, a float64 twin.

Reference line citations are relative to the reference repository
(W/ = CUBEnu/work/main/).
"""
import ctypes
import os

import numpy as np
import scipy.fft as sfft

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
FFT_WORKERS = os.cpu_count()  # threads of the scipy FFTs (the CPU baseline sets 1)
PI_F = np.float32(4.0) * np.arctan(np.float32(1.0))  # pi=4*atan(1.), real(4)


def lib():
    """Load oracle/libcube_oracle.so (built by `make -C oracle`)."""
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "libcube_oracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle library missing: run `make -C oracle`")
        L = ctypes.CDLL(path)
        P = ctypes.c_void_p
        i64, f32, f64, i32 = ctypes.c_int64, ctypes.c_float, ctypes.c_double, ctypes.c_int32
        L.oc_alloc.restype = P
        L.oc_free.argtypes = [P]
        L.oc_create.argtypes = [P, i64, i64, i64, i64, i64, i64, i64, i64, f32, f32, f32, i64]
        L.oc_info.argtypes = [P, P]
        L.oc_set_npglobal.argtypes = [P, i64]
        L.oc_set_disjoint.argtypes = [P, i64, P, P, P, P, P]
        L.oc_get_disjoint.argtypes = [P, P, P, P, P, P]
        L.oc_spine_image.argtypes = [P]
        L.oc_buffer_grid.argtypes = [P]
        L.oc_buffer_xvp.argtypes = [P, ctypes.c_int, ctypes.c_int]
        L.oc_update_xp.argtypes = [P, f32, f32, f32, f32, P, P]
        L.oc_fine_deposit.argtypes = [P, i64, i64, i64, f32, P]
        L.oc_fine_deposit.restype = f64
        L.oc_fine_kick.argtypes = [P, i64, i64, i64, P, f32, f32, f32, f32]
        L.oc_coarse_deposit.argtypes = [P, f32, P]
        L.oc_coarse_deposit.restype = f64
        L.oc_coarse_kick.argtypes = [P, P, f32, f32, f32, P]
        L.oc_pp_force.argtypes = [P, f32, f32, f32, f32, P]
        L.oc_decode_x.argtypes = [i64, i32]
        L.oc_decode_x.restype = f64
        L.oc_decode_v.argtypes = [i64, i32, f32]
        L.oc_decode_v.restype = f64
        L.oc_encode_v.argtypes = [i64, f64, f32]
        L.oc_encode_v.restype = i32
        L.oc_encode_x_drift.argtypes = [i64, i32, f64]
        L.oc_encode_x_drift.restype = i32
        L.oc_tanf_code.argtypes = [i64, i32]
        L.oc_tanf_code.restype = f32
        L.oc_state_from_xv.argtypes = [P, i64, P, P, P, f32]
        L.oc_get_overhead.argtypes = [P, ctypes.c_int]
        L.oc_get_overhead.restype = f64
        L.oc_get_pp_ftemp.argtypes = [P]
        L.oc_get_pp_ftemp.restype = f32
        L.oc_set_pp_ftemp.argtypes = [P, f32]
        for name in ("oc_xp_ptr", "oc_vp_ptr", "oc_pid_ptr", "oc_rhoc_ptr", "oc_vfield_ptr"):
            getattr(L, name).argtypes = [P]
            getattr(L, name).restype = P
        L.oc_spine_ptr.argtypes = [P, ctypes.c_int]
        L.oc_spine_ptr.restype = P
        L.oc_expansion_n.argtypes = [i64, P, P, f32, f32, f32, P, P]
        L.oc_set_threads.argtypes = [ctypes.c_int]
        L.oc_green_mul.argtypes = [P, P, P, i64]
        L.oc_force_extract.argtypes = [P, i64, i64, i64, i64, P]
        L.oc_f2_max.argtypes = [P, i64]
        L.oc_f2_max.restype = f32
        L.oc_get_threads.restype = ctypes.c_int
        _LIB = L
    return _LIB


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


ERRORS = {1: "tile overflow", 2: "image overflow", 3: "pp overflow", 4: "checksum",
          5: "count mismatch", 6: "displacement beyond buffer", 11: "alloc", 12: "invalid"}


class OracleError(RuntimeError):
    def __init__(self, code, where):
        super().__init__(f"{where}: oracle status {code} ({ERRORS.get(code, '?')})")
        self.code = code


def _check(code, where):
    if code != 0:
        raise OracleError(code, where)


# --------------------------------------------------------------------------
# Green's functions (W/kernel_f.f90, W/kernel_c.f90)
# --------------------------------------------------------------------------
def load_tables(datadir=None):
    """Read wfxyzf.3.ascii (16^3 x 3, list-directed) and wfxyzc.2.ascii
    (4^3 x 3, format (3i4,3e16.8)) as W/kernel_f.f90:12-21 and
    W/kernel_c.f90:31-38 do.  Returns fk[i,j,k,d] and ck[d,i,j,k] (0-based
    offsets)."""
    if datadir is None:
        datadir = os.path.join(os.path.dirname(_HERE), "cafproject_amd", "data")
    fk = np.zeros((16, 16, 16, 3), np.float32)
    rows = np.loadtxt(os.path.join(datadir, "wfxyzf.3.ascii"))
    n = 0
    for k in range(16):
        for j in range(16):
            for i in range(16):
                r = rows[n]
                if (int(r[0]), int(r[1]), int(r[2])) != (i + 1, j + 1, k + 1):
                    raise ValueError("wfxyzf table index mismatch")  # W/kernel_f.f90:19
                fk[i, j, k] = r[3:6]
                n += 1
    ck = np.zeros((3, 4, 4, 4), np.float32)
    with open(os.path.join(datadir, "wfxyzc.2.ascii")) as f:
        lines = f.read().splitlines()
    n = 0
    for k in range(4):
        for j in range(4):
            for i in range(4):
                s = lines[n]
                vals = [float(s[12 + 16 * q:12 + 16 * (q + 1)]) for q in range(3)]
                ck[:, i, j, k] = vals
                n += 1
    return fk, ck


def kernel_f(nfe, fk, pp_on=False, pp_range=1, dtype=np.float32):
    """kern_f(nfe/2+1,nfe,nfe,3) of W/kernel_f.f90:25-37, returned as
    [d][kz][ky][kx].  The 16^3 table is placed at offsets 0..15, optionally
    the (pp_range+1)^3 corner zeroed, odd-mirrored per axis, r2c, imag."""
    out = np.zeros((3, nfe, nfe, nfe // 2 + 1), dtype)
    for d in range(3):
        rs = np.zeros((nfe, nfe, nfe), dtype)
        rs[:16, :16, :16] = fk[:, :, :, d].transpose(2, 1, 0)
        if pp_on:
            rs[:pp_range + 1, :pp_range + 1, :pp_range + 1] = 0
        m = [(-1 if ax == d else 1) for ax in range(3)]  # x, y, z
        rs[:, :, nfe - 15:] = m[0] * rs[:, :, 15:0:-1]
        rs[:, nfe - 15:, :] = m[1] * rs[:, 15:0:-1, :]
        rs[nfe - 15:, :, :] = m[2] * rs[15:0:-1, :, :]
        out[d] = _rfftn(rs).imag
    return out


def _rfftn(a):
    if a.dtype == np.float32:
        return sfft.rfftn(a, workers=FFT_WORKERS)
    return np.fft.rfftn(a)


def _irfftn_unnorm(c, n):
    if c.dtype == np.complex64:
        return sfft.irfftn(c, s=(n, n, n), norm="forward", workers=FFT_WORKERS)
    return np.fft.irfftn(c, s=(n, n, n), axes=(0, 1, 2), norm="forward")


def kernel_c(nc, ckt, ncell=4, dtype=np.float32, lrck=True, table=True):
    """kern_c(nc/2+1,nc,nc,3) of W/kernel_c.f90 with LRCKCORR, single image,
    returned as [d][kz][ky][kx].  lrck/table switch off the low-k correction
    (the #else branch, :326-332) or the 4^3 table octants (for the tests)."""
    f = dtype
    idx = (np.arange(1, nc + 1) + nc // 2 - 1) % nc - nc // 2
    rx = (ncell * idx).astype(f)
    X, Y, Z = np.meshgrid(rx, rx, rx, indexing="ij")  # [i][j][k]
    r = np.sqrt(X * X + Y * Y + Z * Z).astype(f)
    r3 = r * (r * r)
    with np.errstate(divide="ignore", invalid="ignore"):
        ck = np.stack([np.where(r == 0, f(0), -X / r3), np.where(r == 0, f(0), -Y / r3),
                       np.where(r == 0, f(0), -Z / r3)]).astype(f)      # [d][i][j][k]
    plain = ck.copy()
    t = ckt.astype(f)
    if not table:
        t = plain[:, :4, :4, :4].copy()
    R = slice(3, 0, -1)   # Fortran 4:2:-1 -> 0-based 3,2,1
    H = slice(nc - 3, nc)  # Fortran nc-2:nc
    Q = slice(0, 4)
    ck[:, Q, Q, Q] = t                                           # 1
    ck[0, H, Q, Q] = -t[0, R, :, :]; ck[1:3, H, Q, Q] = t[1:3, R, :, :]      # 5 x
    ck[0::2, Q, H, Q] = t[0::2, :, R, :]; ck[1, Q, H, Q] = -t[1, :, R, :]    # 3 y
    ck[0:2, Q, Q, H] = t[0:2, :, :, R]; ck[2, Q, Q, H] = -t[2, :, :, R]      # 2 z
    ck[0, Q, H, H] = t[0, :, R, R]; ck[1:3, Q, H, H] = -t[1:3, :, R, R]      # 4 yz
    ck[0::2, H, Q, H] = -t[0::2, R, :, R]; ck[1, H, Q, H] = t[1, R, :, R]    # 6 xz
    ck[0:2, H, H, Q] = -t[0:2, R, R, :]; ck[2, H, H, Q] = t[2, R, R, :]      # 7 xy
    ck[:, H, H, H] = -t[:, R, R, R]                                        # 8 xyz
    kern = np.zeros((3, nc, nc, nc // 2 + 1), f)
    # frequencies of cxyz(i,j,k): kx=mod(i+nc/2-1,nc)-nc/2 for i=1..nc/2+1
    kxi = ((np.arange(1, nc // 2 + 2) + nc // 2 - 1) % nc - nc // 2).astype(f)
    kyj = ((np.arange(1, nc + 1) + nc // 2 - 1) % nc - nc // 2).astype(f)
    KZ, KY, KX = np.meshgrid(kyj, kyj, kxi, indexing="ij")   # [kz][ky][kx]
    kr = np.sqrt(KX * KX + KY * KY + KZ * KZ).astype(f)
    pi = f(PI_F)
    sx = (f(2) * np.sin(pi * KX / f(nc))).astype(f)
    sy = (f(2) * np.sin(pi * KY / f(nc))).astype(f)
    sz = (f(2) * np.sin(pi * KZ / f(nc))).astype(f)
    ssum = ((sx * sx + sy * sy) + sz * sz).astype(f)
    for d, (kd, sd) in enumerate(((KX, sx), (KY, sy), (KZ, sz))):
        kc = _rfftn(np.ascontiguousarray(ck[d].transpose(2, 1, 0))).imag.astype(f)
        cu = _rfftn(np.ascontiguousarray(plain[d].transpose(2, 1, 0))).imag.astype(f)
        with np.errstate(divide="ignore", invalid="ignore"):
            corr = ((((kc * f(0.25)) * pi) * sd) / ssum) / cu
        kern[d] = np.where((kr > 8.0) | (kd == 0), kc, corr) if lrck else kc
    return kern


def apply_green(crho, kern_d):
    """rho_f(::2)=-crho(2::2)*kern, rho_f(2::2)=crho(::2)*kern (W/pm.f90:119-120)."""
    f = kern_d.dtype.type
    if f is np.float32 and crho.dtype == np.complex64 and crho.flags.c_contiguous and kern_d.flags.c_contiguous:
        out = np.empty_like(crho)
        lib().oc_green_mul(_p(crho), _p(kern_d), _p(out), crho.size)
        return out
    re = (-crho.imag * kern_d).astype(f)
    im = (crho.real * kern_d).astype(f)
    return (re + 1j * im).astype(np.complex64 if f is np.float32 else np.complex128)


def fine_force(rho, kern, nfe, nfb):
    """Fine force of one tile, W/pm.f90:112-126.  rho: [z][y][x] nfe^3.
    Returns force_f as [z][y][x][d] over nfb:nfe-nfb+1 and f2_max_fine."""
    f = kern.dtype.type
    crho = _rfftn(rho.astype(f))
    n2 = nfe - 2 * nfb + 2
    out = np.zeros((n2, n2, n2, 3), f)
    if f is np.float32:
        for d in range(3):
            r = np.ascontiguousarray(_irfftn_unnorm(apply_green(crho, kern[d]), nfe), np.float32)
            lib().oc_force_extract(_p(r), nfe, nfb - 1, n2, d, _p(out))
        return out, float(lib().oc_f2_max(_p(out), n2 ** 3))
    for d in range(3):
        r = _irfftn_unnorm(apply_green(crho, kern[d]), nfe).astype(f)
        r = r / f(nfe)
        r = r / f(nfe)
        r = r / f(nfe)
        out[..., d] = r[nfb - 1:nfe - nfb + 1, nfb - 1:nfe - nfb + 1, nfb - 1:nfe - nfb + 1]
    f2 = ((out[..., 0] * out[..., 0] + out[..., 1] * out[..., 1]) + out[..., 2] * out[..., 2])
    return np.ascontiguousarray(out), float(f2.max())


def coarse_force(r3, kern, nc):
    """W/pm.f90:275-315: pencil FFT of r3, i*kern_c, backward /ng^3, halo."""
    f = kern.dtype.type
    c = _rfftn(r3.astype(f))
    fc = np.zeros((nc, nc, nc, 3), f)
    for d in range(3):
        r = _irfftn_unnorm(apply_green(c, kern[d]), nc).astype(f)
        r = r / f(nc)
        r = r / f(nc)
        r = r / f(nc)
        fc[..., d] = r
    halo = np.pad(fc, ((1, 1), (1, 1), (1, 1), (0, 0)), mode="wrap")
    f2 = ((halo[..., 0] * halo[..., 0] + halo[..., 1] * halo[..., 1]) + halo[..., 2] * halo[..., 2])
    return np.ascontiguousarray(halo), float(f2.max())


# --------------------------------------------------------------------------
# state + reference subroutines
# --------------------------------------------------------------------------
class OracleCube:
    """One image (nn=1) of the reference state with its buffered layout."""

    def __init__(self, nc, nnt, ncb=6, izipx=2, izipv=2, np_nc=2, pp_on=False,
                 pp_range=1, rsoft=0.3, image_buffer=1.5, tile_buffer=2.5, has_pid=True):
        L = lib()
        self._c = L.oc_alloc()
        _check(L.oc_create(self._c, nc, nnt, ncb, izipx, izipv, np_nc, int(pp_on), pp_range,
                           rsoft, image_buffer, tile_buffer, int(has_pid)), "oc_create")
        self.nc, self.nnt, self.ncb, self.izipx, self.izipv = nc, nnt, ncb, izipx, izipv
        self.pp_on, self.pp_range, self.has_pid = pp_on, pp_range, has_pid
        info = self.info()
        self.nt, self.nte, self.nft, self.nfb, self.nfe = info[:5]
        self.np_image_max, self.np_tile_max = info[5], info[6]
        self.kern_f = None
        self.kern_c = None

    def __del__(self):
        try:
            lib().oc_free(self._c)
        except Exception:
            pass

    def info(self):
        o = np.zeros(16, np.int64)
        lib().oc_info(self._c, _p(o))
        return [int(v) for v in o]

    @property
    def nplocal(self):
        return self.info()[8]

    def init_kernels(self, fk=None, ck=None, dtype=np.float32):
        """W/kernel_f.f90 and W/kernel_c.f90 (W/initialize.f90:37-38)."""
        if fk is None:
            fk, ck = load_tables()
        self.kern_f = kernel_f(self.nfe, fk, self.pp_on, self.pp_range, dtype)
        self.kern_c = kernel_c(self.nc, ck, dtype=dtype)

    # -- state --------------------------------------------------------------
    def set_state(self, xp, vp, pid, rhoc, vfield, npglobal=None):
        n = xp.shape[0]
        xp = np.ascontiguousarray(xp, np.int16)
        vp = np.ascontiguousarray(vp, np.int16)
        pid = None if pid is None else np.ascontiguousarray(pid, np.int32)
        rhoc = np.ascontiguousarray(rhoc, np.int32)
        vfield = np.ascontiguousarray(vfield, np.float32)
        _check(lib().oc_set_disjoint(self._c, n, _p(xp), _p(vp), _p(pid), _p(rhoc), _p(vfield)),
               "set_state")
        lib().oc_set_npglobal(self._c, n if npglobal is None else npglobal)

    def state_from_xv(self, x, v, pid=None, sigma_vi=1.0):
        x = np.ascontiguousarray(x, np.float64)
        v = np.ascontiguousarray(v, np.float64)
        pid = None if pid is None else np.ascontiguousarray(pid, np.int32)
        _check(lib().oc_state_from_xv(self._c, x.shape[0], _p(x), _p(v), _p(pid), sigma_vi),
               "state_from_xv")

    def get_state(self):
        """Disjoint state in checkpoint order (W/checkpoint.f90:38-44)."""
        n = self.nplocal
        ncell = self.nc ** 3
        xp = np.zeros((n, 3), np.int16)
        vp = np.zeros((n, 3), np.int16)
        pid = np.zeros(n, np.int32) if self.has_pid else None
        rhoc = np.zeros(ncell, np.int32)
        vf = np.zeros((ncell, 3), np.float32)
        _check(lib().oc_get_disjoint(self._c, _p(xp), _p(vp), _p(pid), _p(rhoc), _p(vf)),
               "get_state")
        return dict(xp=xp, vp=vp, pid=pid, rhoc=rhoc, vfield=vf, nplocal=n)

    # -- W/main.f90:102-118 call shape ----------------------------------------
    def buffer_grid(self):
        _check(lib().oc_buffer_grid(self._c), "buffer_grid")

    def buffer_x(self):
        lib().oc_buffer_xvp(self._c, 1, 0)

    def buffer_v(self):
        lib().oc_buffer_xvp(self._c, 0, 1)

    def update_x(self, dt_old, dt, vz_max, sigma_vi):
        st = np.zeros(6, np.float64)
        sv = np.zeros(1, np.float32)
        code = lib().oc_update_xp(self._c, dt_old, dt, vz_max, sigma_vi, _p(st), _p(sv))
        _check(code, "update_x")
        return float(sv[0]), st

    def ext_pp_force(self, a_mid, dt, sigma_vi, mass_p):
        f2 = np.zeros(1, np.float32)
        _check(lib().oc_pp_force(self._c, a_mid, dt, sigma_vi, mass_p, _p(f2)), "ext_pp_force")
        return float(f2[0])

    def fine_deposit(self, itx, ity, itz, mass_p):
        nfe = self.nfe
        rho = np.zeros((nfe, nfe, nfe + 2), np.float32)
        s = lib().oc_fine_deposit(self._c, itx, ity, itz, mass_p, _p(rho))
        return rho, s

    def coarse_deposit(self, mass_p):
        r3 = np.zeros((self.nc,) * 3, np.float32)
        s = lib().oc_coarse_deposit(self._c, mass_p, _p(r3))
        return r3, s

    def particle_mesh(self, a_mid, dt, sigma_vi, sigma_vi_new, mass_p, keep_fields=False):
        """W/pm.f90:2-415 with the reference tile loop.  Returns the dt
        constraints' inputs and, optionally, per-tile fields."""
        nnt, nfe, nfb = self.nnt, self.nfe, self.nfb
        out = dict(f2_max_fine=0.0, testrhof=0.0, tiles={})
        for itz in range(1, nnt + 1):
            for ity in range(1, nnt + 1):
                for itx in range(1, nnt + 1):
                    rho, s = self.fine_deposit(itx, ity, itz, mass_p)
                    out["testrhof"] += s
                    ff, f2 = fine_force(rho[:, :, :nfe], self.kern_f, nfe, nfb)
                    out["f2_max_fine"] = max(out["f2_max_fine"], f2)
                    lib().oc_fine_kick(self._c, itx, ity, itz, _p(ff), a_mid, dt, sigma_vi,
                                       sigma_vi_new)
                    if keep_fields:
                        out["tiles"][(itx, ity, itz)] = dict(rho_f=rho, force_f=ff)
        sigma_vi = sigma_vi_new
        r3, sc = self.coarse_deposit(mass_p)
        fc, f2c = coarse_force(r3, self.kern_c, self.nc)
        vmax = np.zeros(3, np.float32)
        lib().oc_coarse_kick(self._c, _p(fc), a_mid, dt, sigma_vi, _p(vmax))
        out.update(testrhoc=sc, f2_max_coarse=f2c, vmax=vmax, sigma_vi=sigma_vi)
        if keep_fields:
            out.update(r3=r3, force_c=fc)
        return out

    def dt_constraints(self, pm, a_mid):
        """W/pm.f90:391-399 in real(4)."""
        f32 = np.float32
        GG = f32(1.0) / f32(6.0) / PI_F
        dt_fine = np.sqrt(f32(1.0) / (np.sqrt(f32(pm["f2_max_fine"])) * f32(a_mid) * GG))
        dt_coarse = np.sqrt(f32(4.0) / (np.sqrt(f32(pm["f2_max_coarse"])) * f32(a_mid) * GG))
        dt_vmax = f32(0.9) * f32(20) / pm["vmax"].max()
        return dict(dt_fine=float(dt_fine), dt_coarse=float(dt_coarse), dt_vmax=float(dt_vmax),
                    vz_max=float(pm["vmax"][2]))


# codec helpers --------------------------------------------------------------
def decode_x(izipx, code):
    return lib().oc_decode_x(izipx, int(code))


def decode_v(izipv, code, sigma):
    return lib().oc_decode_v(izipv, int(code), sigma)


def encode_v(izipv, v, sigma):
    return lib().oc_encode_v(izipv, float(v), sigma)


def encode_x_drift(izipx, code, dtv):
    return lib().oc_encode_x_drift(izipx, int(code), float(dtv))


def tanf_code(izipv, code):
    return lib().oc_tanf_code(izipv, int(code))


def expansion(a0, dt0, omega_m, omega_l, wde=-1.0):
    """W/timestep.f90:137-177 with the reference's kinds (oc_expansion):
    arrays of real(4) a0, dt0 -> real(4) da1, da2."""
    a0 = np.ascontiguousarray(np.atleast_1d(a0), np.float32)
    dt0 = np.ascontiguousarray(np.broadcast_to(np.asarray(dt0, np.float32), a0.shape))
    da1 = np.zeros_like(a0)
    da2 = np.zeros_like(a0)
    lib().oc_expansion_n(a0.size, _p(a0), _p(dt0), omega_m, omega_l, wde, _p(da1), _p(da2))
    return da1, da2


# --------------------------------------------------------------------------
# whole-box mesh fields at a chosen precision (the FP64 error statement)
# --------------------------------------------------------------------------
def cic_fine(pos, nc, mass, dtype=np.float64):
    """Fine-mesh CIC of W/pm.f90:50-108 on the whole periodic nf^3 mesh
    (nf = 4 nc), [z][y][x]: fine coordinate X = ncell*x with x the global
    coarse position, no half-cell shift (tempx = ncell*(cell-1) +
    ncell*offset, :63), weights 1-d and d on cells floor(X) and floor(X)+1."""
    nf = 4 * nc
    X = 4.0 * np.asarray(pos, np.float64)
    i0 = np.floor(X).astype(np.int64)
    d = (X - i0).astype(dtype)
    return _cic_scatter(i0, d, nf, mass, dtype)


def cic_coarse(pos, nc, mass, dtype=np.float64):
    """Coarse-mesh CIC of W/pm.f90:205-267 on the periodic nc^3 mesh:
    cell-centred, m = x - 0.5 (tempx = (cell-1) + offset - 0.5, :221)."""
    m = np.asarray(pos, np.float64) - 0.5
    i0 = np.floor(m).astype(np.int64)
    d = (m - i0).astype(dtype)
    return _cic_scatter(i0, d, nc, mass, dtype)


def _cic_scatter(i0, d, n, mass, dtype):
    rho = np.zeros(n ** 3, dtype)
    one = dtype(1)
    for cz in (0, 1):
        wz = d[:, 2] if cz else one - d[:, 2]
        for cy in (0, 1):
            wy = d[:, 1] if cy else one - d[:, 1]
            for cx in (0, 1):
                wx = d[:, 0] if cx else one - d[:, 0]
                idx = (((i0[:, 2] + cz) % n) * n + (i0[:, 1] + cy) % n) * n + (i0[:, 0] + cx) % n
                rho += np.bincount(idx, weights=(wx * wy * wz).astype(np.float64), minlength=n ** 3).astype(dtype)
    return (rho * dtype(mass)).reshape(n, n, n)


def fine_force_box(rho, fk, pp_on=False, pp_range=1):
    """Fine force of the whole periodic nf^3 mesh (W/pm.f90:112-124 with the
    kernel of W/kernel_f.f90 built for nf instead of one tile's nfe): the
    tile solve equals this inside each tile, the kernel reaching 15 fine
    cells and the tile buffer 24.  Precision follows rho.dtype; returns
    [z][y][x][3]."""
    nf = rho.shape[0]
    f = rho.dtype.type
    kern = kernel_f(nf, fk, pp_on, pp_range, dtype=rho.dtype)
    crho = _rfftn(rho)
    out = np.zeros(rho.shape + (3,), rho.dtype)
    for d in range(3):
        r = _irfftn_unnorm(apply_green(crho, kern[d]), nf).astype(f)
        out[..., d] = ((r / f(nf)) / f(nf)) / f(nf)
    return out


def fine_potential_box(rho, wg1, pp_on=False, pp_range=1):
    """Fine potential of the whole periodic nf^3 mesh: the 16^3 table of
    CUBEnu/kernels/wg1.ascii (README:8-9; wg1 in (i,j,k) order, i fastest)
    at offsets 0..15, mirrored evenly, its near corner zeroed with PP on
    like the force kernel (W/kernel_f.f90:29-31), applied in k space:
    phi = C2R(Re K * R2C(rho)) / nf^3.  Precision follows rho.dtype."""
    nf = rho.shape[0]
    f = rho.dtype.type
    t = np.asarray(wg1, rho.dtype).reshape(16, 16, 16)  # [k][j][i]
    rs = np.zeros((nf, nf, nf), rho.dtype)
    rs[:16, :16, :16] = t
    if pp_on:
        rs[:pp_range + 1, :pp_range + 1, :pp_range + 1] = 0
    rs[:, :, nf - 15:] = rs[:, :, 15:0:-1]
    rs[:, nf - 15:, :] = rs[:, 15:0:-1, :]
    rs[nf - 15:, :, :] = rs[15:0:-1, :, :]
    kre = _rfftn(rs).real.astype(f)
    crho = _rfftn(rho)
    prod = (crho.real * kre).astype(f) + 1j * (crho.imag * kre).astype(f)
    r = _irfftn_unnorm(prod.astype(crho.dtype), nf).astype(f)
    return ((r / f(nf)) / f(nf)) / f(nf)
