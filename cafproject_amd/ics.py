"""Synthetic Zel'dovich initial conditions in the reference checkpoint format.

This is the harness-side counterpart of the reference IC generator
(CUBEnu/work/utilities/initial_conditions.f90): seeded unit-variance white
noise, an Eisenstein & Hu (1998) no-wiggle LCDM spectrum normalised to sigma8
with a top-hat of 8 Mpc/h (the reference's normalisation, :218-225), the
growth factor Dgrow and velocity factor vfactor of the reference (:953-1003),
potential phi = -4 pi delta / k~^2 with the finite-difference k~ (:433-447),
Zel'dovich displacement x = q - grad(phi)/(8 pi ncell) and velocity
v = -grad(phi)/(8 pi) vfactor(a) from central differences (:719-724,
:754-757), then the count sort into the disjoint cell-sorted state with
xp = floor(x/x_resolution) and vp = enc(v - vfield) (:737-769).

Differences from the reference, all deliberate: the noise is this project's
own seeded numpy generator (the reference's random_number stream is
compiler-specific); the spectrum is EH no-wiggle instead of a CAMB table; the
displacement is evaluated on the particle lattice (np_nc per coarse cell)
rather than on the fine grid, which lets the 512^3-particle workload be made
in seconds; the within-cell particle order is lattice order.
"""
import math

import numpy as np

PI_F = float(np.float32(4.0) * np.arctan(np.float32(1.0)))


# --------------------------------------------------------------------------
# cosmology
# --------------------------------------------------------------------------
def eh_nowiggle_transfer(k_hmpc, omega_m, omega_b, h, tcmb=2.7255):
    """Eisenstein & Hu 1998 (ApJ 496, 605) eqs. 26-31, zero-baryon-wiggle form."""
    k = np.asarray(k_hmpc, np.float64) * h  # 1/Mpc
    theta = tcmb / 2.7
    om_h2 = omega_m * h * h
    ob_h2 = omega_b * h * h
    fb = omega_b / omega_m
    s = 44.5 * np.log(9.83 / om_h2) / np.sqrt(1.0 + 10.0 * ob_h2 ** 0.75)
    alpha = 1.0 - 0.328 * np.log(431.0 * om_h2) * fb + 0.38 * np.log(22.3 * om_h2) * fb * fb
    gamma = omega_m * h * (alpha + (1.0 - alpha) / (1.0 + (0.43 * k * s) ** 4))
    q = k_hmpc * theta * theta / gamma
    L0 = np.log(2.0 * math.e + 1.8 * q)
    C0 = 14.2 + 731.0 / (1.0 + 62.5 * q)
    return L0 / (L0 + C0 * q * q)


def tophat(x):
    x = np.asarray(x, np.float64)
    out = np.ones_like(x)
    nz = x != 0
    xn = x[nz]
    out[nz] = 3.0 * (np.sin(xn) - np.cos(xn) * xn) / xn ** 3
    return out


def dgrow(a, om, ol):
    """Dgrow of WU/initial_conditions.f90:953-966 (Carroll, Press & Turner)."""
    hsq = om / a ** 3 + (1 - om - ol) / a ** 2 + ol
    oma = om / (a ** 3 * hsq)
    ola = ol / hsq
    g = 2.5 * om / (om ** (4.0 / 7) - ol + (1 + om / 2) * (1 + ol / 70))
    ga = 2.5 * oma / (oma ** (4.0 / 7) - ola + (1 + oma / 2) * (1 + ola / 70))
    return a * ga / g


def vfactor(a, om, ol):
    """vfactor of WU/initial_conditions.f90:990-1003 without neutrinos."""
    lm = ol / om
    km = (1 - om - ol) / om
    H = 2.0 / (3.0 * math.sqrt(a ** 3)) * math.sqrt(1 + a * km + a ** 3 * lm)
    return a * a * H


class Cosmology:
    def __init__(self, omega_m=0.31, omega_b=0.049, h=0.68, n_s=0.96, sigma8=0.81):
        self.omega_m, self.omega_b, self.h, self.n_s, self.sigma8 = omega_m, omega_b, h, n_s, sigma8
        self.omega_l = 1.0 - omega_m
        lk = np.linspace(np.log(1e-5), np.log(1e3), 20000)
        k = np.exp(lk)
        d2 = k ** (3 + n_s) * eh_nowiggle_transfer(k, omega_m, omega_b, h) ** 2
        v8 = np.trapezoid(d2 * tophat(8.0 * k) ** 2, lk)
        self._norm = sigma8 ** 2 / v8

    def delta2(self, k, a=1.0):
        """Dimensionless power k^3 P(k)/(2 pi^2) at scale factor a."""
        k = np.asarray(k, np.float64)
        t = eh_nowiggle_transfer(k, self.omega_m, self.omega_b, self.h)
        return self._norm * k ** (3 + self.n_s) * t * t * dgrow(a, self.omega_m, self.omega_l) ** 2

    def pk(self, k, a=1.0):
        k = np.asarray(k, np.float64)
        with np.errstate(divide="ignore", invalid="ignore"):
            return np.where(k > 0, 2 * math.pi ** 2 * self.delta2(k, a) / np.maximum(k, 1e-30) ** 3, 0.0)


# --------------------------------------------------------------------------
# Zel'dovich particles
# --------------------------------------------------------------------------
def zeldovich(nc, box, z_i, seed, cosmo=None, np_nc=2, ncell=4, dtype=np.float32, workers=None):
    """Particle positions (global coarse-cell units, periodic in [0,nc)) and
    velocities (fine cells per unit time) on the np_nc-per-cell lattice."""
    import os

    import scipy.fft as sfft
    workers = workers or os.cpu_count()
    if cosmo is None:
        cosmo = Cosmology()
    a = 1.0 / (1.0 + z_i)
    ng = nc * np_nc                     # lattice points per dim
    dx_fine = ncell / np_nc             # lattice spacing in fine cells
    cell = box / ng                     # Mpc/h per lattice cell
    rng = np.random.default_rng(seed)
    w = rng.standard_normal((ng, ng, ng), dtype=np.float32)
    wk = sfft.rfftn(w, workers=workers)
    del w
    kf = 2 * np.pi * np.fft.fftfreq(ng)                 # radians per lattice cell
    kx = 2 * np.pi * np.fft.rfftfreq(ng)
    # P(k) tabulated once on a log grid and interpolated (k in h/Mpc)
    kmin = 2 * np.pi / box * 0.5
    kmax = np.sqrt(3.0) * np.pi / cell * 1.01
    lk = np.linspace(np.log(kmin), np.log(kmax), 4096)
    lpk = np.log(np.maximum(cosmo.pk(np.exp(lk), a), 1e-300))
    sy = (2 * np.sin(kf / 2)) ** 2
    sx = (2 * np.sin(kx / 2)) ** 2
    ky2 = (kf / cell) ** 2
    kx2 = (kx / cell) ** 2
    for iz in range(ng):
        kz = kf[iz]
        kk = np.sqrt(ky2[:, None] + kx2[None, :] + (kz / cell) ** 2)
        amp = np.exp(0.5 * np.interp(np.log(np.maximum(kk, kmin)), lk, lpk)) / cell ** 1.5
        k2 = sy[:, None] + sx[None, :] + (2 * np.sin(kz / 2)) ** 2
        k2[k2 == 0] = 1.0
        # phi = -4 pi delta / k~^2 in lattice units (WU/initial_conditions.f90:433-447)
        wk[iz] *= (amp * (-4 * np.pi / k2)).astype(np.float32)
    wk[0, 0, 0] = 0
    phi = sfft.irfftn(wk, s=(ng, ng, ng), axes=(0, 1, 2), workers=workers).astype(np.float32)
    del wk
    # displacement: -(phi(i+1)-phi(i-1))/(8 pi) lattice cells (central difference)
    q = ((np.arange(ng, dtype=np.float64) + 0.25) / np_nc).astype(np.float32)
    x = np.empty((ng, ng, ng, 3), dtype)
    v = np.empty((ng, ng, ng, 3), dtype)
    vf = np.float32(vfactor(a, cosmo.omega_m, cosmo.omega_l))
    scale = np.float32(-dx_fine / (8 * np.pi))
    for axis, comp in ((2, 0), (1, 1), (0, 2)):
        g = np.roll(phi, -1, axis=axis)
        g -= np.roll(phi, 1, axis=axis)
        g *= scale                                          # fine cells
        shape = [1, 1, 1]
        shape[axis] = ng
        xc = q.reshape(shape) + g / np.float32(ncell)
        np.mod(xc, np.float32(nc), out=xc)
        x[..., comp] = xc
        v[..., comp] = g * vf
        del g, xc
    return x.reshape(-1, 3), v.reshape(-1, 3), a


# --------------------------------------------------------------------------
# disjoint state (reference checkpoint layout)
# --------------------------------------------------------------------------
def _wrap(bytes_, v):
    if bytes_ == 1:
        return v.astype(np.int64).astype(np.uint8).astype(np.int8).astype(np.int16)
    return v.astype(np.int64).astype(np.uint16).astype(np.int16)


def encode_v(izipv, v, sigma):
    """vp = nint((nvbin-1) atan(sqrt(pi/2)/(sigma*2.5) v)/pi)."""
    c = float(np.float32(math.sqrt(np.float32(PI_F / 2)))) / (float(np.float32(sigma)) * 2.5)
    nvm1 = float((1 << (8 * izipv)) - 1)
    y = (nvm1 * np.arctan(c * np.asarray(v, np.float64))) / PI_F
    r = np.trunc(y) + np.where(np.abs(y - np.trunc(y)) >= 0.5, np.sign(y), 0.0)
    return _wrap(izipv, r)


def cell_ids(x, nc, nnt):
    """Disjoint-order cell id (tile, k, j, i) of global coarse positions in
    [0, nc): cell ceiling(x) with x == 0 in the last cell."""
    nt = nc // nnt
    g = np.ceil(x).astype(np.int32)
    g -= 1
    g[g < 0] += nc
    g[g >= nc] -= nc
    t = g // nt
    g -= t * nt
    tile = t[:, 0] + nnt * (t[:, 1] + nnt * t[:, 2])
    return (g[:, 0] + nt * (g[:, 1] + nt * (g[:, 2] + nt * tile.astype(np.int64)))).astype(np.int64)


def disjoint_state(x, v, nc, nnt, izipx=2, izipv=2, sigma_vi=None, pid=None):
    """Count-sort particles into the reference's disjoint state.  Returns a
    dict with xp, vp, pid, rhoc, vfield, sigma_vi."""
    n = x.shape[0]
    ncells = nc ** 3
    cid = cell_ids(x, nc, nnt)
    order = np.argsort(cid, kind="stable")
    cid_s = cid[order]
    rhoc = np.bincount(cid, minlength=ncells).astype(np.int32)
    vfield = np.zeros((ncells, 3), np.float32)
    for d in range(3):
        s = np.bincount(cid, weights=v[:, d].astype(np.float64), minlength=ncells)
        vfield[:, d] = (s / np.maximum(rhoc, 1)).astype(np.float32)
    xs = np.mod(x[order].astype(np.float64), nc)
    res = 1 << (8 * izipx)
    xp = _wrap(izipx, np.floor(xs * res))
    vrel = v[order].astype(np.float64) - vfield[cid_s].astype(np.float64)
    if sigma_vi is None:
        sigma_vi = float(np.sqrt(np.mean(vrel * vrel)) )
        sigma_vi = float(np.float32(sigma_vi))
    vp = encode_v(izipv, vrel, sigma_vi)
    pid_s = (np.arange(1, n + 1, dtype=np.int32) if pid is None else pid)[order]
    return dict(xp=np.ascontiguousarray(xp), vp=np.ascontiguousarray(vp), pid=pid_s.astype(np.int32),
                rhoc=rhoc, vfield=vfield, sigma_vi=sigma_vi, nplocal=n)
