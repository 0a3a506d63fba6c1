"""GPU Poisson solves (R2C, Green's function multiply, 3 x C2R) at the mesh
sizes of the register-resident FFT path, against the oracle and against
closed forms.

* fine force of a seeded random density, whole-box periodic on the GPU,
  against the oracle's per-tile solve (W/pm.f90:112-126) on the same
  density cut around each tile: the kernel reaches 15 fine cells, inside
  the 24-cell tile buffer, so the two agree to FP32 rounding;
* coarse force of a random density against the oracle's pencil solve
  (W/pm.f90:275-315);
* at the 1024^3 and 2048^3 fine meshes (nc=256, 512), where the oracle is
  too slow for a test: a density of plane waves cos(2 pi k.r/N) has the
  exact force -K_d(k) sin(2 pi k.r/N) per wave, with K_d(k) read back from
  the library's own kern_f (octant storage, odd in k_d).
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import rel_l2

pytestmark = pytest.mark.gpu


def _tile_rho(rho, nc, nnt, t, nfe, nfb):
    nf = 4 * nc
    nft = 4 * (nc // nnt)
    idx = [(T * nft + np.arange(nfe) - nfb) % nf for T in t]
    return np.ascontiguousarray(rho[np.ix_(idx[2], idx[1], idx[0])])


@pytest.mark.parametrize("nc,nnt", [(16, 1), (32, 2), (64, 1), (128, 2)])
def test_fine_force_random_density(nc, nnt):
    from cafproject_amd.cube import Cube
    nf = 4 * nc
    rng = np.random.default_rng(nc + nnt)
    rho = rng.random((nf, nf, nf), dtype=np.float32) * 2
    g = Cube(nc, nnt, np_nc=1)
    g.init_kernels()
    g.set_field("rho_f", rho)
    f2g = g.fine_force()
    fg = g.get_field("force_f")
    fk, ck = O.load_tables()
    nft, nfb = 4 * (nc // nnt), 24
    nfe = nft + 2 * nfb
    kern = O.kernel_f(nfe, fk)
    fo = np.zeros_like(fg)
    f2o = 0.0
    for tz in range(nnt):
        for ty in range(nnt):
            for tx in range(nnt):
                ff, f2 = O.fine_force(_tile_rho(rho, nc, nnt, (tx, ty, tz), nfe, nfb), kern, nfe, nfb)
                fo[tz * nft:(tz + 1) * nft, ty * nft:(ty + 1) * nft, tx * nft:(tx + 1) * nft] = \
                    ff[1:nft + 1, 1:nft + 1, 1:nft + 1]
                f2o = max(f2o, f2)
    err = rel_l2(fg, fo)
    assert err < 1e-5, err
    # the oracle's max also covers the one-cell tile halo, which is the
    # neighbour's physical cell: the same set of values
    assert abs(f2g / f2o - 1) < 1e-4


@pytest.mark.parametrize("nc", [32, 64, 128])
def test_coarse_force_random_density(nc):
    from cafproject_amd.cube import Cube
    rng = np.random.default_rng(nc)
    r3 = rng.random((nc, nc, nc), dtype=np.float32) * 8
    g = Cube(nc, 1, np_nc=1)
    g.init_kernels()
    g.set_field("rho_c", r3)
    f2g = g.coarse_force()
    fk, ck = O.load_tables()
    halo, f2o = O.coarse_force(r3, O.kernel_c(nc, ck), nc)
    err = rel_l2(g.get_field("force_c"), halo[1:-1, 1:-1, 1:-1])
    assert err < 1e-5, err
    assert abs(f2g / f2o - 1) < 1e-4


def _kd(kern, k, N):
    """K_d at integer wave vector k=(kx,ky,kz), 0 < kx < N/2, from the
    octant storage [d][|kz|][|ky|][kx] with the odd sign of component d."""
    kx, ky, kz = k
    ay, az = min(ky % N, N - ky % N), min(kz % N, N - kz % N)
    sy = 1.0 if ky % N <= N // 2 else -1.0
    sz = 1.0 if kz % N <= N // 2 else -1.0
    return np.array([kern[0, az, ay, kx], kern[1, az, ay, kx] * sy, kern[2, az, ay, kx] * sz], np.float64)


def test_plane_waves_1024():
    from cafproject_amd.cube import Cube
    nc, N = 256, 1024
    waves = [((3, 5, 7), 1.0, 0.3), ((1, N - 2, 4), 0.5, 1.1), ((17, 9, N - 30), 0.25, 2.0)]
    g = Cube(nc, 2, np_nc=1)
    g.init_kernels()
    kern = g.get_field("kern_f")
    x = np.arange(N, dtype=np.float64)
    rho = np.empty((N, N, N), np.float32)
    for z in range(N):
        acc = np.zeros((N, N))
        for (kx, ky, kz), amp, ph in waves:
            acc += amp * np.cos(2 * np.pi * (kx * x[None, :] + ky * x[:, None] + kz * z) / N + ph)
        rho[z] = acc
    g.set_field("rho_f", rho)
    del rho
    g.fine_force()
    f = g.get_field("force_f")
    for z in (0, 1, 333, 1023):
        want = np.zeros((N, N, 3))
        for k, amp, ph in waves:
            kx, ky, kz = k
            s = np.sin(2 * np.pi * (kx * x[None, :] + ky * x[:, None] + kz * z) / N + ph)
            want += -amp * s[..., None] * _kd(kern, k, N)[None, None, :]
        err = rel_l2(f[z], want)
        assert err < 1e-5, (z, err)


def test_plane_waves_2048():
    """The same closed form at the 2048^3 fine mesh (nc=512, the compact
    layout: one force spectrum at a time), whose strided passes run with 8
    columns per workgroup and 32 x 64 register lines.  The density is formed
    on the GPU by torch, 32 z planes at a time, and crosses to the library
    through host memory (32 GB); the force comes back whole (96 GB) and four
    z planes are checked."""
    import gc
    import torch
    from cafproject_amd.cube import Cube
    nc, N = 512, 2048
    waves = [((3, 5, 7), 1.0, 0.3), ((1, N - 2, 4), 0.5, 1.1), ((17, 9, N - 30), 0.25, 2.0)]
    rho = np.empty((N, N, N), np.float32)
    dev = torch.device("cuda:0")
    xi = torch.arange(N, device=dev, dtype=torch.int64)
    ch = 32
    for z0 in range(0, N, ch):
        zi = torch.arange(z0, z0 + ch, device=dev, dtype=torch.int64)
        acc = torch.zeros((ch, N, N), device=dev, dtype=torch.float64)
        for (kx, ky, kz), amp, ph in waves:
            m = (kz * zi[:, None, None] + ky * xi[None, :, None] + kx * xi[None, None, :]) % N
            acc += amp * torch.cos(m.to(torch.float64) * (2 * np.pi / N) + ph)
        rho[z0:z0 + ch] = acc.to(torch.float32).cpu().numpy()
        del acc, m
    del xi, zi
    torch.cuda.empty_cache()
    print("plane waves 2048: density formed", flush=True)
    g = Cube(nc, 4, np_nc=1)
    g.init_kernels()
    g.set_field("rho_f", rho)
    del rho
    gc.collect()
    g.fine_force()
    print("plane waves 2048: force solved", flush=True)
    f = g.get_field("force_f")
    kern = g.get_field("kern_f")
    kd = [_kd(kern, k, N) for k, _, _ in waves]
    del kern, g
    gc.collect()
    x = np.arange(N, dtype=np.int64)
    for z in (0, 1, 777, 2047):
        want = np.zeros((N, N, 3))
        for ((kx, ky, kz), amp, ph), K in zip(waves, kd):
            m = (kx * x[None, :] + ky * x[:, None] + kz * z) % N
            s = np.sin(2 * np.pi * m / N + ph)
            want += -amp * s[..., None] * K[None, None, :]
        err = rel_l2(f[z], want)
        print(f"plane waves 2048: z={z} rel L2 {err:.2e}", flush=True)
        assert err < 1e-5, (z, err)


@pytest.mark.parametrize("nc", [16, 24])
def test_kern_f_crosses_the_boundary_as_the_octant(nc):
    """kern_f is blocked on the device and crosses cube_get_field /
    cube_set_field as the plain octant.  Setting 2*K must double the force
    exactly (the solve is linear and scaling by 2 is exact), which fails if
    the two conversions are not inverse layouts of the device one.  nc=16
    (nf=64) takes the register z-pass, nc=24 (nf=96) the LDS one."""
    from cafproject_amd.cube import Cube
    nf = 4 * nc
    rho = np.random.default_rng(nc).random((nf, nf, nf), dtype=np.float32)
    g = Cube(nc, 1, np_nc=1)
    g.init_kernels()
    g.set_field("rho_f", rho)
    g.fine_force()
    f1 = g.get_field("force_f")
    k = g.get_field("kern_f")
    g.set_field("kern_f", 2 * k)
    assert np.array_equal(g.get_field("kern_f"), 2 * k)
    g.set_field("rho_f", rho)
    g.fine_force()
    assert np.array_equal(g.get_field("force_f"), 2 * f1)
