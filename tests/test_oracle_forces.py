"""Pin the oracle's force parts: FFT conventions and the Green's functions
(W/kernel_f.f90, W/kernel_c.f90, W/pm.f90:112-126,275-315) through closed
forms, and the PP force (W/ext_pp_force.f90) through a brute-force pair sum."""
import math

import numpy as np
import pytest

from oracle import oracle as O


def test_fft_roundtrip_normalisation():
    """CUBEnu/pencil_fft/pencil_fft_test.f90:91-119: forward, backward,
    /N^3 returns the input (forward unnormalised, backward /N^3)."""
    rng = np.random.default_rng(0)
    r0 = rng.standard_normal((32, 32, 32)).astype(np.float32)
    c = O._rfftn(r0)
    assert abs(c[0, 0, 0] - r0.astype(np.float64).sum()) < 1e-3
    back = O._irfftn_unnorm(c, 32) / np.float32(32) / np.float32(32) / np.float32(32)
    assert np.abs(back - r0).max() < 1e-5


@pytest.mark.parametrize("pp_on", [False, True])
def test_fine_force_of_point_mass_is_the_table(pp_on):
    """A unit mass in one fine cell: the fine force field equals the
    real-space kernel -- the 16^3 table odd-mirrored (W/kernel_f.f90:25-31),
    with the (pp_range+1)^3 corner zeroed when PP is on -- placed around the
    mass.  Pins kernel_f, the r2c/c2r conventions, i*K and /nfe^3."""
    fk, ck = O.load_tables()
    nfe, nfb = 112, 24
    kern = O.kernel_f(nfe, fk, pp_on=pp_on, pp_range=1)
    rho = np.zeros((nfe, nfe, nfe), np.float32)
    src = (40, 50, 60)  # 0-based x, y, z
    rho[src[2], src[1], src[0]] = 1.0
    ff, _ = O.fine_force(rho, kern, nfe, nfb)
    # ff[z,y,x] covers 0-based indices nfb-1 .. nfe-nfb
    for (ox, oy, oz) in [(1, 0, 0), (0, -2, 0), (3, 1, -2), (-7, 5, 9), (15, -15, 15), (16, 0, 0), (0, 0, 1)]:
        x, y, z = src[0] + ox, src[1] + oy, src[2] + oz
        got = ff[z - (nfb - 1), y - (nfb - 1), x - (nfb - 1)]
        # force on a particle at the field point from the mass at src is
        # the kernel at offset (field - src) ... the table is f(offset)
        a = [abs(ox), abs(oy), abs(oz)]
        want = np.zeros(3)
        if max(a) <= 15 and not (pp_on and max(a) <= 1):
            want = fk[a[0], a[1], a[2]].astype(np.float64).copy()
            for d, o in enumerate((ox, oy, oz)):
                if o < 0:
                    want[d] = -want[d]
        assert np.allclose(got, want, atol=2e-6), ((ox, oy, oz), got, want)
    # FORCETEST (W/main.f90:104-111): unit mass one fine cell away along x
    # pulls with the table value -0.99957371 (wfxyzf.3.ascii line 2)
    if not pp_on:
        got = ff[src[2] - (nfb - 1), src[1] - (nfb - 1), src[0] + 1 - (nfb - 1), 0]
        assert abs(got - (-0.99957371)) < 2e-6


def test_coarse_force_of_point_mass_without_lrck():
    """Without the low-k correction kern_c = Im FFT(ck) of the odd real
    kernel, so a unit mass gives back ck itself: -r/r^3 in fine units, with
    the 4^3 table copied into the 8 octants next to the source
    (W/kernel_c.f90:16-72, the #else branch :326-332)."""
    fk, ckt = O.load_tables()
    nc = 32
    kern = O.kernel_c(nc, ckt, lrck=False)
    rho = np.zeros((nc, nc, nc), np.float32)
    rho[0, 0, 0] = 1
    fc, _ = O.coarse_force(rho, kern, nc)
    for (ox, oy, oz) in [(6, 0, 0), (0, 7, 2), (-5, 5, 5), (9, -3, 1), (2, 0, 0), (-1, 0, 0), (1, -2, 3)]:
        got = fc[1 + oz % nc, 1 + oy % nc, 1 + ox % nc].astype(np.float64)
        a = [abs(ox), abs(oy), abs(oz)]
        if max(a) <= 3 and all((o >= 0 and o <= 3) or (o < 0 and o >= -3) for o in (ox, oy, oz)):
            want = ckt[:, a[0], a[1], a[2]].astype(np.float64).copy()
            for d, o in enumerate((ox, oy, oz)):
                if o < 0:
                    want[d] = -want[d]
        else:
            r = 4.0 * np.array([ox, oy, oz], np.float64)
            want = -r / np.linalg.norm(r) ** 3
        assert np.allclose(got, want, atol=2e-6), ((ox, oy, oz), got, want)


def test_lrckcorr_closed_form():
    """LRCKCORR (W/kernel_c.f90:301-325) rescales the modes with |k|<=8 and
    k_d != 0 by 0.25 pi sin-FD/sum sin^2 over Im FFT(-r/r^3); with the
    table left out the two transforms coincide, so those modes must equal
    the finite-difference gradient Green's function 0.25 pi s_d/sum s^2,
    s = 2 sin(pi k/nc)."""
    _, ckt = O.load_tables()
    nc = 32
    kern = O.kernel_c(nc, ckt, table=False).astype(np.float64)
    k1 = (np.arange(nc) + nc // 2) % nc - nc // 2
    kx = k1[: nc // 2 + 1]
    KZ, KY, KX = np.meshgrid(k1, k1, kx, indexing="ij")
    s = [2 * np.sin(np.pi * K / nc) for K in (KX, KY, KZ)]
    ssum = s[0] ** 2 + s[1] ** 2 + s[2] ** 2
    kr = np.sqrt(KX ** 2 + KY ** 2 + KZ ** 2)
    for d, K in enumerate((KX, KY, KZ)):
        m = (kr <= 8) & (K != 0)
        want = 0.25 * np.pi * s[d][m] / ssum[m]
        assert np.allclose(kern[d][m], want, rtol=2e-5, atol=1e-7)


def test_kernel_symmetries():
    """kern_f is the imaginary part of the FFT of an odd real kernel: odd in
    the component's own wave number, even in the others."""
    fk, _ = O.load_tables()
    nfe = 64
    kern = O.kernel_f(nfe, fk).astype(np.float64)   # [d][kz][ky][kx]
    ky = np.arange(nfe)
    neg = (-ky) % nfe
    # K_y(kx,-ky,kz) = -K_y(kx,ky,kz);  K_x(kx,-ky,kz) = K_x(kx,ky,kz)
    assert np.allclose(kern[1][:, neg, :], -kern[1], atol=1e-5)
    assert np.allclose(kern[0][:, neg, :], kern[0], atol=1e-5)
    assert np.allclose(kern[2][neg, :, :], -kern[2], atol=1e-5)
    assert np.abs(kern[0][:, :, 0]).max() < 1e-5   # odd in kx: zero at kx=0


def _pp_state(n, nc, seed):
    rng = np.random.default_rng(seed)
    centers = rng.uniform(0, nc, (n // 8, 3))
    pos = (centers[:, None, :] + rng.normal(0, 0.15, (n // 8, 8, 3))).reshape(-1, 3) % nc
    return pos


def test_pp_force_against_brute_force():
    """Each particle's PP kick equals the brute-force FP64 pair sum over
    particles in the 27 neighbouring fine cells, F = m r/r^3 (1-7/4(r/16)^3
    +3/4(r/16)^5) for r > rsoft (W/ext_pp_force.f90:94-116); checked through
    the decoded velocity change, which must match to within one velocity
    quantum, and through f2_max_pp."""
    nc, nnt = 16, 1
    pos = _pp_state(4096, nc, 1)
    oc = O.OracleCube(nc, nnt, pp_on=True)
    sigma = 5.0
    oc.state_from_xv(pos, np.zeros_like(pos), None, sigma)
    oc.buffer_grid(); oc.buffer_x(); oc.buffer_v()
    st0 = oc.get_state()
    mass, a_mid, dt = 1.0, 0.5, 0.3
    f2 = oc.ext_pp_force(a_mid, dt, sigma, mass)
    st1 = oc.get_state()
    # decoded positions in fine units
    cid = np.repeat(np.arange(nc ** 3), st0["rhoc"])
    g = np.stack([cid % nc, (cid // nc) % nc, cid // nc ** 2], 1)
    xf = 4.0 * (g + ((st0["xp"].astype(np.int64) & 0xFFFF) + 0.5) / 65536)
    fcell = np.floor(xf).astype(np.int64)
    nf = 4 * nc
    F = np.zeros_like(xf)
    for p in range(len(xf)):
        dfc = (fcell - fcell[p] + nf // 2) % nf - nf // 2
        near = np.all(np.abs(dfc) <= 1, axis=1)
        d = xf[near] - xf[p]
        d = (d + nf / 2) % nf - nf / 2
        r = np.sqrt((d * d).sum(1))
        r = np.where(r == 0, 1.0, r)
        rc = r / 16
        pc = 1 - 1.75 * rc ** 3 + 0.75 * rc ** 5
        f = mass * d / r[:, None] ** 3 * pc[:, None]
        f[r <= np.float32(0.3)] = 0
        F[p] = f.sum(0)
    vt = np.array([O.decode_v(2, c, sigma) for c in range(-32768, 32768)])
    dv = vt[st1["vp"].astype(np.int64) + 32768] - vt[st0["vp"].astype(np.int64) + 32768]
    want = F * a_mid * dt / 6 / float(np.float32(np.pi))
    code = st1["vp"].astype(np.int64) + 32768
    q = np.abs(vt[np.clip(code + 1, 0, 65535)] - vt[np.clip(code - 1, 0, 65535)])
    assert np.all(np.abs(dv - want) <= q + 1e-12)
    assert abs(f2 / float(np.float32((F * F).sum(1).max())) - 1) < 1e-6


def test_threaded_force_glue_equals_numpy():
    """The oracle's threaded C glue of the FP32 fine force (Green multiply,
    /nfe three times, extraction, max |F|^2) gives the same bits as the
    element-wise numpy expressions of W/pm.f90:119-126."""
    import scipy.fft as sfft
    from oracle import oracle as O
    rng = np.random.default_rng(3)
    nfe, nfb = 40, 8
    rho = rng.standard_normal((nfe, nfe, nfe)).astype(np.float32)
    kern = rng.standard_normal((3, nfe, nfe, nfe // 2 + 1)).astype(np.float32)
    out, f2 = O.fine_force(rho, kern, nfe, nfb)
    f = np.float32
    crho = sfft.rfftn(rho)
    n2 = nfe - 2 * nfb + 2
    ref = np.zeros((n2, n2, n2, 3), f)
    for d in range(3):
        g = ((-crho.imag * kern[d]).astype(f) + 1j * (crho.real * kern[d]).astype(f)).astype(np.complex64)
        r = sfft.irfftn(g, s=(nfe,) * 3, norm="forward").astype(f)
        r = r / f(nfe); r = r / f(nfe); r = r / f(nfe)
        ref[..., d] = r[nfb - 1:nfe - nfb + 1, nfb - 1:nfe - nfb + 1, nfb - 1:nfe - nfb + 1]
    assert np.array_equal(out.view(np.uint32), ref.view(np.uint32))
    f2r = ((ref[..., 0] * ref[..., 0] + ref[..., 1] * ref[..., 1]) + ref[..., 2] * ref[..., 2]).max()
    assert np.float32(f2) == f2r
