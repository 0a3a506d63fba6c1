"""Matter power spectrum of a checkpoint state, as the reference's analysis
tools compute it (WU/cicpower.f90:96-187, WU/powerspectrum.f90:28-135;
WU/ = CUBEnu/work/utilities/):

  cell-centred CIC of the particles on an ng^3 mesh (pos*ng/nc - 0.5),
  delta = rho/mean - 1, delta_k = FFT(delta)/ng^3, Hermitian duplicates of
  the kx=0 and kx=ng/2 planes skipped, linear bins ibin = nint(|k|),
  Delta^2 = |delta_k|^2 / sinc^4 * 4 pi |k|^3 averaged per bin,
  k_phys = <|k|> 2 pi / box.
"""
import numpy as np


def decode_positions(st, nc, nnt, izipx):
    """Global coarse-cell positions of a disjoint state (float64)."""
    nt = nc // nnt
    cid = np.repeat(np.arange(nc ** 3, dtype=np.int64), st["rhoc"])
    i = cid % nt
    j = (cid // nt) % nt
    k = (cid // nt ** 2) % nt
    t = cid // nt ** 3
    tx, ty, tz = t % nnt, (t // nnt) % nnt, t // nnt ** 2
    g = np.stack([tx * nt + i, ty * nt + j, tz * nt + k], axis=1).astype(np.float64)
    n = 1 << (8 * izipx)
    return g + ((st["xp"].astype(np.int64) & (n - 1)) + 0.5) / n


def cic_density(pos, nc, ng):
    """WU/cicpower.f90:104-146: periodic cell-centred CIC, float64."""
    p = pos * (ng / nc) - 0.5
    i1 = np.floor(p).astype(np.int64)
    d2 = p - i1
    d1 = 1.0 - d2
    rho = np.zeros(ng ** 3)
    for cz in (0, 1):
        wz = d2[:, 2] if cz else d1[:, 2]
        for cy in (0, 1):
            wy = d2[:, 1] if cy else d1[:, 1]
            for cx in (0, 1):
                wx = d2[:, 0] if cx else d1[:, 0]
                idx = ((i1[:, 2] + cz) % ng) * ng * ng + ((i1[:, 1] + cy) % ng) * ng + (i1[:, 0] + cx) % ng
                rho += np.bincount(idx, weights=wx * wy * wz, minlength=ng ** 3)
    return rho.reshape(ng, ng, ng)


def power(delta, box):
    """WU/powerspectrum.f90:28-135 auto power with linear k bins.
    Returns (k_phys, Delta^2, nmodes) per bin 1..nbin."""
    ng = delta.shape[0]
    ck = np.fft.rfftn(delta) / ng ** 3
    ny = ng // 2
    f = np.fft.fftfreq(ng) * ng
    fx = np.arange(ny + 1)
    KZ, KY, KX = np.meshgrid(f, f, fx, indexing="ij")
    # Fortran mod(i+ny-1,ng)-ny maps index ny (Nyquist) to -ny
    KY = np.where(KY == ny, -ny, KY)
    KZ = np.where(KZ == ny, -ny, KZ)
    KXs = np.where(KX == ny, -ny, KX)
    iy = np.arange(ng)[None, :, None] * np.ones_like(KX, dtype=int)
    iz = np.arange(ng)[:, None, None] * np.ones_like(KX, dtype=int)
    ix = fx[None, None, :] * np.ones_like(KX, dtype=int)
    skip = (ix == 0) & (iy == 0) & (iz == 0)
    edge_x = (ix == 0) | (ix == ny)
    skip |= edge_x & (iy > ny)
    skip |= edge_x & ((iy == 0) | (iy == ny)) & (iz > ny)
    kr = np.sqrt(KXs ** 2 + KY ** 2 + KZ ** 2)

    def sinc(k):
        a = np.pi * k / ng
        return np.where(k == 0, 1.0, np.sin(a) / np.where(a == 0, 1, a))

    s = sinc(KXs) * sinc(KY) * sinc(KZ)
    amp = (np.abs(ck) ** 2) / s ** 4 * 4 * np.pi * kr ** 3
    ibin = np.rint(kr).astype(np.int64)
    keep = ~skip
    nbin = int(round(ny * np.sqrt(3.0)))
    cnt = np.bincount(ibin[keep], minlength=nbin + 2)[: nbin + 1]
    ksum = np.bincount(ibin[keep], weights=kr[keep], minlength=nbin + 2)[: nbin + 1]
    psum = np.bincount(ibin[keep], weights=amp[keep], minlength=nbin + 2)[: nbin + 1]
    with np.errstate(invalid="ignore", divide="ignore"):
        k_phys = ksum / cnt * 2 * np.pi / box
        d2 = psum / cnt
    return k_phys[1:], d2[1:], cnt[1:]


def state_power(st, nc, nnt, izipx, box, ng=None):
    ng = ng or nc
    pos = decode_positions(st, nc, nnt, izipx)
    rho = cic_density(pos, nc, ng)
    delta = rho / rho.mean() - 1.0
    return power(delta, box)
