"""Shared helpers for the tests: synthetic states and comparisons."""
import numpy as np

from cafproject_amd import ics


def zeldovich_state(nc, nnt, seed=42, z_i=49.0, box=None, izipx=2, izipv=2):
    """Disjoint reference-format state from Zel'dovich ICs (numpy count sort)."""
    box = box or 3.125 * nc
    x, v, a = ics.zeldovich(nc, box, z_i, seed=seed)
    st = ics.disjoint_state(x, v, nc, nnt, izipx=izipx, izipv=izipv)
    st["a"] = a
    st["mass_p"] = float(np.float32((4 * nc) ** 3 / x.shape[0]))
    return st


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def global_force_to_tile(force, nc, nnt, itx, ity, itz):
    """force_f(3, nfb:nfe-nfb+1, ...) of tile (itx,ity,itz) cut from a global
    periodic fine force [z][y][x][3] (tile frame local fine 0..nft+1)."""
    nf = 4 * nc
    nft = 4 * (nc // nnt)
    idx = [((np.arange(nft + 2) - 1 + (t - 1) * nft) % nf) for t in (itx, ity, itz)]
    return np.ascontiguousarray(force[np.ix_(idx[2], idx[1], idx[0])])


def global_from_tiles(tiles, nc, nnt, key="force_f", inner=True):
    """Assemble a global fine field from per-tile oracle fields (physical part)."""
    nf = 4 * nc
    nft = 4 * (nc // nnt)
    out = None
    for (itx, ity, itz), d in tiles.items():
        f = d[key]
        if out is None:
            out = np.zeros((nf, nf, nf) + f.shape[3:], f.dtype)
        sl = f[1:nft + 1, 1:nft + 1, 1:nft + 1]
        out[(itz - 1) * nft:itz * nft, (ity - 1) * nft:ity * nft, (itx - 1) * nft:itx * nft] = sl
    return out
