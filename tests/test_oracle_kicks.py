"""Pin the oracle's fine and coarse kicks, its v_max reduction and the dt
constraints against the mathematics (W/pm.f90:127-159, 319-387, 391-399).

CIC interpolation reproduces a linear function exactly, so with a force
field that is linear in the mesh coordinates every particle's kick must be
F(x_p)*a_mid*dt/(6 pi), x_p its CIC coordinate.  The expected velocity code
is computed here from that closed form through the codec (pinned by its own
known-answer and exhaustive tests); the oracle's codes must equal it except
where the exact value sits within FP32 rounding of a code boundary.  The
field has distinct gradients per component and axis and a constant part, and
the decode and encode dispersions differ, so a transposed operand, a swapped
CIC weight, a shifted node index, a wrong factor or the wrong sigma fails.
v_max is checked against a brute-force max of |v + v_c| over particles."""
import math

import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import zeldovich_state

PI_F = np.float32(4.0) * np.arctan(np.float32(1.0))


def _state(nc, nnt, izipx, izipv, seed=11):
    st = zeldovich_state(nc, nnt, seed=seed, izipx=izipx, izipv=izipv)
    oc = O.OracleCube(nc, nnt, izipx=izipx, izipv=izipv)
    oc.set_state(st["xp"], st["vp"], st["pid"], st["rhoc"], st["vfield"])
    oc.buffer_grid(); oc.buffer_x(); oc.buffer_v()
    return st, oc


def _cells(st, nc, nnt):
    nt = nc // nnt
    cid = np.repeat(np.arange(nc ** 3), st["rhoc"])
    ijk = np.stack([cid % nt, (cid // nt) % nt, (cid // nt ** 2) % nt], axis=1)
    t = cid // nt ** 3
    txyz = np.stack([t % nnt, (t // nnt) % nnt, t // nnt ** 2], axis=1)
    return cid, ijk, txyz


def _offsets(st, izipx):
    return np.vectorize(lambda c: O.decode_x(izipx, int(c)))(st["xp"].astype(np.int64))


def _codes(izipv, v, sigma):
    return np.array([O.encode_v(izipv, float(x), sigma) for x in v.reshape(-1)]).reshape(v.shape)


def _decode(izipv, codes, sigma):
    return np.array([O.decode_v(izipv, int(c), sigma) for c in codes.reshape(-1)]).reshape(codes.shape)


def _wrap(izipv, c):
    m = 1 << (8 * izipv)
    return ((c + m // 2) % m) - m // 2


def _compare(izipv, got, want):
    d = np.abs(_wrap(izipv, got.astype(np.int64) - want.astype(np.int64)))
    assert d.max() <= 1, d.max()
    assert (d == 0).mean() > 0.995, (d == 0).mean()


# F_d(p) = C[d] + G[d] . p, p the mesh coordinate of a node
C = np.array([0.7, -1.3, 0.4])
G = np.array([[0.11, -0.05, 0.023], [0.031, 0.09, -0.07], [-0.06, 0.017, 0.13]])


@pytest.mark.parametrize("izipx,izipv", [(2, 2), (1, 2), (2, 1)])
def test_fine_kick_reproduces_a_linear_force(izipx, izipv):
    nc, nnt = 16, 2
    st, oc = _state(nc, nnt, izipx, izipv)
    nt, nfb, nft = nc // nnt, oc.nfb, oc.nft
    a_mid, dt = np.float32(0.05), np.float32(3.0)
    s0 = np.float32(st["sigma_vi"])
    s1 = np.float32(s0 * np.float32(1.25))
    # tile array force_f(3, nfb:nfb+nft+1, ...) as [z][y][x][d]; node n sits
    # at tile-local fine coordinate n - nfb - 1
    n = np.arange(nft + 2, dtype=np.float64) - 1.0
    Z, Y, X = np.meshgrid(n, n, n, indexing="ij")
    P = np.stack([X, Y, Z], axis=-1)
    F = (C + P @ G.T).astype(np.float32)
    F = np.ascontiguousarray(F)
    cid, ijk, txyz = _cells(st, nc, nnt)
    off = _offsets(st, izipx)
    x_loc = (4.0 * ijk + 4.0 * off).astype(np.float32).astype(np.float64)  # tempx, real(4)
    v0 = _decode(izipv, st["vp"], s0)
    dv = (C + x_loc @ G.T) * float(a_mid) * float(dt) / (6.0 * math.pi)
    want = _codes(izipv, v0 + dv, s1)
    for itz in range(1, nnt + 1):
        for ity in range(1, nnt + 1):
            for itx in range(1, nnt + 1):
                O.lib().oc_fine_kick(oc._c, itx, ity, itz, O._p(F), a_mid, dt, s0, s1)
    got = oc.get_state()["vp"]
    _compare(izipv, got, want)


@pytest.mark.parametrize("izipx,izipv", [(2, 2), (1, 2)])
def test_coarse_kick_and_vmax(izipx, izipv):
    nc, nnt = 16, 2
    st, oc = _state(nc, nnt, izipx, izipv, seed=12)
    nt = nc // nnt
    a_mid, dt = np.float32(0.05), np.float32(3.0)
    s = np.float32(st["sigma_vi"])
    # force_c(3, 0:nc+1, 0:nc+1, 0:nc+1) as [z][y][x][d]; node n at coarse
    # coordinate n - 1 (cell-centred CIC: the particle coordinate is x - 0.5)
    n = np.arange(nc + 2, dtype=np.float64) - 1.0
    Z, Y, X = np.meshgrid(n, n, n, indexing="ij")
    P = np.stack([X, Y, Z], axis=-1)
    F = np.ascontiguousarray((10 * C + 4 * P @ G.T).astype(np.float32))
    cid, ijk, txyz = _cells(st, nc, nnt)
    off = _offsets(st, izipx)
    xg = ((txyz * nt + ijk).astype(np.float64) + off - 0.5).astype(np.float32).astype(np.float64)
    v0 = _decode(izipv, st["vp"], s)
    dv = (10 * C + 4 * xg @ G.T) * float(a_mid) * float(dt) / (6.0 * math.pi)
    vreal = v0 + dv
    want = _codes(izipv, vreal, s)
    vmax = np.zeros(3, np.float32)
    O.lib().oc_coarse_kick(oc._c, O._p(F), a_mid, dt, s, O._p(vmax))
    got = oc.get_state()["vp"]
    _compare(izipv, got, want)
    # brute-force v_max = max_p |v_p + v_c(cell p)| per component (W/pm.f90:349)
    vm = np.abs(vreal + st["vfield"][cid].astype(np.float64)).max(axis=0)
    assert np.allclose(vmax, vm, rtol=1e-5, atol=0), (vmax, vm)
    # ... and not max |v| or max |v_c| alone
    assert not np.allclose(vm, np.abs(vreal).max(axis=0), rtol=1e-4)


def test_dt_constraints_closed_form():
    """W/pm.f90:395-398: dt_fine = (1/(|F|max a GG))^(1/2), dt_coarse =
    (ncell/(|F_c|max a GG))^(1/2), GG = 1/(6 pi), dt_vmax = 0.9*20/max v_max,
    vz_max = v_max(3); here against the FP64 closed forms."""
    oc = O.OracleCube(16, 1)
    rng = np.random.default_rng(5)
    for _ in range(100):
        f2f, f2c = (float(np.float32(10 ** rng.uniform(-3, 4))) for _ in range(2))
        a = float(np.float32(rng.uniform(0.02, 1.0)))
        vm = rng.uniform(0.1, 50, 3).astype(np.float32)
        r = oc.dt_constraints(dict(f2_max_fine=f2f, f2_max_coarse=f2c, vmax=vm), a)
        assert abs(r["dt_fine"] / math.sqrt(6 * math.pi / (math.sqrt(f2f) * a)) - 1) < 1e-6
        assert abs(r["dt_coarse"] / math.sqrt(4 * 6 * math.pi / (math.sqrt(f2c) * a)) - 1) < 1e-6
        assert abs(r["dt_vmax"] / (18.0 / float(vm.max())) - 1) < 1e-6
        assert r["vz_max"] == float(vm[2])
