"""Pin the oracle's buffered-layout machinery (spine_image, buffer_grid,
buffer_xvp, update_xp, the deposits) against brute-force restatements that
work on plain per-particle arrays with no tiles, spines or buffer copies."""
import math

import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import zeldovich_state

CASES = [dict(nc=16, nnt=1), dict(nc=24, nnt=2)]


def _ids(c):
    return f"nc{c['nc']}_nnt{c['nnt']}"


def _vtable(izipv, sigma):
    n = 1 << (8 * izipv)
    return np.array([O.decode_v(izipv, c, sigma) for c in range(-(n // 2), n // 2)])


def _decode_state(st, nc, nnt, izipx=2):
    """Global positions (coarse units, FP64) and per-particle cell indices."""
    nt = nc // nnt
    rhoc = st["rhoc"]
    cid = np.repeat(np.arange(nc ** 3), rhoc)
    i = cid % nt
    j = (cid // nt) % nt
    k = (cid // nt ** 2) % nt
    t = cid // nt ** 3
    tx, ty, tz = t % nnt, (t // nnt) % nnt, t // nnt ** 2
    g = np.stack([tx * nt + i, ty * nt + j, tz * nt + k], axis=1)
    n = 1 << (8 * izipx)
    off = ((st["xp"].astype(np.int64) & (n - 1)) + 0.5) / n
    return g, off, cid


@pytest.mark.parametrize("case", CASES, ids=_ids)
def test_buffers_mirror_periodic_neighbours(case):
    """After buffer_grid/buffer_x/buffer_v every buffer cell of every tile
    holds exactly the particles (same order) of its periodic image."""
    nc, nnt = case["nc"], case["nnt"]
    st = zeldovich_state(nc, nnt)
    oc = O.OracleCube(nc, nnt)
    oc.set_state(st["xp"], st["vp"], st["pid"], st["rhoc"], st["vfield"])
    oc.buffer_grid(); oc.buffer_x(); oc.buffer_v()
    L = O.lib()
    nt, ncb, nte = oc.nt, oc.ncb, oc.nte
    nb = nte ** 3 * nnt ** 3
    rhoc = np.ctypeslib.as_array(
        (O.ctypes.c_int32 * nb).from_address(L.oc_rhoc_ptr(oc._c))).reshape(nnt, nnt, nnt, nte, nte, nte)
    idx_b_r = np.ctypeslib.as_array((O.ctypes.c_int64 * (nte * nte * nnt ** 3)).from_address(
        L.oc_spine_ptr(oc._c, 1))).reshape(nnt, nnt, nnt, nte, nte)
    npmax = oc.np_image_max
    xp = np.ctypeslib.as_array((O.ctypes.c_int16 * (3 * npmax)).from_address(L.oc_xp_ptr(oc._c))).reshape(-1, 3)
    pid = np.ctypeslib.as_array((O.ctypes.c_int32 * npmax).from_address(L.oc_pid_ptr(oc._c)))
    start = np.concatenate([[0], np.cumsum(st["rhoc"])])

    def phys_particles(gx, gy, gz):
        gx, gy, gz = gx % nc, gy % nc, gz % nc
        c = (gx % nt) + nt * ((gy % nt) + nt * ((gz % nt) + nt * (
            (gx // nt) + nnt * ((gy // nt) + nnt * (gz // nt)))))
        return st["pid"][start[c]:start[c + 1]]

    rng = np.random.default_rng(0)
    for _ in range(400):
        tz, ty, tx = rng.integers(0, nnt, 3)
        i, j, k = rng.integers(1 - ncb, nt + ncb + 1, 3)
        b = 1 - ncb
        np_ = rhoc[tz, ty, tx, k - b, j - b, i - b]
        nzero = idx_b_r[tz, ty, tx, k - b, j - b] - rhoc[tz, ty, tx, k - b, j - b, i - b:].sum()
        got = pid[nzero:nzero + np_]
        want = phys_particles(tx * nt + i - 1, ty * nt + j - 1, tz * nt + k - 1)
        assert np.array_equal(got, want)


@pytest.mark.parametrize("case", CASES, ids=_ids)
@pytest.mark.parametrize("dt", [30.0, 70.0])
def test_update_x_against_brute_force(case, dt):
    """update_xp (W/update_particle.f90:16-256) vs a per-particle restatement:
    new cell = ceiling(x+dt_mid v/ncell), order inside a cell =
    (ilayer,k,j,i,l) of the copy in the destination tile's frame,
    vfield_new=(sum v + 0.1 v_old)/(n+0.1), v codes relative to it,
    sigma_vi_new = rms residual/sqrt(3), x codes by integer wrap."""
    nc, nnt = case["nc"], case["nnt"]
    nt, ncb = nc // nnt, 6
    st = zeldovich_state(nc, nnt, seed=5)
    sig = st["sigma_vi"]
    vz = float(np.abs(st["vfield"][:, 2]).max() * 1.5 + 0.02)
    oc = O.OracleCube(nc, nnt)
    oc.set_state(st["xp"], st["vp"], st["pid"], st["rhoc"], st["vfield"])
    oc.buffer_grid(); oc.buffer_x(); oc.buffer_v()
    sig_new, sums = oc.update_x(dt, dt, vz, sig)
    new = oc.get_state()
    # brute force on the old state
    g, off, cid = _decode_state(st, nc, nnt)
    vt = _vtable(2, sig)
    v = vt[st["vp"].astype(np.int64) + 32768] + st["vfield"][cid].astype(np.float64)
    dt_mid = np.float32((np.float32(dt) + np.float32(dt)) / np.float32(2))
    nlayer = 2 * math.ceil(np.float32(dt_mid * np.float32(vz)) / np.float32(4)) + 1
    dx = (float(dt_mid) * v) / 4.0
    # home-frame evaluation (copies in other frames differ only on ties)
    gl = g % nt + 1
    ig = np.ceil(((gl - 1.0) + off) + dx).astype(np.int64)
    gnew = (g - gl + ig) % nc
    tile = gnew // nt
    tnew = tile[:, 0] + nnt * (tile[:, 1] + nnt * tile[:, 2])
    loc = gnew % nt
    cnew = loc[:, 0] + nt * (loc[:, 1] + nt * (loc[:, 2] + nt * tnew))
    assert np.array_equal(np.bincount(cnew, minlength=nc ** 3), new["rhoc"])
    # order key: source coordinate in the destination frame = dest local - displacement
    disp = ig - gl
    src_local = (loc + 1) - disp
    ilayer = (src_local[:, 2] - (1 - ncb)) % nlayer
    start = np.concatenate([[0], np.cumsum(st["rhoc"])])
    lsrc = np.arange(len(cid)) - start[cid]
    key = np.lexsort((lsrc, src_local[:, 0], src_local[:, 1], src_local[:, 2], ilayer, cnew))
    assert np.array_equal(st["pid"][key], new["pid"])
    # positions: integer wrap of the code
    q = (float(dt_mid) * v) / (4.0 / 65536.0)
    nq = np.trunc(q) + np.where(np.abs(q - np.trunc(q)) >= 0.5, np.sign(q), 0)
    xnew = ((st["xp"].astype(np.int64) + nq.astype(np.int64) + 32768) % 65536) - 32768
    assert np.array_equal(xnew[key], new["xp"])
    # vfield_new
    n_new = new["rhoc"].astype(np.float64)
    vsum = np.stack([np.bincount(cnew, weights=v[:, d], minlength=nc ** 3) for d in range(3)], 1)
    wv = float(np.float32(0.1))
    want = (vsum + st["vfield"].astype(np.float64) * wv) / (n_new[:, None] + wv)
    assert np.allclose(new["vfield"], want, rtol=2e-6, atol=1e-9)
    # velocity codes decode back to v - vfield_new within one code
    cnew_sorted = cnew[key]
    vres = vt[new["vp"].astype(np.int64) + 32768]
    err = np.abs(vres + new["vfield"][cnew_sorted] - v[key])
    step = np.abs(vt[np.clip(new["vp"].astype(np.int64) + 32769, 0, 65535)] - vres)
    assert np.all(err <= step + 1e-12)
    # statistics
    vr_new = vt[new["vp"].astype(np.int64) + 32768]
    rms = math.sqrt((vr_new ** 2).sum() / len(vr_new))
    assert abs(sig_new / (rms / math.sqrt(np.float32(3))) - 1) < 1e-6


@pytest.mark.parametrize("case", CASES, ids=_ids)
def test_deposits_against_brute_force(case):
    """Fine (W/pm.f90:50-108) and coarse (:205-267) CIC deposits vs a float64
    scatter of the decoded particles on the global periodic meshes; and the
    masses (testrhof, testrhoc) equal N * mass_p."""
    nc, nnt = case["nc"], case["nnt"]
    st = zeldovich_state(nc, nnt, seed=9)
    m = st["mass_p"]
    oc = O.OracleCube(nc, nnt)
    oc.set_state(st["xp"], st["vp"], st["pid"], st["rhoc"], st["vfield"])
    oc.buffer_grid(); oc.buffer_x(); oc.buffer_v()
    g, off, _ = _decode_state(st, nc, nnt)
    pos = g + off
    for fine in (True, False):
        n = 4 * nc if fine else nc
        x = pos * 4 if fine else pos - 0.5
        i1 = np.floor(x).astype(np.int64)
        d2 = x - i1
        d1 = 1 - d2
        rho = np.zeros((n, n, n))
        for cz in (0, 1):
            for cy in (0, 1):
                for cx in (0, 1):
                    w = (d2[:, 0] if cx else d1[:, 0]) * (d2[:, 1] if cy else d1[:, 1]) * (d2[:, 2] if cz else d1[:, 2])
                    np.add.at(rho, ((i1[:, 2] + cz) % n, (i1[:, 1] + cy) % n, (i1[:, 0] + cx) % n), w * m)
        if fine:
            nft, nfb = oc.nft, oc.nfb
            got = np.zeros_like(rho)
            tot = 0.0
            for itz in range(1, nnt + 1):
                for ity in range(1, nnt + 1):
                    for itx in range(1, nnt + 1):
                        r, s = oc.fine_deposit(itx, ity, itz, m)
                        tot += s
                        got[(itz - 1) * nft:itz * nft, (ity - 1) * nft:ity * nft, (itx - 1) * nft:itx * nft] = \
                            r[nfb:nfb + nft, nfb:nfb + nft, nfb:nfb + nft]
        else:
            got, tot = oc.coarse_deposit(m)
        assert np.allclose(got, rho, rtol=1e-5, atol=1e-5 * m)
        assert abs(tot / (len(pos) * m) - 1) < 1e-6


def test_single_particle_cic_closed_form():
    """One particle at fine position (x,y,z): the 8 cells get the products of
    (1-f, f) along each axis; corner convention of the fine mesh and -0.5
    cell-centring of the coarse one (W/pm.f90:63,221)."""
    nc, nnt = 16, 1
    oc = O.OracleCube(nc, nnt)
    pos = np.array([[5.3, 7.71, 11.02]])   # coarse units
    oc.state_from_xv(pos, np.zeros((1, 3)), None, 1.0)
    oc.buffer_grid(); oc.buffer_x(); oc.buffer_v()
    st = oc.get_state()
    off = (st["xp"][0].astype(np.int64) % 65536 + 0.5) / 65536
    p = np.floor(pos[0]) + off
    r, _ = oc.fine_deposit(1, 1, 1, 1.0)
    nfb = oc.nfb
    fx = 4 * p
    i1 = np.floor(fx).astype(int)
    f = fx - i1
    for cz in (0, 1):
        for cy in (0, 1):
            for cx in (0, 1):
                w = (f[0] if cx else 1 - f[0]) * (f[1] if cy else 1 - f[1]) * (f[2] if cz else 1 - f[2])
                got = r[nfb + i1[2] + cz, nfb + i1[1] + cy, nfb + i1[0] + cx]
                assert abs(got - w) < 2e-6
    nft = oc.nft
    assert abs(r[nfb:nfb + nft, nfb:nfb + nft, nfb:nfb + nft].sum() - 1) < 1e-6
    r3, s = oc.coarse_deposit(1.0)
    xc = p - 0.5
    j1 = np.floor(xc).astype(int)
    fc = xc - j1
    for cz in (0, 1):
        for cy in (0, 1):
            for cx in (0, 1):
                w = (fc[0] if cx else 1 - fc[0]) * (fc[1] if cy else 1 - fc[1]) * (fc[2] if cz else 1 - fc[2])
                assert abs(r3[(j1[2] + cz) % nc, (j1[1] + cy) % nc, (j1[0] + cx) % nc] - w) < 2e-6


def test_displacement_past_buffer_is_an_error():
    """A drift longer than the double buffer of rhoce is out of bounds in
    the reference; the oracle reports it as its own status."""
    nc, nnt = 16, 1
    oc = O.OracleCube(nc, nnt)
    pos = np.array([[5.3, 7.7, 11.0], [9.1, 2.2, 3.3]])
    vel = np.array([[400.0, 0, 0], [0, 0, 0]])
    oc.state_from_xv(pos, vel, None, 1.0)
    oc.buffer_grid(); oc.buffer_x(); oc.buffer_v()
    with pytest.raises(O.OracleError) as e:
        oc.update_x(1.0, 1.0, 0.0, 1.0)
    assert e.value.code == 6


def test_tile_overflow_is_an_error():
    """W/update_particle.f90:88-94: a tile whose extended region holds more
    than np_tile_max particles stops the run."""
    nc, nnt = 16, 2
    oc = O.OracleCube(nc, nnt, np_nc=1, tile_buffer=0.05)
    rng = np.random.default_rng(1)
    pos = rng.uniform(0, nc, (4096, 3))
    oc.state_from_xv(pos, np.zeros_like(pos), None, 1.0)
    oc.buffer_grid(); oc.buffer_x(); oc.buffer_v()
    with pytest.raises(O.OracleError) as e:
        oc.update_x(1.0, 1.0, 0.0, 1.0)
    assert e.value.code == 1


def test_parallel_oracle_equals_serial():
    """The oracle's OpenMP loops (the reference's layered slabs, cells of the
    kicks and PP) give the same bits on 1 and on several threads through a
    full step: drift, PP, both deposits and kicks."""
    from tests.helpers import zeldovich_state
    nc, nnt = 24, 2
    st = zeldovich_state(nc, nnt, seed=21, izipx=1, izipv=2)
    out = []
    nthr = O.lib().oc_get_threads()
    try:
        for threads in (1, 4):
            O.lib().oc_set_threads(threads)
            oc = O.OracleCube(nc, nnt, izipx=1, izipv=2, pp_on=True)
            oc.set_state(st["xp"], st["vp"], st["pid"], st["rhoc"], st["vfield"])
            oc.init_kernels()
            oc.buffer_grid(); oc.buffer_x(); oc.buffer_v()
            vz = float(np.abs(st["vfield"][:, 2]).max() * 1.5 + 0.05)
            s2, stats = oc.update_x(20.0, 20.0, vz, st["sigma_vi"])
            oc.buffer_grid(); oc.buffer_x()
            f2 = oc.ext_pp_force(st["a"], 20.0, st["sigma_vi"], st["mass_p"])
            pm = oc.particle_mesh(st["a"], 20.0, st["sigma_vi"], s2, st["mass_p"])
            out.append((oc.get_state(), s2, stats, f2, pm["vmax"].copy(), pm["testrhof"], pm["testrhoc"]))
    finally:
        O.lib().oc_set_threads(nthr)
    a, b = out
    for k in ("xp", "vp", "pid", "rhoc", "vfield"):
        assert np.array_equal(a[0][k], b[0][k]), k
    assert a[1] == b[1] and np.array_equal(a[2], b[2]) and a[3] == b[3]
    assert np.array_equal(a[4], b[4]) and a[5] == b[5] and a[6] == b[6]
