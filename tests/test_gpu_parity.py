"""GPU parity of each stage of the step against the oracle, on identical
inputs (tests/helpers.zeldovich_state).  Integer results (cells, counts,
order, codes) must be bit-exact; deposits too (the GPU gathers in the
reference's own summation order); forces within 1e-5 relative L2 in FP32."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import zeldovich_state, rel_l2, global_force_to_tile, global_from_tiles

pytestmark = pytest.mark.gpu

CASES = [dict(nc=24, nnt=2), dict(nc=16, nnt=1), dict(nc=24, nnt=1, izipx=1, izipv=2),
         dict(nc=16, nnt=2, izipx=2, izipv=1)]


def _ids(c):
    return "nc{nc}_nnt{nnt}_x{ix}v{iv}".format(ix=c.get("izipx", 2), iv=c.get("izipv", 2), **c)


def _pair(case, pp_on=False, seed=42):
    from cafproject_amd.cube import Cube
    ix, iv = case.get("izipx", 2), case.get("izipv", 2)
    st = zeldovich_state(case["nc"], case["nnt"], seed=seed, izipx=ix, izipv=iv)
    oc = O.OracleCube(case["nc"], case["nnt"], izipx=ix, izipv=iv, pp_on=pp_on)
    oc.set_state(st["xp"], st["vp"], st["pid"], st["rhoc"], st["vfield"])
    oc.buffer_grid(); oc.buffer_x(); oc.buffer_v()
    g = Cube(case["nc"], case["nnt"], izipx=ix, izipv=iv, pp_on=pp_on)
    g.set_state(st["xp"], st["vp"], st["pid"], st["rhoc"], st["vfield"])
    return st, oc, g


def _same_state(a, b):
    for k in ("rhoc", "xp", "vp", "pid", "vfield"):
        assert np.array_equal(a[k], b[k]), k


@pytest.mark.parametrize("case", CASES, ids=_ids)
def test_set_get_roundtrip(case):
    st, oc, g = _pair(case)
    _same_state(g.get_state(), oc.get_state())


@pytest.mark.parametrize("case", CASES, ids=_ids)
@pytest.mark.parametrize("dt", [1.0, 25.0, 70.0])
def test_update_x_bit_exact(case, dt):
    st, oc, g = _pair(case)
    vz = float(np.abs(st["vfield"][:, 2]).max() * 1.5 + 0.02)
    s_o, stats = oc.update_x(dt, dt, vz, st["sigma_vi"])
    s_g = g.update_x(dt, dt, vz, st["sigma_vi"])
    _same_state(g.get_state(), oc.get_state())
    assert s_g["nplocal"] == oc.nplocal
    assert abs(s_g["sum_r"] / stats[1] - 1) < 1e-12
    assert abs(s_g["sigma_vi_new"] / s_o - 1) < 1e-6


@pytest.mark.parametrize("case", CASES[:2], ids=_ids)
@pytest.mark.parametrize("agg_cap", [None, "0", "300"])
def test_two_drifts_bit_exact(case, agg_cap, monkeypatch):
    """The drift takes slots in new cells through LDS counters per source row
    (rows up to CUBE_DRIFT_AGG_CAP particles) or global atomics; all-LDS,
    all-global and mixed runs must match the oracle bit for bit."""
    if agg_cap is not None:
        monkeypatch.setenv("CUBE_DRIFT_AGG_CAP", agg_cap)
    st, oc, g = _pair(case)
    vz = float(np.abs(st["vfield"][:, 2]).max() * 1.5 + 0.02)
    sig = st["sigma_vi"]
    for dt in (40.0, 55.0):
        s_o, _ = oc.update_x(dt, dt, vz, sig)
        g.update_x(dt, dt, vz, sig)
        oc.buffer_grid(); oc.buffer_x(); oc.buffer_v()
        g.buffer_grid(); g.buffer_x(); g.buffer_v()
        sig = s_o
    _same_state(g.get_state(), oc.get_state())


# nt = 9 takes the coarse deposit's per-cell kernel, even nt its 2x2x2-block one
DEP_CASES = CASES + [dict(nc=20, nnt=2), dict(nc=18, nnt=2, izipx=1, izipv=2)]


@pytest.mark.parametrize("case", DEP_CASES, ids=_ids)
@pytest.mark.parametrize("heavy", [None, "0"])
def test_fine_and_coarse_deposit_bit_exact(case, heavy, monkeypatch):
    """Deposits bit for bit; with CUBE_CD_HEAVY=0 and CUBE_FD_HEAVY=0 every
    coarse 2^3 block and every fine-deposit coarse cell is formed by its
    whole wave (the path of clustered regions)."""
    if heavy is not None:
        monkeypatch.setenv("CUBE_CD_HEAVY", heavy)
        monkeypatch.setenv("CUBE_FD_HEAVY", heavy)
    st, oc, g = _pair(case)
    m = st["mass_p"]
    tf = g.fine_deposit(m)
    rho_g = g.get_field("rho_f")
    nc, nnt = case["nc"], case["nnt"]
    nfb, nfe, nft = oc.nfb, oc.nfe, oc.nft
    tiles = {}
    s_o = 0.0
    for itz in range(1, nnt + 1):
        for ity in range(1, nnt + 1):
            for itx in range(1, nnt + 1):
                rho, s = oc.fine_deposit(itx, ity, itz, m)
                s_o += s
                tiles[(itx, ity, itz)] = dict(rho=rho[nfb - 1:nfb + nft + 1, nfb - 1:nfb + nft + 1,
                                                      nfb - 1:nfb + nft + 1])
    rho_o = global_from_tiles(tiles, nc, nnt, key="rho")
    assert np.array_equal(rho_g, rho_o)
    assert abs(tf / s_o - 1) < 1e-12
    tc = g.coarse_deposit(m)
    r3_o, sc = oc.coarse_deposit(m)
    assert np.array_equal(g.get_field("rho_c"), r3_o)
    assert abs(tc / sc - 1) < 1e-12


FORCE_CASES = [dict(nc=24, nnt=2), dict(nc=24, nnt=1, izipx=1, izipv=2)]


@pytest.mark.parametrize("case", FORCE_CASES, ids=_ids)
def test_forces_within_1e5(case):
    st, oc, g = _pair(case)
    m = st["mass_p"]
    oc.init_kernels()
    g.init_kernels()
    pm = oc.particle_mesh(0.02, 0.0, st["sigma_vi"], st["sigma_vi"], m, keep_fields=True)
    g.fine_deposit(m)
    f2 = g.fine_force()
    fo = global_from_tiles(pm["tiles"], case["nc"], case["nnt"])
    fg = g.get_field("force_f")
    assert rel_l2(fg, fo) < 1e-5, rel_l2(fg, fo)
    assert abs(f2 / pm["f2_max_fine"] - 1) < 1e-4
    g.coarse_deposit(m)
    f2c = g.coarse_force()
    fc_o = pm["force_c"][1:-1, 1:-1, 1:-1]
    assert rel_l2(g.get_field("force_c"), fc_o) < 1e-5
    assert abs(f2c / pm["f2_max_coarse"] - 1) < 1e-4


@pytest.mark.parametrize("case", CASES, ids=_ids)
def test_kicks_bit_exact_given_forces(case):
    st, oc, g = _pair(case)
    nc, nnt = case["nc"], case["nnt"]
    rng = np.random.default_rng(7)
    ff = rng.standard_normal((4 * nc,) * 3 + (3,)).astype(np.float32) * 3
    g.set_field("force_f", ff)
    a_mid, dt, s0, s1 = 0.03, 2.0, st["sigma_vi"], st["sigma_vi"] * 1.1
    g.fine_kick(a_mid, dt, s0, s1)
    for itz in range(1, nnt + 1):
        for ity in range(1, nnt + 1):
            for itx in range(1, nnt + 1):
                tile = global_force_to_tile(ff, nc, nnt, itx, ity, itz)
                O.lib().oc_fine_kick(oc._c, itx, ity, itz, O._p(tile), a_mid, dt, s0, s1)
    _same_state(g.get_state(), oc.get_state())
    fc = rng.standard_normal((nc,) * 3 + (3,)).astype(np.float32) * 50
    g.set_field("force_c", fc)
    vg = g.coarse_kick(a_mid, dt, s1)
    halo = np.ascontiguousarray(np.pad(fc, ((1, 1), (1, 1), (1, 1), (0, 0)), mode="wrap"))
    vo = np.zeros(3, np.float32)
    O.lib().oc_coarse_kick(oc._c, O._p(halo), a_mid, dt, s1, O._p(vo))
    _same_state(g.get_state(), oc.get_state())
    assert np.array_equal(vg, vo)


@pytest.mark.parametrize("case", CASES, ids=_ids)
@pytest.mark.parametrize("stage_max,cell_cap", [(None, None), ("0", None), ("40", "16"), ("0", "1"), ("40", "200")])
def test_pp_bit_exact(case, stage_max, cell_cap, monkeypatch):
    """PP stages each 4^3-cell block's neighbourhood in LDS; blocks with more
    than CUBE_PP_STAGE_MAX particles there go to the cell path (work items of
    one coarse cell and up to 256 of its particles), which stages its
    neighbourhood CUBE_PP_CELL_CAP particles at a time.  Every mix of the
    paths and window sizes must match the oracle bit for bit."""
    if stage_max is not None:
        monkeypatch.setenv("CUBE_PP_STAGE_MAX", stage_max)
    if cell_cap is not None:
        monkeypatch.setenv("CUBE_PP_CELL_CAP", cell_cap)
    st, oc, g = _pair(case, pp_on=True)
    m = st["mass_p"]
    f_o = oc.ext_pp_force(0.02, 3.0, st["sigma_vi"], m)
    f_g = g.ext_pp_force(0.02, 3.0, st["sigma_vi"], m)
    _same_state(g.get_state(), oc.get_state())
    assert f_g == f_o


def _clump_state(nc=16, seed=5, nclump=6000, nrest=10000):
    """A collapsed clump (Gaussian, sigma 0.7 coarse cells) in a uniform
    background, as continuous positions and velocities."""
    rng = np.random.default_rng(seed)
    clump = np.array([7.3, 8.1, 8.6]) + rng.normal(0, 0.7, (nclump, 3))
    rest = rng.uniform(0, nc, (nrest, 3))
    pos = np.concatenate([clump, rest]) % nc
    vel = rng.normal(0, 0.3, (nclump + nrest, 3)).astype(np.float32)
    return pos, vel


def test_drift_clustered_bit_exact():
    """The drift of the clump state: its cells of 49-1024 particles take the
    wave-per-cell sort, the few below and above the small and workgroup
    sorts; state after two drifts bit for bit against the oracle."""
    from cafproject_amd.cube import Cube
    nc, nnt = 16, 1
    pos, vel = _clump_state(nc)
    sig = 0.3
    g = Cube(nc, nnt)
    g.set_state_xv(pos, vel, sig)
    oc = O.OracleCube(nc, nnt)
    oc.state_from_xv(pos, vel.astype(np.float64), None, sig)
    rh = oc.get_state()["rhoc"]
    assert ((rh > 48) & (rh <= 1024)).sum() > 20
    oc.buffer_grid(); oc.buffer_x(); oc.buffer_v()
    for dt in (1.5, 2.5):
        oc.update_x(dt, dt, 3.0, sig)
        g.update_x(dt, dt, 3.0, sig)
        oc.buffer_grid(); oc.buffer_x(); oc.buffer_v()
        g.buffer_grid(); g.buffer_x(); g.buffer_v()
        _same_state(g.get_state(), oc.get_state())


@pytest.mark.parametrize("heavy", ["0", "40", "2048", "100000000"])
def test_deposits_clustered_bit_exact(heavy, monkeypatch):
    """The clump state: fine and coarse meshes bit for bit against the
    oracle, with the clustered coarse blocks and fine-deposit cells formed
    by their waves always (0), often (40), at 2048 source particles and
    never; the totals are the same numbers whichever lanes gathered."""
    from cafproject_amd.cube import Cube
    nc, nnt = 16, 1
    pos, vel = _clump_state(nc)
    sig = 0.3
    g = Cube(nc, nnt)
    g.set_state_xv(pos, vel, sig)
    oc = O.OracleCube(nc, nnt)
    oc.state_from_xv(pos, vel.astype(np.float64), None, sig)
    oc.buffer_grid(); oc.buffer_x(); oc.buffer_v()
    m = float(np.float32((4 * nc) ** 3 / len(pos)))
    tc_default = g.coarse_deposit(m)
    tf_default = g.fine_deposit(m)
    monkeypatch.setenv("CUBE_CD_HEAVY", heavy)
    monkeypatch.setenv("CUBE_FD_HEAVY", heavy)
    tc = g.coarse_deposit(m)
    r3_o, sc = oc.coarse_deposit(m)
    assert np.array_equal(g.get_field("rho_c"), r3_o)
    assert tc == tc_default
    assert abs(tc / sc - 1) < 1e-12
    assert g.fine_deposit(m) == tf_default
    rho, _ = oc.fine_deposit(1, 1, 1, m)
    nfb, nft = oc.nfb, oc.nft
    tiles = {(1, 1, 1): dict(rho=rho[nfb - 1:nfb + nft + 1, nfb - 1:nfb + nft + 1, nfb - 1:nfb + nft + 1])}
    assert np.array_equal(g.get_field("rho_f"), global_from_tiles(tiles, nc, nnt, key="rho"))


@pytest.mark.parametrize("stage_max,cell_cap,heavy",
                         [(None, None, None), ("0", None, None), (None, "300", None), (None, None, "0"),
                          (None, None, "16")])
def test_pp_clustered_bit_exact(stage_max, cell_cap, heavy, monkeypatch):
    """A collapsed clump (6000 particles, Gaussian with sigma 0.7 coarse
    cells) in a uniform background of 10000: the blocks around it go to the
    cell path, its coarse cells hold several 256-particle items each, and
    their neighbourhoods hold more than one staging window; the block
    kernel's staged cells above CUBE_PP_HEAVY particles are counted and
    placed by their whole wave (0: every cell).  Velocities and max |f|^2
    match the oracle bit for bit."""
    from cafproject_amd.cube import Cube
    if stage_max is not None:
        monkeypatch.setenv("CUBE_PP_STAGE_MAX", stage_max)
    if cell_cap is not None:
        monkeypatch.setenv("CUBE_PP_CELL_CAP", cell_cap)
    if heavy is not None:
        monkeypatch.setenv("CUBE_PP_HEAVY", heavy)
    nc, nnt = 16, 1
    pos, vel = _clump_state(nc)
    sig = 0.3
    g = Cube(nc, nnt, pp_on=True)
    g.set_state_xv(pos, vel, sig)
    oc = O.OracleCube(nc, nnt, pp_on=True)
    oc.state_from_xv(pos, vel.astype(np.float64), None, sig)
    oc.buffer_grid(); oc.buffer_x(); oc.buffer_v()
    so = oc.get_state()
    _same_state(g.get_state(), so)
    assert so["rhoc"].max() > 512  # coarse cells of three or more items
    m = float(np.float32((4 * nc) ** 3 / 16000))
    f_o = oc.ext_pp_force(0.05, 2.0, sig, m)
    f_g = g.ext_pp_force(0.05, 2.0, sig, m)
    _same_state(g.get_state(), oc.get_state())
    assert f_g == f_o


@pytest.mark.parametrize("case", CASES, ids=_ids)
@pytest.mark.parametrize("heavy", ["0", "6"])
def test_pp_wave_staging_bit_exact(case, heavy, monkeypatch):
    """The block kernel's staged coarse cells above CUBE_PP_HEAVY particles
    are counted and placed by their whole wave, 64 particles at a time, in
    the chain order of the one-thread walk: with every cell (0) or the
    fuller ones (6) on that path the forces match the oracle bit for bit."""
    monkeypatch.setenv("CUBE_PP_HEAVY", heavy)
    st, oc, g = _pair(case, pp_on=True)
    m = st["mass_p"]
    f_o = oc.ext_pp_force(0.02, 3.0, st["sigma_vi"], m)
    f_g = g.ext_pp_force(0.02, 3.0, st["sigma_vi"], m)
    _same_state(g.get_state(), oc.get_state())
    assert f_g == f_o


@pytest.mark.parametrize("case", CASES[:2], ids=_ids)
def test_pp_three_divisions_bit_exact(case, monkeypatch):
    """The pair force's three IEEE divisions x/r^3 (CUBE_PP_FASTDIV=0) and
    the default single division with corrected products give the oracle's
    bits alike."""
    monkeypatch.setenv("CUBE_PP_FASTDIV", "0")
    st, oc, g = _pair(case, pp_on=True)
    m = st["mass_p"]
    f_o = oc.ext_pp_force(0.02, 3.0, st["sigma_vi"], m)
    f_g = g.ext_pp_force(0.02, 3.0, st["sigma_vi"], m)
    _same_state(g.get_state(), oc.get_state())
    assert f_g == f_o


@pytest.mark.parametrize("pp_range", [0, 2])
@pytest.mark.parametrize("stage_max,cell_cap", [(None, None), ("0", None), ("0", "5")])
def test_pp_other_ranges_bit_exact(pp_range, stage_max, cell_cap, monkeypatch):
    """pp_range 0 and 2: 1 and 25 (kk, jj) rows of 1 and 5 fine cells per
    particle; staged per block, on the cell path ("0"), and on the cell path
    in windows of 5 particles."""
    from cafproject_amd.cube import Cube
    if stage_max is not None:
        monkeypatch.setenv("CUBE_PP_STAGE_MAX", stage_max)
    if cell_cap is not None:
        monkeypatch.setenv("CUBE_PP_CELL_CAP", cell_cap)
    nc, nnt = 24, 2
    st = zeldovich_state(nc, nnt, seed=43)
    oc = O.OracleCube(nc, nnt, pp_on=True, pp_range=pp_range)
    oc.set_state(st["xp"], st["vp"], st["pid"], st["rhoc"], st["vfield"])
    oc.buffer_grid(); oc.buffer_x(); oc.buffer_v()
    g = Cube(nc, nnt, pp_on=True, pp_range=pp_range)
    g.set_state(st["xp"], st["vp"], st["pid"], st["rhoc"], st["vfield"])
    m = st["mass_p"]
    f_o = oc.ext_pp_force(0.02, 3.0, st["sigma_vi"], m)
    f_g = g.ext_pp_force(0.02, 3.0, st["sigma_vi"], m)
    _same_state(g.get_state(), oc.get_state())
    assert f_g == f_o


def test_coarse_force_small_box_is_finite():
    """nc < 24 is outside the reference's range (W/parameters.f90:29): its
    LRCKCORR divides by the zero Nyquist mode.  The library keeps those
    modes uncorrected, so the coarse force stays finite."""
    st, oc, g = _pair(dict(nc=16, nnt=1))
    g.init_kernels()
    g.coarse_deposit(st["mass_p"])
    f2 = g.coarse_force()
    assert np.isfinite(f2) and np.all(np.isfinite(g.get_field("force_c")))


@pytest.mark.parametrize("case", FORCE_CASES + [dict(nc=16, nnt=1, izipx=2, izipv=1)], ids=_ids)
def test_particle_mesh_outputs_bit_exact(case):
    """cube_particle_mesh's scalar outputs (W/pm.f90:126,315,349,395-398):
    f2_max_fine/coarse equal maxval(sum(F**2,1)) of the GPU's own force
    meshes evaluated on the host in the reference's order; the kicks and
    v_max equal the oracle's given those meshes; dt_fine, dt_coarse,
    dt_vmax and vz_max equal the oracle's dt_constraints of those values,
    bit for bit."""
    from cafproject_amd.cube import Cube
    st, oc, g = _pair(case)
    ix, iv = case.get("izipx", 2), case.get("izipv", 2)
    nc, nnt = case["nc"], case["nnt"]
    m = st["mass_p"]
    a_mid, dt = np.float32(0.03), np.float32(2.0)
    s0 = np.float32(st["sigma_vi"])
    s1 = np.float32(s0 * np.float32(1.1))
    # the GPU's force meshes, from the stage entry points on a twin context
    g2 = Cube(nc, nnt, izipx=ix, izipv=iv)
    g2.set_state(st["xp"], st["vp"], st["pid"], st["rhoc"], st["vfield"])
    g2.init_kernels()
    g2.fine_deposit(m)
    f2f = np.float32(g2.fine_force())
    ff = g2.get_field("force_f")
    g2.coarse_deposit(m)
    f2c = np.float32(g2.coarse_force())
    fc = g2.get_field("force_c")

    def f2max(f):
        f = f.astype(np.float32)
        return np.max((f[..., 0] * f[..., 0] + f[..., 1] * f[..., 1]) + f[..., 2] * f[..., 2])

    assert f2max(ff) == f2f and f2max(fc) == f2c
    g.init_kernels()
    pm = g.particle_mesh(float(a_mid), float(dt), float(s0), float(s1), m)
    assert np.float32(pm["f2_max_fine"]) == f2f
    assert np.float32(pm["f2_max_coarse"]) == f2c
    for itz in range(1, nnt + 1):
        for ity in range(1, nnt + 1):
            for itx in range(1, nnt + 1):
                tile = global_force_to_tile(ff, nc, nnt, itx, ity, itz)
                O.lib().oc_fine_kick(oc._c, itx, ity, itz, O._p(tile), a_mid, dt, s0, s1)
    halo = np.ascontiguousarray(np.pad(fc, ((1, 1), (1, 1), (1, 1), (0, 0)), mode="wrap"))
    vo = np.zeros(3, np.float32)
    O.lib().oc_coarse_kick(oc._c, O._p(halo), a_mid, dt, s1, O._p(vo))
    _same_state(g.get_state(), oc.get_state())
    assert np.array_equal(np.asarray(pm["vmax"], np.float32), vo)
    want = oc.dt_constraints(dict(f2_max_fine=f2f, f2_max_coarse=f2c, vmax=vo), a_mid)
    for k in ("dt_fine", "dt_coarse", "dt_vmax", "vz_max"):
        assert np.float32(pm[k]) == np.float32(want[k]), (k, pm[k], want[k])


# tile widths nt with nt % 4 != 0 take PP's per-row kernel (k_pp) instead of
# the 4^3-block one
PP_ROW_CASES = [dict(nc=18, nnt=2), dict(nc=20, nnt=2, izipx=1, izipv=2), dict(nc=12, nnt=2)]


@pytest.mark.parametrize("case", PP_ROW_CASES, ids=_ids)
@pytest.mark.parametrize("pp_range", [1, 2])
def test_pp_row_kernel_bit_exact(case, pp_range):
    from cafproject_amd.cube import Cube
    ix, iv = case.get("izipx", 2), case.get("izipv", 2)
    nc, nnt = case["nc"], case["nnt"]
    st = zeldovich_state(nc, nnt, seed=44, izipx=ix, izipv=iv)
    oc = O.OracleCube(nc, nnt, izipx=ix, izipv=iv, pp_on=True, pp_range=pp_range)
    oc.set_state(st["xp"], st["vp"], st["pid"], st["rhoc"], st["vfield"])
    oc.buffer_grid(); oc.buffer_x(); oc.buffer_v()
    g = Cube(nc, nnt, izipx=ix, izipv=iv, pp_on=True, pp_range=pp_range)
    g.set_state(st["xp"], st["vp"], st["pid"], st["rhoc"], st["vfield"])
    m = st["mass_p"]
    f_o = oc.ext_pp_force(0.02, 3.0, st["sigma_vi"], m)
    f_g = g.ext_pp_force(0.02, 3.0, st["sigma_vi"], m)
    _same_state(g.get_state(), oc.get_state())
    assert f_g == f_o


@pytest.mark.parametrize("case", [dict(nc=18, nnt=2), dict(nc=20, nnt=1, izipx=1, izipv=2)], ids=_ids)
def test_fine_force_other_mesh_sizes(case):
    """Fine meshes 72^3 and 80^3 (radix-3 and radix-5 line FFTs) within
    1e-5 of the oracle's per-tile FFT solve."""
    st, oc, g = _pair(case)
    m = st["mass_p"]
    oc.init_kernels()
    g.init_kernels()
    tiles = {}
    f2o = 0.0
    for itz in range(1, case["nnt"] + 1):
        for ity in range(1, case["nnt"] + 1):
            for itx in range(1, case["nnt"] + 1):
                rho, _ = oc.fine_deposit(itx, ity, itz, m)
                ff, f2 = O.fine_force(rho[:, :, :oc.nfe], oc.kern_f, oc.nfe, oc.nfb)
                f2o = max(f2o, f2)
                tiles[(itx, ity, itz)] = dict(force_f=ff)
    g.fine_deposit(m)
    f2g = g.fine_force()
    fo = global_from_tiles(tiles, case["nc"], case["nnt"])
    assert rel_l2(g.get_field("force_f"), fo) < 1e-5
    assert abs(f2g / f2o - 1) < 1e-4
