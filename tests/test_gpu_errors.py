"""Each reference `stop` condition reaches the caller as its own CUBE_ERR_*
code through the library, forced from the host with the parameters or the
time step, so every kernel still exits normally:

  tile overflow   W/update_particle.f90:88-94   (tile_buffer small)
  image overflow  W/buffer_grid.f90:148-154     (image_buffer small)
  PP overflow     W/ext_pp_force.f90:41-45      (np_pp_max = np_tile_max small)
  displacement    a particle past the rhoce double buffer (a long drift)

The oracle, given the same state, stops with the same status."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _codes(make_gpu, make_oracle, call_gpu, call_oracle):
    from cafproject_amd.cube import CubeError
    g = make_gpu()
    with pytest.raises(CubeError) as eg:
        call_gpu(g)
    oc = make_oracle()
    with pytest.raises(O.OracleError) as eo:
        call_oracle(oc)
    return eg.value.code, eo.value.code


def test_tile_overflow():
    from cafproject_amd.cube import Cube
    nc, nnt = 16, 2
    rng = np.random.default_rng(1)
    pos = rng.uniform(0, nc, (4096, 3))

    def mk_g():
        g = Cube(nc, nnt, np_nc=1, tile_buffer=0.05)
        g.set_state_xv(pos, np.zeros_like(pos, dtype=np.float32), 1.0)
        return g

    def mk_o():
        oc = O.OracleCube(nc, nnt, np_nc=1, tile_buffer=0.05)
        oc.state_from_xv(pos, np.zeros_like(pos), None, 1.0)
        oc.buffer_grid(); oc.buffer_x(); oc.buffer_v()
        return oc

    cg, co = _codes(mk_g, mk_o, lambda g: g.update_x(1.0, 1.0, 0.0, 1.0), lambda o: o.update_x(1.0, 1.0, 0.0, 1.0))
    assert cg == co == 1


def test_image_overflow():
    from cafproject_amd.cube import Cube
    from tests.helpers import zeldovich_state
    nc, nnt = 24, 2
    st = zeldovich_state(nc, nnt, seed=3)

    def mk_g():
        g = Cube(nc, nnt, image_buffer=0.4)
        g.set_state(st["xp"], st["vp"], st["pid"], st["rhoc"], st["vfield"])
        return g

    def mk_o():
        oc = O.OracleCube(nc, nnt, image_buffer=0.4)
        oc.set_state(st["xp"], st["vp"], st["pid"], st["rhoc"], st["vfield"])
        return oc

    cg, co = _codes(mk_g, mk_o, lambda g: g.buffer_grid(), lambda o: o.buffer_grid())
    assert cg == co == 2


def test_pp_overflow():
    from cafproject_amd.cube import Cube
    from tests.helpers import zeldovich_state
    nc, nnt = 24, 2
    st = zeldovich_state(nc, nnt, seed=4)

    def mk_g():
        g = Cube(nc, nnt, pp_on=True, tile_buffer=0.15)
        g.set_state(st["xp"], st["vp"], st["pid"], st["rhoc"], st["vfield"])
        g.buffer_grid(); g.buffer_x(); g.buffer_v()
        return g

    def mk_o():
        oc = O.OracleCube(nc, nnt, pp_on=True, tile_buffer=0.15)
        oc.set_state(st["xp"], st["vp"], st["pid"], st["rhoc"], st["vfield"])
        oc.buffer_grid(); oc.buffer_x(); oc.buffer_v()
        return oc

    args = (0.02, 3.0, st["sigma_vi"], st["mass_p"])
    cg, co = _codes(mk_g, mk_o, lambda g: g.ext_pp_force(*args), lambda o: o.ext_pp_force(*args))
    assert cg == co == 3


def test_displacement_past_buffer():
    from cafproject_amd.cube import Cube
    nc, nnt = 16, 1
    pos = np.array([[5.3, 7.7, 11.0], [9.1, 2.2, 3.3]])
    vel = np.array([[400.0, 0, 0], [0, 0, 0]])

    def mk_g():
        g = Cube(nc, nnt)
        g.set_state_xv(pos, vel.astype(np.float32), 1.0)
        return g

    def mk_o():
        oc = O.OracleCube(nc, nnt)
        oc.state_from_xv(pos, vel, None, 1.0)
        oc.buffer_grid(); oc.buffer_x(); oc.buffer_v()
        return oc

    cg, co = _codes(mk_g, mk_o, lambda g: g.update_x(1.0, 1.0, 0.0, 1.0), lambda o: o.update_x(1.0, 1.0, 0.0, 1.0))
    assert cg == co == 6


def test_cell_above_lds_sort_capacity_bit_exact():
    """A cell with more particles than the per-cell LDS sort holds (16384)
    is sorted by merging LDS-sorted runs; the drift of a state with one
    such cell (20000 particles, as in a collapsed halo core) and the IC
    count sort of it match the oracle bit for bit."""
    from cafproject_amd.cube import Cube
    nc, nnt = 16, 1
    rng = np.random.default_rng(9)
    n_dense, n_rest = 20000, 12768
    dense = np.array([7.0, 8.0, 9.0]) + rng.uniform(0.02, 0.98, (n_dense, 3))
    rest = rng.uniform(0, nc, (n_rest, 3))
    pos = np.concatenate([dense, rest])
    vel = np.concatenate([rng.normal(0, 0.05, (n_dense, 3)), rng.normal(0, 0.5, (n_rest, 3))])
    sig = 0.3
    g = Cube(nc, nnt)
    g.set_state_xv(pos, vel.astype(np.float32), sig)
    oc = O.OracleCube(nc, nnt)
    oc.state_from_xv(pos, vel.astype(np.float32).astype(np.float64), None, sig)
    so = oc.get_state()
    assert so["rhoc"].max() >= n_dense  # the dense cell (and the few uniform particles that fall in it)
    sg = g.get_state()
    for k in ("rhoc", "xp", "vp"):
        assert np.array_equal(sg[k], so[k]), k
    oc.buffer_grid(); oc.buffer_x(); oc.buffer_v()
    for dt in (2.0, 3.0):
        vz = 3.0
        oc.update_x(dt, dt, vz, sig)
        g.update_x(dt, dt, vz, sig)
        oc.buffer_grid(); oc.buffer_x(); oc.buffer_v()
        g.buffer_grid(); g.buffer_x(); g.buffer_v()
        so, sg = oc.get_state(), g.get_state()
        assert so["rhoc"].max() > 16384
        for k in ("rhoc", "xp", "vp", "pid", "vfield"):
            assert np.array_equal(sg[k], so[k]), (dt, k)
