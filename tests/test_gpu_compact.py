"""The compact fine-mesh layout (one force spectrum that the three force
components pass through one after the other; the drift's sort scratch
inside the force mesh), which lets a 2048^3 fine mesh fit one GPU, gives the
same bits as the full layout: the same state after several full steps (with
PP, so the forces are formed beside PP on the side streams), the same
forces and max |F|^2 from the stage entry point, the same IC count sort,
the same potential."""
import numpy as np
import pytest

from tests.helpers import zeldovich_state

pytestmark = pytest.mark.gpu

CASES = [dict(nc=24, nnt=2, izipx=2, izipv=2, pp=True), dict(nc=16, nnt=1, izipx=1, izipv=2, pp=False)]


def _run(case, layout, st, nsteps=3):
    from cafproject_amd.cube import Cube
    from cafproject_amd.timestep import Clock
    g = Cube(case["nc"], case["nnt"], izipx=case["izipx"], izipv=case["izipv"], pp_on=case["pp"],
             mem_layout=layout)
    g.init_kernels()
    g.set_state(st["xp"], st["vp"], st["pid"], st["rhoc"], st["vfield"])
    clock = Clock(st["a"], [0.0], 0.31, 0.69)
    m, sig = st["mass_p"], st["sigma_vi"]
    vz = float(np.abs(st["vfield"][:, 2]).max() * 1.5 + 0.02)
    dtc = dict(dt_fine=1000.0, dt_coarse=1000.0, dt_pp=1000.0, dt_vmax=1000.0)
    outs = []
    for _ in range(nsteps):
        clock.timestep(**dtc)
        a_mid, dt = float(clock.a_mid), float(clock.dt)
        vs = g.update_x(float(clock.dt_old), dt, vz, sig)
        g.buffer_grid(); g.buffer_x()
        if case["pp"]:
            g.ext_pp_force(a_mid, dt, sig, m)
        pm = g.particle_mesh(a_mid, dt, sig, vs["sigma_vi_new"], m)
        g.buffer_v()
        dtc.update(dt_fine=pm["dt_fine"], dt_coarse=pm["dt_coarse"], dt_vmax=pm["dt_vmax"])
        vz, sig = pm["vz_max"], vs["sigma_vi_new"]
        outs.append((vs, pm))
    g.fine_deposit(m)
    f2 = g.fine_force()
    out = dict(state=g.get_state(), outs=outs, force=g.get_field("force_f"), f2=f2, phi=g.fine_potential())
    del g
    return out


@pytest.mark.parametrize("case", CASES, ids=lambda c: "nc{nc}_nnt{nnt}_x{izipx}_pp{pp:d}".format(**c))
def test_compact_layout_bit_identical(case):
    print(f"compact layout: {case}", flush=True)
    st = zeldovich_state(case["nc"], case["nnt"], seed=31, izipx=case["izipx"], izipv=case["izipv"])
    full = _run(case, 1, st)
    comp = _run(case, 2, st)
    for k in ("xp", "vp", "pid", "rhoc", "vfield"):
        assert np.array_equal(full["state"][k], comp["state"][k]), k
    assert full["outs"] == comp["outs"]
    assert np.array_equal(full["force"].view(np.uint32), comp["force"].view(np.uint32))
    assert full["f2"] == comp["f2"]
    assert np.array_equal(full["phi"].view(np.uint32), comp["phi"].view(np.uint32))


def test_compact_layout_count_sort():
    """The IC count sort (cube_set_state_xv) with its scratch carved out of
    the force mesh equals the full layout's."""
    from cafproject_amd.cube import Cube
    rng = np.random.default_rng(4)
    nc = 16
    x = rng.uniform(0, nc, (30000, 3))  # under the 16^3 x 2^3 = 32768 particles the context holds
    v = rng.normal(0, 1, (30000, 3)).astype(np.float32)
    s = []
    for layout in (1, 2):
        g = Cube(nc, 2, mem_layout=layout)
        g.set_state_xv(x, v, 0.8)
        s.append(g.get_state())
        del g
    for k in ("xp", "vp", "pid", "rhoc", "vfield"):
        assert np.array_equal(s[0][k], s[1][k]), k
