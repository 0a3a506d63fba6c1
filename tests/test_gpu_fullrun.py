"""Full run z=49 -> 0, GPU against the oracle from identical initial state
(48^3 particles, nc=24, 2^3 tiles, 200 Mpc/h, int16 x and v, PP on, pids).
Both runs take the oracle's (a, dt) sequence (the host time integration of
W/timestep.f90 fed with the oracle's dt constraints), so they differ only by
the GPU's own arithmetic (FP32 FFTs).  Gates: matter power spectrum within
0.1% up to half the particle Nyquist wavenumber, and the final positions of
the same particles (matched by pid) within 1% of a fine cell at the 99th
percentile."""
import numpy as np
import pytest

from cafproject_amd import power
from oracle import oracle as O
from tests.helpers import zeldovich_state

pytestmark = pytest.mark.gpu


def test_full_run_power_spectrum_and_positions():
    from cafproject_amd.cube import Cube
    from cafproject_amd.timestep import Clock
    nc, nnt, box = 24, 2, 200.0
    st = zeldovich_state(nc, nnt, seed=11, izipx=2, izipv=2, box=box)
    oc = O.OracleCube(nc, nnt, izipx=2, izipv=2, pp_on=True)
    oc.set_state(st["xp"], st["vp"], st["pid"], st["rhoc"], st["vfield"])
    oc.init_kernels()
    oc.buffer_grid(); oc.buffer_x(); oc.buffer_v()
    g = Cube(nc, nnt, izipx=2, izipv=2, pp_on=True)
    g.init_kernels()
    g.set_state(st["xp"], st["vp"], st["pid"], st["rhoc"], st["vfield"])
    om = 0.31
    clock = Clock(st["a"], [0.0], om, 1 - om)
    m = st["mass_p"]
    sig_o = sig_g = st["sigma_vi"]
    vz_o = vz_g = float(np.abs(st["vfield"][:, 2]).max() * 1.5 + 0.02)
    dtc = dict(dt_fine=1000.0, dt_coarse=1000.0, dt_pp=1000.0, dt_vmax=1000.0)
    GG = np.float32(1.0) / np.float32(6.0) / np.float32(np.pi)
    nstep = 0
    while not clock.final_step:
        clock.timestep(**dtc)
        a_mid, dt, dt_old = float(clock.a_mid), float(clock.dt), float(clock.dt_old)
        s_o, _ = oc.update_x(dt_old, dt, vz_o, sig_o)
        vs = g.update_x(dt_old, dt, vz_g, sig_g)
        oc.buffer_grid(); oc.buffer_x()
        g.buffer_grid(); g.buffer_x()
        f2 = oc.ext_pp_force(a_mid, dt, sig_o, m)
        g.ext_pp_force(a_mid, dt, sig_g, m)
        pm_o = oc.particle_mesh(a_mid, dt, sig_o, s_o, m)
        pm_g = g.particle_mesh(a_mid, dt, sig_g, vs["sigma_vi_new"], m)
        oc.buffer_v(); g.buffer_v()
        d = oc.dt_constraints(pm_o, a_mid)
        with np.errstate(divide="ignore"):
            dt_pp = float(np.float32(1.0) / (np.sqrt(np.float32(f2)) * np.float32(a_mid) * GG))
        dtc.update(dt_fine=d["dt_fine"], dt_coarse=d["dt_coarse"], dt_vmax=d["dt_vmax"], dt_pp=dt_pp)
        vz_o, vz_g = d["vz_max"], pm_g["vz_max"]
        sig_o, sig_g = s_o, vs["sigma_vi_new"]
        nstep += 1
        assert nstep < 2000
    so, sg = oc.get_state(), g.get_state()
    assert so["nplocal"] == sg["nplocal"]

    # power spectra on the particle-grid mesh, bins up to k_Nyq/2
    ng = 2 * nc
    k_o, p_o, n_o = power.state_power(so, nc, nnt, 2, box, ng=ng)
    k_g, p_g, n_g = power.state_power(sg, nc, nnt, 2, box, ng=ng)
    sel = slice(0, ng // 4)
    dp = np.abs(p_g[sel] / p_o[sel] - 1)
    print(f"{nstep} steps to a={float(clock.a):.4f}; P(k) max rel diff to k_Nyq/2: {dp.max():.2e}")
    assert np.all(n_o[sel] == n_g[sel])
    assert dp.max() < 1e-3, dp

    # positions of the same particles, in fine cells (4 per coarse cell)
    xo = power.decode_positions(so, nc, nnt, 2)
    xg = power.decode_positions(sg, nc, nnt, 2)
    io = np.argsort(so["pid"])
    ig = np.argsort(sg["pid"])
    assert np.array_equal(so["pid"][io], sg["pid"][ig])
    dx = xg[ig] - xo[io]
    dx -= nc * np.round(dx / nc)
    err = 4 * np.sqrt((dx ** 2).sum(1))
    q99 = np.quantile(err, 0.99)
    print(f"position difference in fine cells: median {np.median(err):.2e}, 99% {q99:.2e}, max {err.max():.2e}")
    assert q99 < 0.01, q99
