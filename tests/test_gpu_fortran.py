"""The Fortran driver (cafproject_amd/fortran/cube_main.f90) runs the step
loop of W/main.f90:102-118 through the ISO_C_BINDING interface, reading and
writing reference-format checkpoints.  Replaying the scalars it printed
through the ctypes harness on the same checkpoint must give a bit-identical
final state, and its host time integration must take the same (a, dt)
sequence as cafproject_amd.timestep."""
import os
import struct
import subprocess

import numpy as np
import pytest

from cafproject_amd import checkpoint as ck
from tests.helpers import zeldovich_state

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "cafproject_amd", "fortran", "cube_main")
DATA = os.path.join(ROOT, "cafproject_amd", "data")


def _f(h):
    return struct.unpack("<f", struct.pack("<I", int(h, 16)))[0]


@pytest.mark.parametrize("pp", [0, 1])
def test_fortran_driver_matches_ctypes_replay(tmp_path, pp):
    from cafproject_amd.cube import Cube
    from cafproject_amd.timestep import Clock
    if not os.path.exists(DRIVER):
        pytest.fail("Fortran driver not built (cafproject_amd/build.py builds it)")
    nc, nnt, nsteps = 24, 2, 3
    st = zeldovich_state(nc, nnt, seed=5, izipx=1, izipv=2)
    om = 0.31
    vz0 = float(np.abs(st["vfield"][:, 2]).max() * 1.5 + 0.02)
    h = ck.new_header(npglobal=st["xp"].shape[0], izipx=1, izipv=2, image=1, nn=1, nnt=nnt, nt=nc // nnt,
                      ncell=4, ncb=6, a=st["a"], dt_pp=1000.0, dt_fine=1000.0, dt_coarse=1000.0, dt_vmax=1000.0,
                      mass_p_cdm=st["mass_p"], box=3.125 * nc, h0=0.68, omega_m=om, omega_l=1 - om,
                      sigma_vi=st["sigma_vi"], z_i=49.0, vz_max=vz0)
    z_in = 1 / st["a"] - 1
    ck.write(str(tmp_path / "in"), z_in, h, st["xp"], st["vp"], st["rhoc"], st["vfield"], st["pid"])

    # zero steps: checkpoint in -> GPU state -> checkpoint out is the identity
    r = subprocess.run([DRIVER, str(tmp_path / "in"), "%.6f" % z_in, str(tmp_path / "out0"), "0.0", "0",
                        "2", str(pp), DATA], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    _, out0 = ck.read(str(tmp_path / "out0"), 0.0, nc)
    for k in ("xp", "vp", "rhoc", "vfield", "pid"):
        assert np.array_equal(np.asarray(out0[k]).reshape(-1), np.asarray(st[k]).reshape(-1)), k

    r = subprocess.run([DRIVER, str(tmp_path / "in"), "%.6f" % z_in, str(tmp_path / "out"), "0.0", str(nsteps),
                        "2", str(pp), DATA], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    steps = [ln.split() for ln in r.stdout.splitlines() if ln.startswith("step")]
    assert len(steps) == nsteps, r.stdout

    hdr, out = ck.read(str(tmp_path / "out"), 0.0, nc)
    assert int(hdr["timestep"]) == nsteps

    # replay through ctypes with the driver's scalars
    g = Cube(nc, nnt, izipx=1, izipv=2, pp_on=bool(pp))
    g.init_kernels()
    g.set_state(st["xp"], st["vp"], st["pid"], st["rhoc"], st["vfield"])
    clock = Clock(st["a"], [0.0], om, 1 - om)
    sigma, vz = np.float32(st["sigma_vi"]), np.float32(vz0)
    dtc = dict(dt_fine=1000.0, dt_coarse=1000.0, dt_pp=1000.0, dt_vmax=1000.0)
    GG = np.float32(1.0) / np.float32(6.0) / np.float32(np.pi)
    for s in steps:
        a, dt, dt_old, a_mid, sig_new = (_f(x) for x in s[2:7])
        clock.timestep(**dtc)
        assert (np.float32(clock.a), np.float32(clock.dt), np.float32(clock.a_mid)) == \
            (np.float32(a), np.float32(dt), np.float32(a_mid)), "host time integration differs"
        vs = g.update_x(dt_old, dt, float(vz), float(sigma))
        g.buffer_grid()
        g.buffer_x()
        if pp:
            f2 = g.ext_pp_force(a_mid, dt, float(sigma), st["mass_p"])
            dtc["dt_pp"] = float(np.float32(1.0) / (np.sqrt(np.float32(f2)) * np.float32(a_mid) * GG))
        pm = g.particle_mesh(a_mid, dt, float(sigma), vs["sigma_vi_new"], st["mass_p"])
        g.buffer_v()
        sigma = np.float32(vs["sigma_vi_new"])
        assert sigma == np.float32(sig_new), (r.stdout, vs, float(vz), dt_old, dt)
        vz = np.float32(pm["vz_max"])
        dtc.update(dt_fine=pm["dt_fine"], dt_coarse=pm["dt_coarse"], dt_vmax=pm["dt_vmax"])
    mine = g.get_state()
    assert out["nplocal"] == mine["nplocal"]
    for k in ("xp", "vp", "rhoc", "pid"):
        assert np.array_equal(np.asarray(out[k]).reshape(-1), np.asarray(mine[k]).reshape(-1)), k
    assert np.array_equal(out["vfield"].reshape(-1), mine["vfield"].reshape(-1))
