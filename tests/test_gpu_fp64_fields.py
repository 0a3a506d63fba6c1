"""Error of the GPU's FP32 mesh fields against the FP64 oracle, stated per
field: fine density, fine potential, the three fine force components,
coarse density and the three coarse force components, each as a relative
L2 error after one deposit and solve of the same state.  The FP64 twin
(oracle.cic_fine/cic_coarse, fine_force_box, fine_potential_box,
coarse_force with an FP64 kern_c) follows W/pm.f90:50-126 and :205-315 on
the whole periodic box.  Gate: every field within 1e-5 relative (the FP32
bar); the measured errors are printed, one line per field."""
import numpy as np
import pytest

from cafproject_amd import power
from cafproject_amd.cube import load_potential_table
from oracle import oracle as O
from tests.helpers import rel_l2, zeldovich_state

pytestmark = pytest.mark.gpu

CASES = [dict(nc=32, nnt=2, izipx=2, izipv=2, pp=False, z_i=49.0),
         dict(nc=32, nnt=2, izipx=1, izipv=2, pp=True, z_i=5.0),
         dict(nc=24, nnt=1, izipx=2, izipv=2, pp=False, z_i=5.0)]


@pytest.mark.parametrize("case", CASES, ids=lambda c: "nc{nc}_nnt{nnt}_x{izipx}_pp{pp:d}_z{z_i:g}".format(**c))
def test_fields_against_fp64_oracle(case):
    from cafproject_amd.cube import Cube
    nc, nnt, ix, iv, pp = case["nc"], case["nnt"], case["izipx"], case["izipv"], case["pp"]
    st = zeldovich_state(nc, nnt, seed=21, izipx=ix, izipv=iv, z_i=case["z_i"])
    m = st["mass_p"]
    g = Cube(nc, nnt, izipx=ix, izipv=iv, pp_on=pp)
    g.init_kernels()
    g.set_state(st["xp"], st["vp"], st["pid"], st["rhoc"], st["vfield"])
    wg1 = load_potential_table()
    g.fine_deposit(m)
    rho_g = g.get_field("rho_f")
    phi_g = g.fine_potential(wg1)
    g.fine_force()
    F_g = g.get_field("force_f")
    g.coarse_deposit(m)
    rc_g = g.get_field("rho_c")
    g.coarse_force()
    Fc_g = g.get_field("force_c")

    pos = power.decode_positions(st, nc, nnt, ix)
    fk, ck = O.load_tables()
    rho64 = O.cic_fine(pos, nc, m)
    phi64 = O.fine_potential_box(rho64, wg1, pp, 1)
    F64 = O.fine_force_box(rho64, fk, pp, 1)
    rc64 = O.cic_coarse(pos, nc, m)
    Fc64, _ = O.coarse_force(rc64, O.kernel_c(nc, ck, dtype=np.float64), nc)
    Fc64 = Fc64[1:-1, 1:-1, 1:-1]

    err = {"fine rho": rel_l2(rho_g, rho64), "fine phi": rel_l2(phi_g, phi64)}
    for d, ax in enumerate("xyz"):
        err[f"fine F{ax}"] = rel_l2(F_g[..., d], F64[..., d])
    err["coarse rho"] = rel_l2(rc_g, rc64)
    for d, ax in enumerate("xyz"):
        err[f"coarse F{ax}"] = rel_l2(Fc_g[..., d], Fc64[..., d])
    for k, v in err.items():
        print(f"FP32 vs FP64 {case}: {k:10s} rel L2 {v:.3e}", flush=True)
    bad = {k: v for k, v in err.items() if not v < 1e-5}
    assert not bad, bad
