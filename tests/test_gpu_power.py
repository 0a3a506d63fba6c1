"""cube_power (GPU P(k): FP64 CIC, FP32 FFT, FP64 binning) against the host
analysis cafproject_amd.power.state_power (FP64 throughout) on the same
state: identical mode counts and mean |k| per bin, Delta^2 within 1e-4."""
import numpy as np
import pytest

from cafproject_amd import power
from tests.helpers import zeldovich_state

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("nc,nnt,izipx,ngs", [(32, 2, 2, (32, 64, 128)), (24, 2, 1, (24, 48)), (32, 1, 2, (16,))])
def test_power_matches_host(nc, nnt, izipx, ngs):
    from cafproject_amd.cube import Cube
    box = 150.0
    st = zeldovich_state(nc, nnt, seed=5, izipx=izipx, izipv=2, box=box, z_i=3.0)
    g = Cube(nc, nnt, izipx=izipx, izipv=2)
    g.set_state(st["xp"], st["vp"], st["pid"], st["rhoc"], st["vfield"])
    for ng in ngs:
        print(f"nc={nc} nnt={nnt} izipx={izipx} ng={ng}", flush=True)
        k_h, p_h, n_h = power.state_power(st, nc, nnt, izipx, box, ng=ng)
        k_g, p_g, n_g = g.power(ng, box)
        assert np.array_equal(n_g, n_h)
        m = n_h > 0
        assert np.allclose(k_g[m], k_h[m], rtol=1e-12, atol=0)
        rel = np.abs(p_g[m] / p_h[m] - 1)
        assert rel.max() < 1e-4, (ng, rel.max())


def test_power_rejects_bad_mesh():
    from cafproject_amd.cube import Cube, CubeError
    nc, nnt = 16, 1
    st = zeldovich_state(nc, nnt, seed=2)
    g = Cube(nc, nnt)
    g.set_state(st["xp"], st["vp"], st["pid"], st["rhoc"], st["vfield"])
    for ng in (40, 128, 2):
        with pytest.raises(CubeError):
            g.power(ng, 50.0)
