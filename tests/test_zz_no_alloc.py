"""cube_power_noalloc (P(k) with its twiddles and bin sums in the library's
fixed device array, no allocation in the call) against cube_power on the
same state, through the C ABI: identical mode counts, mean |k| to 1e-12 and
Delta^2 to 1e-6 relative (both sum in FP64 with atomics, so the order of
the additions varies).  Two calls in a row, the second on a smaller mesh,
so stale bins or twiddles of the first would show; then a size beyond the
array is refused."""
import ctypes

import numpy as np
import pytest

from tests.helpers import zeldovich_state

pytestmark = pytest.mark.gpu


def _check(g, ng, box):
    print(f"no-alloc P(k): nc={g.nc} ng={ng}", flush=True)
    k_a, p_a, n_a = g.power(ng, box)
    k_b, p_b, n_b = g.power(ng, box, noalloc=True)
    assert np.array_equal(n_a, n_b)
    m = n_a > 0
    assert m.sum() > ng // 4
    assert np.allclose(k_b[m], k_a[m], rtol=1e-12, atol=0)
    assert np.allclose(p_b[m], p_a[m], rtol=1e-6, atol=0)


def test_power_noalloc_matches_power():
    from cafproject_amd.cube import Cube, CubeError, lib
    nc, nnt, box = 32, 2, 150.0
    st = zeldovich_state(nc, nnt, seed=11, izipx=2, izipv=2, box=box, z_i=3.0)
    g = Cube(nc, nnt, izipx=2, izipv=2)
    g.set_state(st["xp"], st["vp"], st["pid"], st["rhoc"], st["vfield"])
    _check(g, 128, box)
    _check(g, 64, box)
    # bins beyond the fixed array: refused before any device work
    nbin = 1775
    ks, ps = np.zeros(nbin), np.zeros(nbin)
    cnt = np.zeros(nbin, np.int64)
    p = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    assert lib().cube_power_noalloc(g._h, 32, nbin, p(ks), p(ps), p(cnt)) == 12
    with pytest.raises(CubeError):
        g.power(40, box, noalloc=True)
