"""CPU checks of the P(k) analysis (cafproject_amd/power.py, the restatement
of WU/cicpower.f90:96-187 and WU/powerspectrum.f90:28-135) against what the
mathematics fixes: CIC mass conservation and weights, the count of
independent Fourier modes of a real field, and the exact power of a single
plane wave (amplitude, sinc^4 window, 4 pi k^3 factor, bin index)."""
import itertools

import numpy as np

from cafproject_amd import power


def test_cic_weights_and_mass():
    nc, ng = 8, 16
    # mesh coordinate pos*ng/nc - 0.5 = 3 exactly: all mass in cell 3
    pos = np.array([[3.5, 3.5, 3.5]]) * nc / ng
    rho = power.cic_density(pos, nc, ng)
    assert rho[3, 3, 3] == 1.0 and rho.sum() == 1.0
    # halfway between cells 3 and 4 on every axis: 1/8 in each of 8 cells
    pos = np.array([[4.0, 4.0, 4.0]]) * nc / ng
    rho = power.cic_density(pos, nc, ng)
    assert np.allclose(rho[3:5, 3:5, 3:5], 0.125) and abs(rho.sum() - 1) < 1e-15
    # periodic wrap at the low edge: mesh coordinate -0.25
    pos = np.array([[0.25, 0.5, 0.5]]) * nc / ng
    rho = power.cic_density(pos, nc, ng)
    assert abs(rho[0, 0, ng - 1] - 0.25) < 1e-15 and abs(rho[0, 0, 0] - 0.75) < 1e-15
    rng = np.random.default_rng(3)
    pos = rng.uniform(0, nc, size=(1000, 3))
    assert abs(power.cic_density(pos, nc, ng).sum() - 1000) < 1e-9


def test_independent_mode_count():
    # a real field on an even ng^3 lattice has (ng^3 - 8)/2 + 8 independent
    # modes (8 self-conjugate ones); the skip rules keep each once, k=0 aside
    for ng in (8, 12, 16):
        _, _, cnt = power.power(np.zeros((ng, ng, ng)), 1.0)
        assert cnt.sum() == (ng ** 3 - 8) // 2 + 7


def test_mode_count_by_brute_force():
    ng = 8
    ny = ng // 2
    seen, per_bin = set(), {}
    for kx, ky, kz in itertools.product(range(-ny + 1, ny + 1), repeat=3):
        if (kx, ky, kz) == (0, 0, 0):
            continue
        # k and -k, both written with components in -ny+1..ny
        key = min((kx, ky, kz), tuple((-c + ny - 1) % ng - ny + 1 for c in (kx, ky, kz)))
        if key in seen:
            continue
        seen.add(key)
        b = int(np.rint(np.sqrt(kx * kx + ky * ky + kz * kz)))
        per_bin[b] = per_bin.get(b, 0) + 1
    _, _, cnt = power.power(np.zeros((ng, ng, ng)), 1.0)
    for b in range(1, len(cnt) + 1):
        assert cnt[b - 1] == per_bin.get(b, 0), b


def test_plane_wave_power():
    ng, box, A = 16, 100.0, 0.3
    x = np.arange(ng)
    for kv in ((3, 0, 0), (0, 2, 1), (1, 1, 1)):
        ph = 2 * np.pi * (kv[0] * x[None, None, :] + kv[1] * x[None, :, None] + kv[2] * x[:, None, None]) / ng
        delta = A * np.cos(ph)
        k, d2, cnt = power.power(delta, box)
        kr = np.sqrt(sum(c * c for c in kv))
        b = int(np.rint(kr))
        s = np.prod([1.0 if c == 0 else np.sin(np.pi * c / ng) / (np.pi * c / ng) for c in kv])
        want = (A * A / 4) / s ** 4 * 4 * np.pi * kr ** 3 / cnt[b - 1]
        assert abs(d2[b - 1] / want - 1) < 1e-12
        others = np.delete(d2, b - 1)
        assert np.all(others < 1e-20 * want)
        # k is the mean |k| of the bin's modes, in 2 pi / box units
        assert 2 * np.pi / box * (b - 0.5) <= k[b - 1] <= 2 * np.pi / box * (b + 0.5)
