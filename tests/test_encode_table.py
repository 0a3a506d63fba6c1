"""The kernels re-encode velocities by counting code boundaries
T[k] = tan((k+0.5)*pi_f/(nvbin-1)) (cube_encode_v_tab in
cafproject_amd/csrc/cube_internal.h) instead of evaluating
nint((nvbin-1)*atan(u)/pi_f) (W/update_particle.f90:117, W/pm.f90:158).
This restates that lookup with the same long-double table and checks it
against the FP64 formula (numpy arctan) on every boundary's neighbours and
on random residuals, for 1- and 2-byte codes."""
import numpy as np
import pytest

PI_F = np.float32(np.pi)


def table(nvbin):
    kmax = nvbin // 2 - 1
    k = np.arange(kmax, dtype=np.longdouble)
    t = np.tan((k + np.longdouble(0.5)) * np.longdouble(PI_F) / np.longdouble(nvbin - 1)).astype(np.float64)
    return np.append(t, np.inf), kmax


def encode_tab(u, T, kmax, nvbin):
    scale = np.float32((nvbin - 1) / np.float64(PI_F))
    a = np.abs(u)
    with np.errstate(over="ignore"):
        af = a.astype(np.float32)
    k0 = np.minimum((scale * np.arctan(af) + np.float32(0.5)).astype(np.int64), kmax)
    lo = np.where(k0 > 0, T[np.maximum(k0 - 1, 0)], -1.0)
    k = k0 - (a < lo) + (a >= T[k0])
    k = np.clip(k, 0, kmax)
    return np.where(u < 0, -k, k)


def encode_formula(u, nvbin):
    y = (np.float64(nvbin - 1) * np.arctan(u)) / np.float64(PI_F)
    return (np.sign(y) * np.floor(np.abs(y) + 0.5)).astype(np.int64)  # nint: half away from zero


@pytest.mark.parametrize("nvbin", [256, 65536])
def test_table_lookup_matches_formula(nvbin):
    T, kmax = table(nvbin)
    rng = np.random.default_rng(7)
    # residuals c*(v - v_cell) are O(1): normal, wide tails, and exact zeros
    u = np.concatenate([rng.standard_normal(400000), rng.standard_normal(100000) * 50,
                        rng.standard_cauchy(100000), [0.0, -0.0, 1e300, -1e300]])
    assert np.array_equal(encode_tab(u, T, kmax, nvbin), encode_formula(u, nvbin))


@pytest.mark.parametrize("nvbin", [256, 65536])
def test_boundaries_are_where_the_code_steps(nvbin):
    T, kmax = table(nvbin)
    b = T[:-1]
    # just below / at-or-above each boundary the code is k / k+1
    below = np.nextafter(b, 0) * (1 - 1e-12)
    above = b * (1 + 1e-12)
    k = np.arange(kmax)
    assert np.array_equal(encode_tab(below, T, kmax, nvbin), k)
    assert np.array_equal(encode_tab(above, T, kmax, nvbin), k + 1)
    # the FP64 formula rounds y = (nvbin-1)*atan(u)/pi_f with an error that,
    # near pi/2, spans ~1e-11 of u: it agrees once a step clears that
    assert np.array_equal(encode_formula(b * (1 - 1e-9), nvbin), k)
    assert np.array_equal(encode_formula(b * (1 + 1e-9), nvbin), k + 1)
    assert np.array_equal(encode_tab(-above, T, kmax, nvbin), -(k + 1))
    # the largest code: u -> inf gives (nvbin-1)/2 * pi/pi_f < kmax + 1/2
    assert encode_tab(np.array([np.inf]), T, kmax, nvbin)[0] == kmax
