"""The FP64 twin of the differentiable PM is a correct adjoint: the
operator transposes satisfy <A u, v> = <u, A^T v>, and the gradient of
L = mean((delta - target)^2) matches central finite differences along
random directions (where no particle crosses a CIC cell edge)."""
import numpy as np

from cafproject_amd import dpm
from tests.dpm_ref import Ref, kfac


def _ref(ng=8, nf=16, nsteps=3, box=64.0):
    ic, steps = dpm.schedule(0.1, 1.0, nsteps)
    amp = dpm.amplitude(ng, box).astype(np.float64)
    return Ref(ng, nf, amp, ic, steps)


def test_poisson_transpose():
    n = 12
    rng = np.random.default_rng(0)
    u = rng.standard_normal((n, n, n))
    v = rng.standard_normal((3, n, n, n))
    uk = np.fft.rfftn(u)
    Au = [np.fft.irfftn(1j * kfac(n, d) * uk, s=(n, n, n), axes=(0, 1, 2)) for d in range(3)]
    ATv = np.fft.irfftn(sum((-1j) * kfac(n, d) * np.fft.rfftn(v[d]) for d in range(3)), s=(n, n, n), axes=(0, 1, 2))
    lhs = sum((Au[d] * v[d]).sum() for d in range(3))
    assert abs(lhs - (u * ATv).sum()) < 1e-10 * abs(lhs)


def test_nyquist_planes_are_odd_consistent():
    # K_d vanishes on the Nyquist plane of axis d, so i K_d X stays Hermitian
    n = 8
    for d in range(3):
        K = kfac(n, d)
        assert np.all(K[..., n // 2] == 0) if d == 0 else True
    assert np.all(kfac(n, 1)[:, n // 2, :] == 0) and np.all(kfac(n, 2)[n // 2] == 0)


def test_gradient_matches_finite_differences():
    R = _ref()
    rng = np.random.default_rng(1)
    w = rng.standard_normal(R.ng ** 3)
    target = rng.standard_normal((R.nf,) * 3) * 0.1

    def loss(w):
        d = R.forward(w)
        return np.mean((d - target) ** 2), d

    L0, d0 = loss(w)
    g = R.adjoint(2 * (d0 - target) / d0.size)
    eps = 1e-6
    good = 0
    for t in range(64):
        v = rng.standard_normal(w.size)
        fd = (loss(w + eps * v)[0] - loss(w - eps * v)[0]) / (2 * eps)
        an = g @ v
        if abs(fd - an) <= 1e-5 * max(abs(an), 1e-12):
            good += 1
    # a direction whose step moves a particle across a CIC edge sees a kink
    assert good >= 60, good
