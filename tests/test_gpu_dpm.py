"""Differentiable PM on the GPU (include/cube_dpm.h) against its FP64 CPU
twin (tests/dpm_ref.py): forward density, gradient of an MSE loss with
respect to the initial white noise (<= 1e-3 relative), central finite
differences of the FP32 forward along 64 random directions, checkpointed
vs recomputed adjoint, and the torch.autograd harness."""
import numpy as np
import pytest

from cafproject_amd import dpm
from tests.dpm_ref import Ref
from tests.helpers import rel_l2

pytestmark = pytest.mark.gpu

NG, NF, NSTEPS, BOX = 32, 64, 10, 64.0


def _setup(checkpoint=True, seed=3):
    import torch
    ic, steps = dpm.schedule(0.1, 1.0, NSTEPS)
    amp = dpm.amplitude(NG, BOX)
    box = dpm.DPM(NG, NF, NSTEPS, checkpoint=checkpoint)
    box.set_amplitude(amp)
    box.set_schedule(ic, steps)
    rng = np.random.default_rng(seed)
    w = rng.standard_normal(NG ** 3).astype(np.float32)
    target = (np.random.default_rng(seed + 100).standard_normal(NF ** 3) * 0.5).astype(np.float32)
    ref = Ref(NG, NF, amp.astype(np.float64), ic, steps)
    return box, w, target, ref, torch


def _fwd(box, w, torch):
    wt = torch.from_numpy(w).cuda()
    d = torch.empty(NF ** 3, device="cuda")
    torch.cuda.synchronize()
    box.forward(wt.data_ptr(), d.data_ptr())
    box.sync()
    return d.cpu().numpy().astype(np.float64)


def _adj(box, dbar, torch):
    db = torch.from_numpy(np.ascontiguousarray(dbar, np.float32)).cuda()
    nb = torch.empty(NG ** 3, device="cuda")
    torch.cuda.synchronize()
    box.adjoint(db.data_ptr(), nb.data_ptr())
    box.sync()
    return nb.cpu().numpy().astype(np.float64)


def test_forward_matches_fp64_twin():
    box, w, target, ref, torch = _setup()
    dg = _fwd(box, w, torch)
    dr = ref.forward(w.astype(np.float64)).ravel()
    err = rel_l2(dg, dr)
    print(f"forward delta rel L2 vs FP64: {err:.2e}")
    assert err < 1e-3, err


def test_gradient_matches_fp64_adjoint():
    box, w, target, ref, torch = _setup()
    dg = _fwd(box, w, torch)
    gg = _adj(box, 2 * (dg - target) / dg.size, torch)
    dr = ref.forward(w.astype(np.float64)).ravel()
    gr = ref.adjoint(2 * (dr - target) / dr.size)
    err = rel_l2(gg, gr)
    print(f"gradient rel L2 vs FP64 adjoint: {err:.2e}")
    assert err < 1e-3, err


def test_gradient_matches_finite_differences():
    # CIC is piecewise linear in the positions, so a step that moves a
    # particle across a cell edge sees a kink, and FP32 positions need a
    # step well above their resolution: eps = 1e-3 of the unit noise.  Over
    # 64 random directions the FD and adjoint directional derivatives then
    # agree as a vector (the FP64 twin gives 3% rel L2, correlation 0.9995,
    # slope 0.998 at this step); a wrong sign or a dropped term does not.
    box, w, target, ref, torch = _setup()

    def loss(wv):
        d = _fwd(box, wv.astype(np.float32), torch)
        return np.mean((d - target) ** 2), d

    L0, d0 = loss(w.astype(np.float64))
    g = _adj(box, 2 * (d0 - target) / d0.size, torch)
    rng = np.random.default_rng(9)
    eps = 1e-3
    fd, an = [], []
    for t in range(64):
        v = rng.standard_normal(w.size)
        fd.append((loss(w + eps * v)[0] - loss(w - eps * v)[0]) / (2 * eps))
        an.append(float(g @ v))
    fd, an = np.array(fd), np.array(an)
    err = np.linalg.norm(fd - an) / np.linalg.norm(an)
    corr = np.corrcoef(fd, an)[0, 1]
    slope = float(fd @ an / (an @ an))
    print(f"FD vs adjoint over 64 directions: rel L2 {err:.2e}, corr {corr:.5f}, slope {slope:.4f}")
    assert err < 0.1 and corr > 0.995 and abs(slope - 1) < 0.03


def test_recompute_matches_checkpoints():
    box, w, target, ref, torch = _setup(checkpoint=True)
    d1 = _fwd(box, w, torch)
    g1 = _adj(box, 2 * (d1 - target) / d1.size, torch)
    box2, *_ = _setup(checkpoint=False)
    d2 = _fwd(box2, w, torch)
    g2 = _adj(box2, 2 * (d2 - target) / d2.size, torch)
    assert np.array_equal(d2, d1)  # the same steps, order-free sums: the same bits
    assert np.array_equal(g2, g1)


def test_autograd_function_gives_the_adjoint():
    from tests.dpm_autograd import simulate
    box, w, target, ref, torch = _setup()
    wt = torch.from_numpy(w).cuda().requires_grad_(True)
    tt = torch.from_numpy(target).cuda()
    loss = torch.mean((simulate(wt, box) - tt) ** 2)
    loss.backward()
    d = _fwd(box, w, torch)
    g = _adj(box, 2 * (d - target) / d.size, torch)
    assert rel_l2(wt.grad.cpu().numpy(), g) < 1e-5
