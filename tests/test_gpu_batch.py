"""Batched ensemble runner (include/cube_batch.h): the graph replay of the
whole ensemble gives the per-box loop's losses and gradients bit for bit
(the model's sums are order-free fixed point), replay after
replay with eager work in between, and each box equals a standalone box of
the differentiable PM; the full 48-box ensemble of uneven sizes (16x32^3,
16x64^3, 8x96^3, 8x128^3 particles, 10 steps) gives finite losses, the
loop's results bit for bit from its graph, and a 32^3 box's gradient
within 1e-3 of the FP64 twin's."""
import numpy as np
import pytest

from cafproject_amd import dpm
from tests.helpers import rel_l2

pytestmark = pytest.mark.gpu

CLASSES = ((16, 3), (32, 2), (24, 1))


def _grab(t):
    import torch
    torch.cuda.synchronize()
    return [(float(x["loss"].item()), x["noise_bar"].cpu().numpy().copy(), x["delta"].cpu().numpy().copy())
            for x in t]


def test_graph_replays_match_the_per_box_loop():
    import torch
    b, t = dpm.make_ensemble(CLASSES, nsteps=4)
    shapes = b.tune()
    assert shapes.shape == (6, 2) and np.all(shapes > 0)
    b.run(graph=False)
    b.sync()
    ref = _grab(t)
    for rep in range(3):
        for x in t:  # eager work on torch's stream between replays
            x["noise_bar"].zero_()
            x["delta"].fill_(7.0)
        torch.cuda.synchronize()
        b.run(graph=True)
        b.sync()
        got = _grab(t)
        for (l0, g0, d0), (l1, g1, d1) in zip(ref, got):  # deterministic sums: the same bits
            assert l1 == l0
            assert np.array_equal(d1, d0)
            assert np.array_equal(g1, g0)
        b.run(graph=False)  # eager runs of the same boxes between replays
        b.sync()


def test_ensemble_box_equals_standalone_box():
    import torch
    b, t = dpm.make_ensemble(((16, 2),), nsteps=3)
    b.run(graph=True)
    b.sync()
    x = t[1]
    ic, steps = dpm.schedule(0.1, 1.0, 3)
    box = dpm.DPM(16, 32, 3)
    from cafproject_amd import ics
    box.set_amplitude(dpm.amplitude(16, x["box"], ics.Cosmology(sigma8=x["sigma8"])))
    box.set_schedule(ic, steps)
    d = torch.empty(32 ** 3, device="cuda")
    torch.cuda.synchronize()
    box.forward(x["noise"].data_ptr(), d.data_ptr())
    box.sync()
    db = (2 * (d - x["target"]) / d.numel()).contiguous()
    nb = torch.empty(16 ** 3, device="cuda")
    box.adjoint(db.data_ptr(), nb.data_ptr())
    box.sync()
    assert rel_l2(d.cpu().numpy(), x["delta"].cpu().numpy()) < 1e-5
    assert rel_l2(nb.cpu().numpy(), x["noise_bar"].cpu().numpy()) < 1e-4
    loss = float(torch.mean((d - x["target"]) ** 2).item())
    assert abs(loss / float(x["loss"].item()) - 1) < 1e-5


def test_config5_full_ensemble():
    import torch
    from cafproject_amd import ics
    from tests.dpm_ref import Ref
    b, t = dpm.make_ensemble()  # 48 boxes, 10 steps
    assert len(t) == 48
    b.tune()
    b.run(graph=False)
    b.sync()
    loop = _grab(t)
    assert all(np.isfinite(l) for l, _, _ in loop)
    b.run(graph=True)
    b.sync()
    graph = _grab(t)
    for (l0, g0, d0), (l1, g1, d1) in zip(loop, graph):  # deterministic sums: the same bits
        assert np.isfinite(l1) and l1 == l0
        assert np.array_equal(d1, d0)
        assert np.array_equal(g1, g0)
    # box 0 (32^3 particles, 64^3 mesh) against the FP64 twin
    x = t[0]
    ic, steps = dpm.schedule(0.1, 1.0, 10)
    amp = dpm.amplitude(32, x["box"], ics.Cosmology(sigma8=x["sigma8"]))
    ref = Ref(32, 64, amp.astype(np.float64), ic, steps)
    w = x["noise"].cpu().numpy().astype(np.float64)
    tgt = x["target"].cpu().numpy().astype(np.float64)
    dr = ref.forward(w).ravel()
    gr = ref.adjoint(2 * (dr - tgt) / dr.size)
    _, g1, d1 = graph[0]
    print(f"box 0 vs FP64 twin: delta {rel_l2(d1, dr):.2e}, gradient {rel_l2(g1, gr):.2e}", flush=True)
    assert rel_l2(d1, dr) < 1e-3
    assert rel_l2(g1, gr) < 1e-3
    del b, t
    torch.cuda.empty_cache()
