"""The differentiable PM at its workload size: 128^3 particles in a 128
Mpc/h box, 256^3 fine mesh, 10 kick-drift-kick steps from z = 9 to 0, no PP,
MSE loss against the density of the same model from another seed.  The
forward density and the gradient with respect to the initial white noise
are checked against the FP64 CPU twin (tests/dpm_ref.py) through its
fixture tests/golden/dpm_config4.npz (made by
tests/golden/make_dpm_config4.py): the loss, the norms of delta and of the
gradient, 4096 sampled values of each and 16 projections of the gradient,
all within 1e-3 relative."""
import os

import numpy as np
import pytest

from cafproject_amd import dpm
from tests.helpers import rel_l2

pytestmark = pytest.mark.gpu

FIX = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "dpm_config4.npz")


def test_config4_forward_and_gradient_against_fp64_twin():
    import torch
    f = np.load(FIX)
    ng, nf, nsteps, box_mpc, a0, a1, seed, proj_seed, nproj = f["params"]
    ng, nf, nsteps, seed, proj_seed, nproj = int(ng), int(nf), int(nsteps), int(seed), int(proj_seed), int(nproj)
    ic, steps = dpm.schedule(float(a0), float(a1), nsteps)
    box = dpm.DPM(ng, nf, nsteps, checkpoint=True)
    box.set_amplitude(dpm.amplitude(ng, float(box_mpc)))
    box.set_schedule(ic, steps)

    def forward(s):
        w = torch.from_numpy(np.random.default_rng(s).standard_normal(ng ** 3).astype(np.float32)).cuda()
        d = torch.empty(nf ** 3, device="cuda")
        torch.cuda.synchronize()
        box.forward(w.data_ptr(), d.data_ptr())
        box.sync()
        return d

    target = forward(seed + 1).clone()
    d = forward(seed)  # the adjoint below runs back through this forward's checkpoints
    dbar = (2 * (d - target) / d.numel()).contiguous()
    nb = torch.empty(ng ** 3, device="cuda")
    torch.cuda.synchronize()
    box.adjoint(dbar.data_ptr(), nb.data_ptr())
    box.sync()
    dh = d.cpu().numpy().astype(np.float64)
    th = target.cpu().numpy().astype(np.float64)
    gh = nb.cpu().numpy().astype(np.float64)
    loss = float(np.mean((dh - th) ** 2))
    pr = np.random.default_rng(proj_seed)
    proj = np.array([gh @ pr.standard_normal(ng ** 3) for _ in range(nproj)])
    errs = dict(loss=abs(loss / float(f["loss"]) - 1),
                delta_norm=abs(np.linalg.norm(dh) / float(f["delta_norm"]) - 1),
                delta_sampled=rel_l2(dh[f["delta_idx"]], f["delta_val"]),
                grad_norm=abs(np.linalg.norm(gh) / float(f["grad_norm"]) - 1),
                grad_sampled=rel_l2(gh[f["grad_idx"]], f["grad_val"]),
                grad_projections=rel_l2(proj, f["proj"]))
    print("config-4 errors against the FP64 twin: " + ", ".join(f"{k} {v:.2e}" for k, v in errs.items()), flush=True)
    for k, v in errs.items():
        assert v < 1e-3, (k, v)
