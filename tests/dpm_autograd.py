"""torch.autograd.Function over the differentiable PM's C ABI (test
harness only): forward = cube_dpm_forward, backward = cube_dpm_adjoint.
The library runs on its own stream; the harness synchronises around each
call so torch's stream sees the results."""
import torch


class DPMFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, noise, box):
        noise = noise.contiguous()
        delta = torch.empty(box.nf ** 3, device=noise.device, dtype=torch.float32)
        torch.cuda.synchronize()
        box.forward(noise.data_ptr(), delta.data_ptr())
        box.sync()
        ctx.box = box
        return delta

    @staticmethod
    def backward(ctx, gdelta):
        g = gdelta.contiguous()
        nb = torch.empty(ctx.box.ng ** 3, device=g.device, dtype=torch.float32)
        torch.cuda.synchronize()
        ctx.box.adjoint(g.data_ptr(), nb.data_ptr())
        ctx.box.sync()
        return nb, None


def simulate(noise, box):
    return DPMFunction.apply(noise, box)
