"""torch.autograd wrapper of the adjoint pair, for gradient checks only.

RenderFunction.apply(theta, inverse_render, salt) returns the primal image [H][W][3] (float32,
CPU tensor) rendered by pt_forward at theta; its backward calls pt_backward with the incoming
gradient as the adjoint image, replaying the same resetSalt stream. The renders themselves
stay on the HIP path; torch only carries the host-side tensors.
"""
import numpy as np
import torch


class RenderFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, theta, inv, salt):
        img = inv.render(theta.detach().cpu().double().numpy(), salt)
        ctx.inv, ctx.salt = inv, salt
        ctx.save_for_backward(theta)
        return torch.from_numpy(np.ascontiguousarray(img[..., :3]))

    @staticmethod
    def backward(ctx, grad_img):
        (theta,) = ctx.saved_tensors
        g = grad_img.detach().cpu().float().numpy()
        dL = np.zeros(g.shape[:2] + (4,), np.float32)
        dL[..., :3] = g
        grad = ctx.inv.gradient(theta.detach().cpu().double().numpy(), dL, ctx.salt)
        return torch.from_numpy(grad).to(theta.dtype), None, None


def render(theta, inv, salt=1):
    return RenderFunction.apply(theta, inv, salt)
