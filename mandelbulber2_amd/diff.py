"""Differentiable rendering on top of mbx_render_diff / mbx_adjoint.

theta is the vector of MBX_DIFF_NPARAMS scene parameters (power, node rotation x/y/z in
degrees, camera x/y/z, target x/y/z).  `DiffRenderer.render(theta)` returns the FP64-refined
image on the device; `RenderFunction` makes it a torch.autograd.Function whose backward runs the
adjoint kernel.  No part of this module computes pixels itself.
"""
import numpy as np

from . import abi, runtime, scene


def theta_of(p):
    """The differentiable parameters of a parameter block."""
    return np.array([p.fractal.bulb.power, *p.formulaRotation, *p.camera, *p.target], np.float64)


def with_theta(p, theta):
    """A copy of p with the differentiable parameters set from theta."""
    q = scene.copy_params(p)
    t = [float(v) for v in theta]
    q.fractal.bulb.power = t[0]
    abi.set3(q.formulaRotation, t[1:4])
    abi.set3(q.camera, t[4:7])
    abi.set3(q.target, t[7:10])
    return q


class DiffRenderer:
    """One device context rendering a base scene at varying theta."""

    def __init__(self, base_params, device=0):
        self.base = scene.copy_params(base_params)
        self.ctx = runtime.Context(device)
        self.torch = self.ctx.torch
        self._theta = None

    def _upload(self, theta):
        theta = np.asarray(theta, np.float64)
        if self._theta is None or not np.array_equal(theta, self._theta):
            self.ctx.upload(with_theta(self.base, theta))
            self._theta = theta.copy()

    def render(self, theta):
        """(outputs, checkpoints) of the differentiable render at theta."""
        self._upload(theta)
        bufs = self.ctx.alloc_outputs()
        ck = self.ctx.alloc_checkpoints()
        self.ctx.render_diff(bufs, ck)
        return bufs, ck

    def adjoint(self, theta, checkpoints, dL_drgb, pixel_grad=False):
        """dL/dtheta for the render at theta that produced `checkpoints`."""
        self._upload(theta)
        pg = None
        if pixel_grad:
            n = self.base.imageWidth * self.base.imageHeight
            pg = self.torch.empty((n, abi.MBX_DIFF_NPARAMS), dtype=self.torch.float64,
                                  device=dL_drgb.device)
        g = self.ctx.adjoint(checkpoints, dL_drgb.contiguous(), pg)
        return (g, pg) if pixel_grad else g

    def close(self):
        self.ctx.close()


def _autograd_function():
    import torch

    class RenderFunction(torch.autograd.Function):
        """image = RenderFunction.apply(theta, renderer): theta a float64 CPU tensor of
        MBX_DIFF_NPARAMS, image the float32 HxWx3 device tensor of the differentiable render."""

        @staticmethod
        def forward(ctx, theta, renderer):
            th = theta.detach().cpu().numpy().astype(np.float64)
            bufs, ck = renderer.render(th)
            ctx.renderer = renderer
            ctx.theta = th
            ctx.ck = ck
            return bufs["rgb"]

        @staticmethod
        def backward(ctx, grad_rgb):
            g = ctx.renderer.adjoint(ctx.theta, ctx.ck, grad_rgb.float().contiguous())
            return torch.from_numpy(g), None

    return RenderFunction


def render_function():
    """The torch.autograd.Function class (imported lazily so that torch stays optional)."""
    return _autograd_function()
