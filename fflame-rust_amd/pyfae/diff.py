"""Differentiable soft splat for PyTorch-ROCm.

    fits = [(flame_desc, soft_job(256, 256, ...)), ...]
    splat = SoftSplat(engine, fits)
    theta = splat.theta0(device="cuda").requires_grad_()
    density = splat(theta)              # (total pixels,) float32, fits back to back
    loss = ((torch.log1p(density) - torch.log1p(target)) ** 2).sum()
    loss.backward()                     # theta.grad from ffm_softsplat_adjoint

The forward and the adjoint are the library's gfx950 kernels
(ffm_softsplat_forward / ffm_softsplat_adjoint); torch only holds the
tensors and orders the launches on its current stream. theta packs, per
fit, the transforms' (a, b, c, d, e, f) then their variation weights.
"""
import numpy as np
import torch

from . import EngineError, soft_theta


def _stream_launch(splat, fn, args, tensors):
    """Run fn(*args, stream=...) ordered after and before torch's work on the
    current stream. torch's default stream is the null stream, which the
    library reads as "the context's own stream", so that case goes through a
    side stream joined to torch's."""
    cur = torch.cuda.current_stream(tensors[0].device)
    if cur.cuda_stream != 0:
        fn(*args, stream=cur.cuda_stream)
        return
    side = splat._side_stream(tensors[0].device)
    side.wait_stream(cur)
    fn(*args, stream=side.cuda_stream)
    cur.wait_stream(side)
    for t in tensors:
        t.record_stream(side)


class _SoftSplatFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, theta, splat):
        splat._activate()
        density = torch.empty(splat.density_floats, device=theta.device, dtype=torch.float32)
        _stream_launch(splat, splat.engine.softsplat_forward, (theta, density), (theta, density))
        ctx.splat = splat
        ctx.save_for_backward(theta)
        return density

    @staticmethod
    def backward(ctx, grad_density):
        (theta,) = ctx.saved_tensors
        splat = ctx.splat
        splat._activate()
        g = grad_density.detach().to(torch.float32).contiguous()
        grad_theta = torch.empty_like(theta)
        _stream_launch(splat, splat.engine.softsplat_adjoint, (theta, g, grad_theta),
                       (theta, g, grad_theta))
        return grad_theta, None


class SoftSplat:
    """A set of soft-splat fits on one Engine; calling it with theta (float32
    device tensor of theta_floats) returns the densities as a differentiable
    tensor."""

    def __init__(self, engine, fits):
        self.engine = engine
        self.fits = [(fd, job) for fd, job in fits]
        self.theta_floats, self.density_floats = engine.softsplat_set(self.fits)
        engine._soft_owner = self
        self._side = None

    def _activate(self):
        if getattr(self.engine, "_soft_owner", None) is not self:
            self.engine.softsplat_set(self.fits)
            self.engine._soft_owner = self

    def _side_stream(self, device):
        if self._side is None:
            self._side = torch.cuda.Stream(device)
        return self._side

    def theta0(self, device="cuda"):
        """theta as the flame descriptors hold it."""
        th = np.concatenate([soft_theta(fd) for fd, _ in self.fits]).astype(np.float32)
        return torch.from_numpy(th).to(device)

    def shapes(self):
        """(height, width) of each fit's density, in order."""
        return [(job.view.height, job.view.width) for _, job in self.fits]

    def split(self, density):
        """Per-fit (H, W) views of a flat density tensor."""
        out, off = [], 0
        for h, w in self.shapes():
            out.append(density[off:off + h * w].view(h, w))
            off += h * w
        return out

    def __call__(self, theta):
        if not (isinstance(theta, torch.Tensor) and theta.is_cuda and
                theta.dtype == torch.float32 and theta.is_contiguous() and
                theta.numel() == self.theta_floats):
            raise EngineError(-1, "theta must be a contiguous float32 device tensor of %d values"
                              % self.theta_floats)
        return _SoftSplatFn.apply(theta, self)


def log_density_l2(density, target):
    """The fit's loss: L2 distance of log(1 + density) to log(1 + target)."""
    d = torch.log1p(density) - torch.log1p(target)
    return (d * d).sum()
