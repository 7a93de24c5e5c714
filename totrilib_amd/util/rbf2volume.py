"""Gaussian RBF -> volume splat as a torch.autograd.Function
(reference totri/util/rbf2volume.py:9-50)."""
from __future__ import annotations

import torch

from totrilib_amd.totri_cpp.util import rbf2volume_gaussian_forward, rbf2volume_gaussian_grad
from totrilib_amd.types import VolumeDef


class Rbf2VolumeGaussian(torch.autograd.Function):
    """rbf [B, N, 4+A] rows (x, y, z, sigma, attributes) -> volume
    [B, 1+A, D, H, W]: channel 0 the sum of the Gaussians, channels 1..A the
    Gaussian-weighted attribute means, on volume_def's inclusive grid.

    `truncation` (default 0: none, the reference's behaviour) skips RBFs
    farther than truncation * sigma from a voxel.
    """

    @staticmethod
    def forward(ctx, rbf: torch.Tensor, volume_def: VolumeDef,
                truncation: float = 0.0) -> torch.Tensor:
        num_attributes = rbf.shape[2] - 4
        shape = (rbf.shape[0], 1 + num_attributes) + tuple(volume_def.resolution)
        dev = rbf.device
        if rbf.shape[1] == 0:
            volume = torch.zeros(shape, device=dev, dtype=rbf.dtype)
        else:
            volume = torch.empty(shape, device=dev, dtype=rbf.dtype)
            rbf2volume_gaussian_forward(rbf, volume, volume_def.start, volume_def.end,
                                        truncation)
        ctx.set_materialize_grads(False)
        ctx.save_for_backward(rbf, volume)
        ctx.volume_def = volume_def
        ctx.truncation = truncation
        return volume

    @staticmethod
    def backward(ctx, volume_grad: torch.Tensor):
        rbf_grad = None
        if volume_grad is None:
            return rbf_grad, None, None
        if ctx.needs_input_grad[0]:
            rbf, volume = ctx.saved_tensors
            rbf_grad = torch.zeros_like(rbf)
            if rbf.shape[1] > 0:
                rbf2volume_gaussian_grad(rbf, volume, volume_grad.contiguous(), rbf_grad,
                                         ctx.volume_def.start, ctx.volume_def.end,
                                         ctx.truncation)
        if ctx.needs_input_grad[1]:  # volume_def
            raise NotImplementedError()
        return rbf_grad, None, None
