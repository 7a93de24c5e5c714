"""Backprojection reconstructions as torch.autograd.Functions.

Mirrors totri/reconstruct/backprojection.py of the reference: same classes,
argument order and return shapes; kernels run on the GPU through
totrilib_amd.totri_cpp.reconstruct.  Like the reference these Functions have
no backward.
"""
from __future__ import annotations

import typing

import torch

from totrilib_amd.totri_cpp.reconstruct import (
    backprojection_confocal_velten_forward,
    backprojection_exhaustive_arellano_forward,
    backprojection_exhaustive_velten_forward,
)
from totrilib_amd.types import BinDef, VolumeDef
from totrilib_amd.util.format import origin_of_wallpoints


def second_derivative_filter(volume: torch.Tensor) -> torch.Tensor:
    """max(0, -d^2/dz^2 volume) along dim 1 of [B, D, H, W], with the last
    slice repeated at the end (reference backprojection.py:12-18)."""
    def dz(v):
        v = torch.cat((v, v[:, -1:, :, :]), dim=1)
        return v[:, 1:, :, :] - v[:, :-1, :, :]
    v = dz(dz(volume))
    v[v > 0] = 0
    return -v


def _gpu_device(t: torch.Tensor) -> torch.device:
    return t.device if t.is_cuda else torch.device('cuda', torch.cuda.current_device())


def _origin_3xn(origin, points):
    """Origin broadcasting for the [B, 3, N] point layout
    (reference backprojection.py:57-72)."""
    if origin is None:
        return torch.empty_like(points[:, :, :0])
    if origin.ndim == 2:
        return origin[:, :, None].expand_as(points)
    if origin.ndim == 3:
        return origin.expand_as(points) if origin.shape[2] > 0 else origin
    raise ValueError(f'Shape {origin.shape} of origin is invalid.')


class BackprojectionExhaustiveArellano(torch.autograd.Function):
    """Probability volume by ellipsoid backprojection (Arellano et al. 2017).

    transient [B, N, S, L], scan_points [B, 3, S], laser_points [B, 3, L],
    origins [B, 3, S] / [B, 3, L] / [B, 3] / None -> volume [B, D, H, W].
    The reference's forward lacks `ctx` (backprojection.py:48), which makes its
    .apply unusable; here forward takes ctx like every autograd Function.
    """

    @staticmethod
    def forward(ctx, transient: torch.Tensor, scan_points: torch.Tensor,
                laser_points: torch.Tensor, volume_def: VolumeDef, bin_def: BinDef,
                scan_origin: typing.Optional[torch.Tensor] = None,
                laser_origin: typing.Optional[torch.Tensor] = None) -> torch.Tensor:
        if not torch.cuda.is_available():
            raise RuntimeError('BackprojectionExhaustiveArellano needs a GPU, '
                               'but no devices are available.')
        dev = _gpu_device(transient)
        volume = torch.zeros((transient.shape[0],) + tuple(volume_def.resolution), device=dev)
        scan_origin = _origin_3xn(scan_origin, scan_points)
        laser_origin = _origin_3xn(laser_origin, laser_points)
        backprojection_exhaustive_arellano_forward(
            transient.to(dev), scan_points.to(dev), laser_points.to(dev),
            scan_origin.to(dev), laser_origin.to(dev), volume,
            volume_def.start, volume_def.end, bin_def.width, bin_def.offset)
        return volume.to(transient.device)


class BackprojectionConfocalVelten(torch.autograd.Function):
    """Confocal backprojection (Velten et al. 2012).

    transient [B, T, S], scan_points [B, S, 3], scan_origin [B, S, 3] /
    [B, 1, 3] / [B, 3] / None -> volume [B, D, H, W] on the transient's device.
    """

    @staticmethod
    def forward(ctx, transient: torch.Tensor, scan_points: torch.Tensor,
                volume_def: VolumeDef, bin_def: BinDef,
                scan_origin: typing.Optional[torch.Tensor] = None) -> torch.Tensor:
        if not torch.cuda.is_available():
            raise RuntimeError('BackprojectionConfocalVelten needs a GPU, '
                               'but no devices are available.')
        dev = _gpu_device(transient)
        # float32 as in the reference; a float64 transient gets a float64
        # volume (the reference's kernel dispatches double, but its wrapper
        # allocates float32 and so cannot run it)
        dtype = torch.float64 if transient.dtype == torch.float64 else torch.float32
        volume = torch.zeros((transient.shape[0],) + tuple(volume_def.resolution), device=dev,
                             dtype=dtype)
        scan_origin = origin_of_wallpoints(scan_origin, scan_points, 'scan_origin')
        backprojection_confocal_velten_forward(
            transient.to(dev), scan_points.to(dev), scan_origin.to(dev), volume,
            volume_def.start, volume_def.end, bin_def.width, bin_def.offset)
        return volume.to(transient.device)


class BackprojectionExhaustiveVelten(torch.autograd.Function):
    """Exhaustive (non-confocal) Velten backprojection.

    transient [B, T, S, L], scan_points [B, S, 3], laser_points [B, L, 3],
    origins [B, S|L, 3] / [B, 1, 3] / [B, 3] / None -> volume [B, D, H, W].
    """

    @staticmethod
    def forward(ctx, transient: torch.Tensor, scan_points: torch.Tensor,
                laser_points: torch.Tensor, volume_def: VolumeDef, bin_def: BinDef,
                scan_origin: typing.Optional[torch.Tensor] = None,
                laser_origin: typing.Optional[torch.Tensor] = None) -> torch.Tensor:
        if not torch.cuda.is_available():
            raise RuntimeError('BackprojectionExhaustiveVelten needs a GPU, '
                               'but no devices are available.')
        dev = _gpu_device(transient)
        volume = torch.zeros((transient.shape[0],) + tuple(volume_def.resolution), device=dev)
        scan_origin = origin_of_wallpoints(scan_origin, scan_points, 'scan_origin')
        laser_origin = origin_of_wallpoints(laser_origin, laser_points, 'laser_origin')
        backprojection_exhaustive_velten_forward(
            transient.to(dev), scan_points.to(dev), laser_points.to(dev),
            scan_origin.to(dev), laser_origin.to(dev), volume,
            volume_def.start, volume_def.end, bin_def.width, bin_def.offset)
        return volume.to(transient.device)
