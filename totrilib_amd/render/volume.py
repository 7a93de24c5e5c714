"""Volume transient renderers as torch.autograd.Functions.

Mirrors totri/render/volume.py of the reference: same classes, argument order,
return shapes and needs_input_grad behaviour.  The forward and backward passes
run hand-written HIP kernels through totrilib_amd.totri_cpp.
"""
from __future__ import annotations

import typing

import torch

from totrilib_amd.totri_cpp.render import (
    volume_confocal_forward,
    volume_confocal_grad_measurement_points,
    volume_confocal_grad_volume,
    volume_render_exhaustive_forward,
    volume_render_exhaustive_grad_volume,
)


class VolumeRenderConfocal(torch.autograd.Function):
    """Confocal transient of an albedo volume (reference totri/render/volume.py:12-76).

    volume [X, Y, Z] (z fastest) covers the box volume_bbox_origin +
    [0, volume_bbox_size]; measurement_points [3, N] are confocal wall points;
    bin t collects the one-way radius t * bin_width + bin_offset.  Returns the
    transient [N, bin_num] (dtype of the volume, or `transient_dtype` --
    bfloat16 / float16 storage with FP32 accumulation).
    """

    @staticmethod
    def apply_transposed(transient: torch.Tensor, measurement_points: torch.Tensor,
                         volume_bbox_origin: typing.Tuple, volume_bbox_size: typing.Tuple,
                         volume_resolution: typing.Tuple, bin_width: float,
                         bin_offset: float, reference_window: bool = False) -> torch.Tensor:
        """Backprojection R^T applied to `transient` [N, T] -> volume [X, Y, Z]
        (float32), reference totri/render/volume.py:14-23.  The default is the
        exact transpose of the render; reference_window=True gives the
        reference's own window form (volume_cuda.cu:95-160), which also fills
        the top slice z = Z-1."""
        volume = torch.empty(tuple(volume_resolution), device=transient.device,
                             dtype=torch.float32)
        volume_confocal_grad_volume(volume, measurement_points, transient,
                                    volume_bbox_origin, volume_bbox_size,
                                    bin_width, bin_offset, reference_window=reference_window)
        return volume

    @staticmethod
    def forward(ctx, volume: torch.Tensor, measurement_points: torch.Tensor,
                volume_bbox_origin: typing.Tuple, volume_bbox_size: typing.Tuple,
                bin_num: int, bin_width: float, bin_offset: float,
                transient_dtype: typing.Optional[torch.dtype] = None):
        if volume.ndim != 3:
            raise ValueError('Volume tensor must be 3d')
        if not volume.is_cuda:
            raise NotImplementedError('VolumeRenderConfocal is only implemented for GPU tensors')
        if not measurement_points.is_cuda:
            raise NotImplementedError('VolumeRenderConfocal is only implemented for GPU tensors')
        if measurement_points.ndim != 2:
            raise ValueError('Measurement Point Tensor must be 2d')
        if measurement_points.shape[0] != 3:
            raise ValueError('Measurement Point Tensor must be 3xn')
        if len(volume_bbox_origin) != 3:
            raise ValueError('volume_bbox_origin must have 3 entries')
        if len(volume_bbox_size) != 3:
            raise ValueError('volume_bbox_size must have 3 entries')
        ctx.set_materialize_grads(False)
        ctx.save_for_backward(volume, measurement_points)
        ctx.volume_bbox_origin = volume_bbox_origin
        ctx.volume_bbox_size = volume_bbox_size
        ctx.bin_width = bin_width
        ctx.bin_offset = bin_offset
        dtype = volume.dtype if transient_dtype is None else transient_dtype
        transient = torch.empty((measurement_points.shape[1], bin_num),
                                device=volume.device, dtype=dtype)
        volume_confocal_forward(volume, measurement_points, transient,
                                volume_bbox_origin, volume_bbox_size,
                                bin_width, bin_offset)
        return transient

    @staticmethod
    def backward(ctx, transient_grad: torch.Tensor):
        if transient_grad is None:
            return (None,) * 8
        volume_grad = None
        measurement_points_grad = None
        volume, measurement_points = ctx.saved_tensors
        if ctx.needs_input_grad[0]:
            volume_grad = torch.empty_like(volume)
            volume_confocal_grad_volume(volume_grad, measurement_points, transient_grad,
                                        ctx.volume_bbox_origin, ctx.volume_bbox_size,
                                        ctx.bin_width, ctx.bin_offset)
        if ctx.needs_input_grad[1]:
            measurement_points_grad = torch.empty_like(measurement_points)
            volume_confocal_grad_measurement_points(
                measurement_points_grad, volume, measurement_points, transient_grad,
                ctx.volume_bbox_origin, ctx.volume_bbox_size, ctx.bin_width, ctx.bin_offset)
        return (volume_grad, measurement_points_grad) + (None,) * 6


def _origin_3xn(origin, points, name):
    """Origin broadcasting for [B, 3, N] wall points
    (reference totri/render/volume.py:113-130)."""
    if origin is None:
        return torch.empty_like(points[:, :, :0])
    if origin.ndim == 2:
        return origin[:, :, None].expand_as(points)
    if origin.ndim == 3:
        return origin.expand_as(points) if origin.shape[2] > 0 else origin
    raise ValueError(f'Shape {origin.shape} of {name} is invalid.')


class VolumeRenderExhaustive(torch.autograd.Function):
    """Render an occupancy volume to a non-confocal transient
    (reference totri/render/volume.py:78-170).

    volume [B, D, H, W], scan_points [B, 3, S], laser_points [B, 3, L],
    volume_def (start/end (x, y, z), resolution (D, H, W)), bin_def
    (width/offset of the scan+laser path length) -> transient [B, N, S, L].
    The origins are accepted and, as in the reference, ignored.  The backward
    pass returns R2's exact transpose for the volume (the reference's
    backward kernel is empty).
    """

    @staticmethod
    def forward(ctx, volume: torch.Tensor, scan_points: torch.Tensor,
                laser_points: torch.Tensor, volume_def, bin_def,
                scan_origin: typing.Optional[torch.Tensor] = None,
                laser_origin: typing.Optional[torch.Tensor] = None) -> torch.Tensor:
        if not torch.cuda.is_available():
            raise RuntimeError('VolumeRenderExhaustive needs a GPU, but no devices are available.')
        dev = volume.device if volume.is_cuda else torch.device('cuda', torch.cuda.current_device())
        transient = torch.zeros((volume.shape[0], bin_def.num, scan_points.shape[2],
                                 laser_points.shape[2]), device=dev)
        scan_origin = _origin_3xn(scan_origin, scan_points, 'scan_origin')
        laser_origin = _origin_3xn(laser_origin, laser_points, 'laser_origin')
        ctx.set_materialize_grads(False)
        ctx.save_for_backward(volume, scan_points, laser_points, scan_origin, laser_origin)
        ctx.volume_def = volume_def
        ctx.bin_def = bin_def
        volume_render_exhaustive_forward(
            volume.to(dev), scan_points.to(dev), laser_points.to(dev), scan_origin.to(dev),
            laser_origin.to(dev), transient, volume_def.start, volume_def.end,
            bin_def.width, bin_def.offset)
        return transient.to(volume.device)

    @staticmethod
    def backward(ctx, transient_grad: torch.Tensor):
        n_in = len(ctx.needs_input_grad)
        grad = [None] * n_in
        if transient_grad is None:
            return tuple(grad)
        volume, scan_points, laser_points, scan_origin, laser_origin = ctx.saved_tensors
        if ctx.needs_input_grad[0]:
            dev = transient_grad.device if transient_grad.is_cuda else \
                torch.device('cuda', torch.cuda.current_device())
            volume_grad = torch.empty(volume.shape, device=dev)
            volume_render_exhaustive_grad_volume(
                transient_grad.to(dev).float(), scan_points.to(dev), laser_points.to(dev),
                scan_origin.to(dev), laser_origin.to(dev), volume_grad,
                ctx.volume_def.start, ctx.volume_def.end, ctx.bin_def.width,
                ctx.bin_def.offset)
            grad[0] = volume_grad.to(volume.device)
        for k in range(1, n_in):
            if ctx.needs_input_grad[k]:
                raise NotImplementedError()
        return tuple(grad)
