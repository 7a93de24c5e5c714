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
                         bin_offset: float) -> torch.Tensor:
        """Backprojection R^T applied to `transient` [N, T] -> volume [X, Y, Z]
        (float32), reference totri/render/volume.py:14-23."""
        volume = torch.empty(tuple(volume_resolution), device=transient.device,
                             dtype=torch.float32)
        volume_confocal_grad_volume(volume, measurement_points, transient,
                                    volume_bbox_origin, volume_bbox_size,
                                    bin_width, bin_offset)
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
