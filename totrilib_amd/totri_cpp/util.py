"""totri_cpp.util -- Gaussian RBF splat.

Same names and positional arguments as the reference's pybind11 submodule
(src/totri.cpp:50-53).  `truncation` (keyword, default 0 = none) is an
addition: RBFs farther than truncation * sigma from a voxel are skipped.
"""
from __future__ import annotations

import typing

import torch

from totrilib_amd import _lib
from totrilib_amd.totri_cpp._common import (OutputSlot, require_f32, require_gpu,
                                            stream_ptr, workspace)

Vec3 = typing.Sequence[float]


def _check(rbf: torch.Tensor, volume: torch.Tensor):
    if rbf.dim() != 3 or rbf.shape[2] < 4:
        raise ValueError("rbf must be [B, N, 4+A]")
    if volume.dim() != 5:
        raise ValueError("volume must be [B, 1+A, D, H, W]")
    B, N, C = (int(s) for s in rbf.shape)
    A = C - 4
    if volume.shape[0] != B or volume.shape[1] != 1 + A:
        raise ValueError(f"volume must be [{B}, {1 + A}, D, H, W], got {tuple(volume.shape)}")
    return B, N, A


def rbf2volume_gaussian_forward(rbf: torch.Tensor, volume: torch.Tensor, volume_start: Vec3,
                                volume_end: Vec3, truncation: float = 0.0) -> None:
    """S1: volume[B, 1+A, D, H, W] from rbf[B, N, 4+A] (rbf2volume_cuda.cu:49-120)."""
    dev = require_gpu(rbf, volume, names=("rbf", "volume"))
    B, N, A = _check(rbf, volume)
    require_f32(rbf, "rbf")
    require_f32(volume, "volume")
    D, H, W = (int(s) for s in volume.shape[2:])
    r = rbf.contiguous()
    lib = _lib.load()
    out = OutputSlot(volume)
    with torch.cuda.device(dev):
        ws, wsp = workspace(lib.totri_rbf2volume_gaussian_forward_workspace(
            B, N, A, D, H, W, float(truncation)), dev)
        rc = lib.totri_rbf2volume_gaussian_forward(
            r.data_ptr(), B, N, A, out.ptr(), D, H, W, _lib.f3(volume_start),
            _lib.f3(volume_end), float(truncation), wsp, ws.numel() if ws is not None else 0,
            stream_ptr(dev))
        _lib.check(rc, "rbf2volume_gaussian_forward")
        out.finish()


def rbf2volume_gaussian_grad(rbf: torch.Tensor, volume: torch.Tensor,
                             volume_grad: torch.Tensor, rbf_grad: torch.Tensor,
                             volume_start: Vec3, volume_end: Vec3,
                             truncation: float = 0.0) -> None:
    """S2: rbf_grad[B, N, 4+A] (written) from the forward output and its
    gradient (rbf2volume_cuda.cu:123-232)."""
    dev = require_gpu(rbf, volume, volume_grad, rbf_grad,
                      names=("rbf", "volume", "volume_grad", "rbf_grad"))
    B, N, A = _check(rbf, volume)
    if tuple(volume_grad.shape) != tuple(volume.shape) or tuple(rbf_grad.shape) != tuple(rbf.shape):
        raise ValueError("volume_grad / rbf_grad shapes must match volume / rbf")
    for t, n in ((rbf, "rbf"), (volume, "volume"), (volume_grad, "volume_grad"),
                 (rbf_grad, "rbf_grad")):
        require_f32(t, n)
    D, H, W = (int(s) for s in volume.shape[2:])
    r, v, g = rbf.contiguous(), volume.contiguous(), volume_grad.contiguous()
    lib = _lib.load()
    out = OutputSlot(rbf_grad)
    with torch.cuda.device(dev):
        rc = lib.totri_rbf2volume_gaussian_grad(
            r.data_ptr(), B, N, A, v.data_ptr(), g.data_ptr(), out.ptr(), D, H, W,
            _lib.f3(volume_start), _lib.f3(volume_end), float(truncation), stream_ptr(dev))
        _lib.check(rc, "rbf2volume_gaussian_grad")
        out.finish()
