"""totri_cpp.render -- volume renderers and their adjoints.

Same names and positional arguments as the reference's pybind11 submodule
(src/totri.cpp:35-47); every function writes its output tensor in place and
returns None.  The OptiX mesh functions of that submodule are not provided.
"""
from __future__ import annotations

import typing

import torch

from totrilib_amd import _lib
from totrilib_amd.totri_cpp._common import (
    OutputSlot, require_f32, require_f64, require_gpu, storage_code, stream_ptr, workspace)

Vec3 = typing.Sequence[float]


def volume_confocal_forward(volume: torch.Tensor, measurement_points: torch.Tensor,
                            transient: torch.Tensor, volume_origin: Vec3,
                            volume_size: Vec3, bin_width: float,
                            bin_offset: float) -> None:
    """R1: transient[N, T] = render(volume[X, Y, Z], measurement_points[3, N]).

    Reference: src/render/volume.cpp:7-23, src/render/volume_cuda.cu:11-93.
    """
    dev = require_gpu(volume, measurement_points, transient,
                      names=("volume", "measurement_points", "transient"))
    if volume.dim() != 3 or measurement_points.dim() != 2 or transient.dim() != 2:
        raise ValueError("expected volume [X,Y,Z], measurement_points [3,N], transient [N,T]")
    if measurement_points.shape[1] != transient.shape[0]:
        # volume.cpp:14-16
        raise ValueError("Measurement and transient dimensions do not match.")
    if measurement_points.shape[0] != 3:
        raise ValueError("Measurement points must have shape 3xN.")
    if volume.dtype == torch.float64:
        # the reference's double dispatch (volume_cuda.cu:85)
        require_f64(measurement_points, "measurement_points")
        require_f64(transient, "transient")
        lib = _lib.load()
        X, Y, Z = (int(s) for s in volume.shape)
        N, T = (int(s) for s in transient.shape)
        vol, mp = volume.contiguous(), measurement_points.contiguous()
        out = OutputSlot(transient)
        with torch.cuda.device(dev):
            rc = lib.totri_volume_confocal_forward_f64(
                vol.data_ptr(), X, Y, Z, mp.data_ptr(), N, out.ptr(), T, _lib.f3(volume_origin),
                _lib.f3(volume_size), float(bin_width), float(bin_offset), stream_ptr(dev))
            _lib.check(rc, "volume_confocal_forward")
            out.finish()
        return
    require_f32(volume, "volume")
    require_f32(measurement_points, "measurement_points")
    code = storage_code(transient, "transient")
    lib = _lib.load()
    X, Y, Z = (int(s) for s in volume.shape)
    N, T = (int(s) for s in transient.shape)
    vol = volume.contiguous()
    mp = measurement_points.contiguous()
    out = OutputSlot(transient)
    with torch.cuda.device(dev):
        ws, wsp = workspace(lib.totri_volume_confocal_forward_workspace(X, Y, Z, N, T), dev)
        rc = lib.totri_volume_confocal_forward(
            vol.data_ptr(), X, Y, Z, mp.data_ptr(), N, out.ptr(), T, code,
            _lib.f3(volume_origin), _lib.f3(volume_size), float(bin_width),
            float(bin_offset), wsp, ws.numel() if ws is not None else 0,
            stream_ptr(dev))
        _lib.check(rc, "volume_confocal_forward")
        out.finish()


def volume_confocal_grad_volume(volume: torch.Tensor, measurement_points: torch.Tensor,
                                transient: torch.Tensor, volume_origin: Vec3,
                                volume_size: Vec3, bin_width: float,
                                bin_offset: float, reference_window: bool = False) -> None:
    """A1: volume[X, Y, Z] = R1^T(transient[N, T]); the exact transpose of R1.

    reference_window=True computes the reference's own window form instead
    (volume_cuda.cu:95-160): it adds, on the top slice z = Z-1, samples the
    render never reads, and decides each voxel's bins by its window.

    Reference: src/render/volume.cpp:25-41, src/render/volume_cuda.cu:95-186.
    """
    dev = require_gpu(volume, measurement_points, transient,
                      names=("volume", "measurement_points", "transient"))
    if volume.dim() != 3 or measurement_points.dim() != 2 or transient.dim() != 2:
        raise ValueError("expected volume [X,Y,Z], measurement_points [3,N], transient [N,T]")
    if measurement_points.shape[1] != transient.shape[0]:
        # volume.cpp:32-34
        raise ValueError("Measurement and transient dimensions do not match.")
    if measurement_points.shape[0] != 3:
        raise ValueError("Measurement points must have shape 3xN.")
    if volume.dtype == torch.float64:
        # the reference's double dispatch (volume_cuda.cu:175)
        require_f64(measurement_points, "measurement_points")
        require_f64(transient, "transient")
        lib = _lib.load()
        X, Y, Z = (int(s) for s in volume.shape)
        N, T = (int(s) for s in transient.shape)
        mp, tr = measurement_points.contiguous(), transient.contiguous()
        out = OutputSlot(volume)
        with torch.cuda.device(dev):
            rc = lib.totri_volume_confocal_grad_volume_f64(
                out.ptr(), X, Y, Z, mp.data_ptr(), N, tr.data_ptr(), T, _lib.f3(volume_origin),
                _lib.f3(volume_size), float(bin_width), float(bin_offset),
                1 if reference_window else 0, stream_ptr(dev))
            _lib.check(rc, "volume_confocal_grad_volume")
            out.finish()
        return
    require_f32(volume, "volume")
    require_f32(measurement_points, "measurement_points")
    code = storage_code(transient, "transient")
    lib = _lib.load()
    X, Y, Z = (int(s) for s in volume.shape)
    N, T = (int(s) for s in transient.shape)
    mp = measurement_points.contiguous()
    tr = transient.contiguous()
    out = OutputSlot(volume)
    if reference_window:
        with torch.cuda.device(dev):
            rc = lib.totri_volume_confocal_grad_volume_window(
                out.ptr(), X, Y, Z, mp.data_ptr(), N, tr.data_ptr(), T, code,
                _lib.f3(volume_origin), _lib.f3(volume_size), float(bin_width),
                float(bin_offset), stream_ptr(dev))
            _lib.check(rc, "volume_confocal_grad_volume_window")
            out.finish()
        return
    with torch.cuda.device(dev):
        ws, wsp = workspace(lib.totri_volume_confocal_grad_volume_workspace(X, Y, Z, N, T), dev)
        rc = lib.totri_volume_confocal_grad_volume(
            out.ptr(), X, Y, Z, mp.data_ptr(), N, tr.data_ptr(), T, code,
            _lib.f3(volume_origin), _lib.f3(volume_size), float(bin_width),
            float(bin_offset), wsp, ws.numel() if ws is not None else 0,
            stream_ptr(dev))
        _lib.check(rc, "volume_confocal_grad_volume")
        out.finish()


def volume_confocal_grad_measurement_points(measurement_points_grad: torch.Tensor,
                                            volume: torch.Tensor,
                                            measurement_points: torch.Tensor,
                                            transient_grad: torch.Tensor,
                                            volume_origin: Vec3, volume_size: Vec3,
                                            bin_width: float, bin_offset: float) -> None:
    """A6: measurement_points_grad[3, N] = d<G, R1(volume; w)>/dw.

    Reference: src/render/volume.cpp:43-81, src/render/volume_cuda.cu:188-295.
    """
    if volume.dtype == torch.float64 or measurement_points.dtype == torch.float64:
        raise TypeError("the scan-point gradient is provided in float32 only; the float64 "
                        "forms cover the render, its albedo adjoint and the Velten "
                        "backprojection")
    dev = require_gpu(measurement_points_grad, volume, measurement_points, transient_grad,
                      names=("measurement_points_grad", "volume", "measurement_points",
                             "transient_grad"))
    # volume.cpp:51-74
    if measurement_points_grad.dim() != 2:
        raise ValueError("Gradient must be 2D.")
    if volume.dim() != 3:
        raise ValueError("volume must be 3D.")
    if measurement_points.dim() != 2:
        raise ValueError("Measurement points must be 2D.")
    if transient_grad.dim() != 2:
        raise ValueError("Transient must be 2D.")
    if measurement_points.shape[0] != 3:
        raise ValueError("Measurement points must have shape 3xN.")
    if measurement_points_grad.shape[0] != 3:
        raise ValueError("Gradient must have shape 3xN.")
    if measurement_points.shape[1] != transient_grad.shape[0]:
        raise ValueError("Measurement points and transient dimensions do not match.")
    if measurement_points_grad.shape[1] != measurement_points.shape[1]:
        raise ValueError("Measurement points and gradient dimensions do not match.")
    require_f32(volume, "volume")
    require_f32(measurement_points, "measurement_points")
    require_f32(measurement_points_grad, "measurement_points_grad")
    code = storage_code(transient_grad, "transient_grad")
    lib = _lib.load()
    X, Y, Z = (int(s) for s in volume.shape)
    N, T = (int(s) for s in transient_grad.shape)
    vol = volume.contiguous()
    mp = measurement_points.contiguous()
    g = transient_grad.contiguous()
    out = OutputSlot(measurement_points_grad)
    with torch.cuda.device(dev):
        ws, wsp = workspace(
            lib.totri_volume_confocal_grad_measurement_points_workspace(X, Y, Z, N, T), dev)
        rc = lib.totri_volume_confocal_grad_measurement_points(
            out.ptr(), vol.data_ptr(), X, Y, Z, mp.data_ptr(), N, g.data_ptr(), T, code,
            _lib.f3(volume_origin), _lib.f3(volume_size), float(bin_width),
            float(bin_offset), wsp, ws.numel() if ws is not None else 0, stream_ptr(dev))
        _lib.check(rc, "volume_confocal_grad_measurement_points")
        out.finish()


def _exhaustive_render_shapes(volume_like_B, scan_points, laser_points):
    if scan_points.dim() != 3 or laser_points.dim() != 3 or scan_points.shape[1] != 3 \
            or laser_points.shape[1] != 3:
        raise ValueError("scan_points / laser_points must be [B,3,S] / [B,3,L]")
    B = int(scan_points.shape[0])
    if laser_points.shape[0] != B or volume_like_B != B:
        raise ValueError("batch sizes do not match")
    return B, int(scan_points.shape[2]), int(laser_points.shape[2])


def volume_render_exhaustive_forward(volume: torch.Tensor, scan_points: torch.Tensor,
                                     laser_points: torch.Tensor, scan_origin: torch.Tensor,
                                     laser_origin: torch.Tensor, transient: torch.Tensor,
                                     volume_start: Vec3, volume_end: Vec3, bin_width: float,
                                     bin_offset: float) -> None:
    """R2: transient[B, T, S, L] = exhaustive render of volume[B, D, H, W]
    (volume_cuda.cu:297-412).  The origins are accepted and ignored, as in
    the reference (volume_cuda.cu:312-313)."""
    dev = require_gpu(volume, scan_points, laser_points, transient,
                      names=("volume", "scan_points", "laser_points", "transient"))
    if volume.dim() != 4 or transient.dim() != 4:
        raise ValueError("expected volume [B,D,H,W] and transient [B,T,S,L]")
    B, S, L = _exhaustive_render_shapes(int(volume.shape[0]), scan_points, laser_points)
    if tuple(transient.shape[::2]) != (B, S) or transient.shape[3] != L:
        raise ValueError(f"transient must be [{B},T,{S},{L}], got {tuple(transient.shape)}")
    for t, n in ((volume, "volume"), (scan_points, "scan_points"),
                 (laser_points, "laser_points"), (transient, "transient")):
        require_f32(t, n)
    D, H, W = (int(s) for s in volume.shape[1:])
    T = int(transient.shape[1])
    vol = volume.contiguous()
    sp, lp = scan_points.contiguous(), laser_points.contiguous()
    lib = _lib.load()
    out = OutputSlot(transient)
    with torch.cuda.device(dev):
        ws, wsp = workspace(lib.totri_volume_render_exhaustive_forward_workspace(B, D, H, W), dev)
        rc = lib.totri_volume_render_exhaustive_forward(
            vol.data_ptr(), B, D, H, W, sp.data_ptr(), S, lp.data_ptr(), L, out.ptr(), T,
            _lib.f3(volume_start), _lib.f3(volume_end), float(bin_width), float(bin_offset),
            wsp, ws.numel() if ws is not None else 0, stream_ptr(dev))
        _lib.check(rc, "volume_render_exhaustive_forward")
        out.finish()


def volume_render_exhaustive_grad_volume(transient_grad: torch.Tensor,
                                         scan_points: torch.Tensor,
                                         laser_points: torch.Tensor,
                                         scan_origin: torch.Tensor,
                                         laser_origin: torch.Tensor,
                                         volume_grad: torch.Tensor, volume_start: Vec3,
                                         volume_end: Vec3, bin_width: float,
                                         bin_offset: float) -> None:
    """A5: volume_grad[B, D, H, W] = R2^T(transient_grad[B, T, S, L]).  The
    reference's kernel for this entry is empty (volume_cuda.cu:414-431)."""
    dev = require_gpu(transient_grad, scan_points, laser_points, volume_grad,
                      names=("transient_grad", "scan_points", "laser_points", "volume_grad"))
    if volume_grad.dim() != 4 or transient_grad.dim() != 4:
        raise ValueError("expected transient_grad [B,T,S,L] and volume_grad [B,D,H,W]")
    B, S, L = _exhaustive_render_shapes(int(volume_grad.shape[0]), scan_points, laser_points)
    if transient_grad.shape[0] != B or transient_grad.shape[2] != S \
            or transient_grad.shape[3] != L:
        raise ValueError("transient_grad does not match the scan/laser points")
    for t, n in ((transient_grad, "transient_grad"), (scan_points, "scan_points"),
                 (laser_points, "laser_points"), (volume_grad, "volume_grad")):
        require_f32(t, n)
    D, H, W = (int(s) for s in volume_grad.shape[1:])
    T = int(transient_grad.shape[1])
    g = transient_grad.contiguous()
    sp, lp = scan_points.contiguous(), laser_points.contiguous()
    lib = _lib.load()
    out = OutputSlot(volume_grad)
    with torch.cuda.device(dev):
        ws, wsp = workspace(lib.totri_volume_render_exhaustive_grad_volume_workspace(B, T, S, L),
                            dev)
        rc = lib.totri_volume_render_exhaustive_grad_volume(
            g.data_ptr(), B, T, S, L, sp.data_ptr(), lp.data_ptr(), out.ptr(), D, H, W,
            _lib.f3(volume_start), _lib.f3(volume_end), float(bin_width), float(bin_offset),
            wsp, ws.numel() if ws is not None else 0, stream_ptr(dev))
        _lib.check(rc, "volume_render_exhaustive_grad_volume")
        out.finish()
