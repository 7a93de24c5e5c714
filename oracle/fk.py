"""f-k migration restated in numpy, float64 -- TEST INFRASTRUCTURE ONLY.

Follows totri/reconstruct/fk_migration.py of the reference step by step:
  stolt_grid    :41-49   sample grid and weights of the Stolt interpolation
  grid_sample   :74-81   torch grid_sample, trilinear, align_corners=False,
                         zero padding (restated here, pinned against torch's)
  stolt         :66-85   interpolate the fftshift-ed spectrum and weight it
  fk_migration  :106-120 pad, FFT, Stolt, inverse FFT, magnitude of batch 0
The FFTs are numpy's.  Only tests, the smoke check and the bench's CPU
baseline use this module.
"""
from __future__ import annotations

import numpy as np


def stolt_grid(D: int, H: int, W: int, depth: float, height: float, width: float):
    """Grid coordinates (x, y, zz) and weights, each [2D, H, W]
    (fk_migration.py:41-49)."""
    x = np.linspace(-1, 1, W)[None, None, :]
    y = np.linspace(-1, 1, H)[None, :, None]
    z = np.linspace(-1, 1, 2 * D)[:, None, None]
    a = (W * depth) / (2 * D * width)
    b = (H * depth) / (2 * D * height)
    zz = np.sqrt(x ** 2 * a ** 2 + y ** 2 * b ** 2 + z ** 2)
    shape = (2 * D, H, W)
    wgt = (z > 0) * np.abs(z) / np.maximum(zz, 1e-6)
    return (np.broadcast_to(x, shape), np.broadcast_to(y, shape),
            np.broadcast_to(zz, shape), np.broadcast_to(wgt, shape))


def grid_sample(vol: np.ndarray, gx, gy, gz) -> np.ndarray:
    """Trilinear sampling of vol [C, Dz, Hy, Wx] (real or complex) at the
    normalised coordinates gx (along Wx), gy (Hy), gz (Dz) in [-1, 1], with
    align_corners=False ((g + 1) * size - 1) / 2 and zeros outside."""
    C, Dz, Hy, Wx = vol.shape
    ix = ((np.asarray(gx) + 1) * Wx - 1) / 2
    iy = ((np.asarray(gy) + 1) * Hy - 1) / 2
    iz = ((np.asarray(gz) + 1) * Dz - 1) / 2
    x0, y0, z0 = np.floor(ix), np.floor(iy), np.floor(iz)
    out = np.zeros((C,) + ix.shape, dtype=np.result_type(vol.dtype, np.float64))
    for dz in (0, 1):
        zi = z0 + dz
        wz = 1 - np.abs(iz - zi)
        for dy in (0, 1):
            yi = y0 + dy
            wy = 1 - np.abs(iy - yi)
            for dx in (0, 1):
                xi = x0 + dx
                wx = 1 - np.abs(ix - xi)
                ok = (xi >= 0) & (xi < Wx) & (yi >= 0) & (yi < Hy) & (zi >= 0) & (zi < Dz)
                xc = np.where(ok, xi, 0).astype(np.int64)
                yc = np.where(ok, yi, 0).astype(np.int64)
                zc = np.where(ok, zi, 0).astype(np.int64)
                out += np.where(ok, wx * wy * wz, 0) * vol[:, zc, yc, xc]
    return out


def stolt(fourier: np.ndarray, depth: float, height: float, width: float) -> np.ndarray:
    """StoltInterpolation.forward (fk_migration.py:56-85) on the fftshift-ed
    spectrum fourier [N, C, 2D, H, W]."""
    N, C, D2, H, W = fourier.shape
    assert D2 % 2 == 0
    gx, gy, gz, wgt = stolt_grid(D2 // 2, H, W, depth, height, width)
    f = fourier.reshape(N * C, D2, H, W)
    return (grid_sample(f, gx, gy, gz) * wgt).reshape(N, C, D2, H, W)


def fk_migration(transient: np.ndarray, depth: float, height: float,
                 width: float) -> np.ndarray:
    """FkMigration.forward (fk_migration.py:93-120): transient [N, D, H, W]
    -> the first batch entry's volume, [1, D, H, W]."""
    t = np.asarray(transient, dtype=np.float64)
    N, D, H, W = t.shape
    t = t[:, None]
    gz = np.linspace(0, 1, D)[None, None, :, None, None]
    t = np.sqrt(t) * gz
    t = np.concatenate((t, np.zeros_like(t)), axis=2)
    axes = (2, 3, 4)
    f = np.fft.fftshift(np.fft.fftn(t, axes=axes), axes=axes)
    f = stolt(f, depth, height, width)
    v = np.fft.ifftn(np.fft.ifftshift(f, axes=axes), axes=axes)
    return np.abs(v[:, :, :D])[0]
