"""Confocal NLOS captures of O'Toole et al. 2018 (.mat), reference
totri/data/otooleconfocal2018.py:8-74."""
from __future__ import annotations

import typing

import numpy as np
import torch
from scipy.io import loadmat

from totrilib_amd.data.util import make_wall_grid
from totrilib_amd.types import BinDef, VolumeDef


class OTooleConfocal2018Dataset:
    """transient [1, N, SH, SW] and scan_points [1, SH, SW, 3] of a confocal
    capture, with its bins (4 ps, offset 100 bins, 2048 bins) and the volume
    of the published reconstructions."""
    __slots__ = ['is_confocal', 'bin_def', 'volume_def', 'scan_point_extent',
                 'transient', 'scan_points']

    def __init__(self):
        self.is_confocal = True
        self.bin_def = BinDef(4e-12, 100 * 4e-12, 2048, unit_is_time=True)
        self.volume_def = VolumeDef([-0.4, -0.4, 0.2], [0.4, 0.4, 1.0], [512, 64, 64])
        self.scan_point_extent = (2, 2)
        self.transient: typing.Optional[torch.Tensor] = None
        self.scan_points: typing.Optional[torch.Tensor] = None

    def downsample_temporal(self, factor: int) -> None:
        """Sum groups of `factor` consecutive bins (width grows by factor)."""
        if not isinstance(factor, int):
            raise ValueError('Downsample factor must be integer.')
        if factor < 1:
            raise ValueError('Downsample factor must be positive.')
        if factor == 1:
            return
        if self.transient.shape[1] % factor:
            raise ValueError('Number of bins must be divisible by factor.')
        out = torch.zeros_like(self.transient[:, ::factor])
        for i in range(factor):  # bins j*factor + i, added in phase order
            out += self.transient[:, i::factor]
        self.transient = out
        self.bin_def.width = self.bin_def.width * factor
        self.bin_def.num = self.bin_def.num // factor

    def downsample_spatial(self, resolution) -> None:
        """Bilinear resampling of transient and scan points to `resolution`
        (SH, SW)."""
        self.transient = torch.nn.functional.interpolate(
            self.transient, size=resolution, mode='bilinear', align_corners=False)
        self.scan_points = torch.nn.functional.interpolate(
            self.scan_points.permute(0, 3, 1, 2), size=resolution, mode='bilinear',
            align_corners=False).permute(0, 2, 3, 1)


def OTooleConfocal2018(path: str, dtype: torch.dtype = torch.float32,
                       device: torch.device = torch.device('cpu')) -> OTooleConfocal2018Dataset:
    """Load `rect_data` [W, H, N] of a .mat file as transient [1, N, H, W]
    (rows flipped, the first 600 bins zeroed as in the reference) with the
    64 x 64 wall grid over the volume's x, y extent."""
    ds = OTooleConfocal2018Dataset()
    data = loadmat(path)['rect_data'].astype(np.float32)        # [W, H, N]
    tr = torch.from_numpy(np.ascontiguousarray(data.transpose(2, 1, 0)))   # [N, H, W]
    tr = tr.flip([1])
    tr[:600] = 0
    ds.transient = tr[None].to(dtype=dtype, device=device)
    vd = ds.volume_def
    ds.scan_points = make_wall_grid(vd.start[0], vd.end[0], 64, vd.start[1], vd.end[1], 64,
                                    z_val=0, device=device, flattened=False)[None].to(dtype=dtype)
    return ds
