"""Exhaustive NLOS captures of Velten et al. 2012 (.mat), reference
totri/data/veltennc2012.py:8-140."""
from __future__ import annotations

import math
import typing

import torch
from scipy.io import loadmat

from totrilib_amd.types import BinDef, VolumeDef


def _DefaultTransformVeltenNc2012(x: torch.Tensor, is_normal: bool) -> torch.Tensor:
    """Coordinates [N, 3] of the capture frame into the volume frame:
    translate by (-3.75, -2.20, 0) m, rotate by 25 degrees about z, map
    (x, y, z) -> (y, z, -x) and move z by 2 m.  Normals are only rotated and
    permuted."""
    if not is_normal:
        x = x + torch.tensor([[-3.75, -2.20, 0.0]], device=x.device, dtype=x.dtype)
    a = 25.0 / 360.0 * 2.0 * math.pi
    rot = torch.tensor([[math.cos(a), -math.sin(a), 0.0],
                        [math.sin(a), math.cos(a), 0.0],
                        [0.0, 0.0, 1.0]], device=x.device, dtype=x.dtype)
    x = x @ rot.T
    x = torch.stack((x[:, 1], x[:, 2], -x[:, 0]), dim=1)
    if not is_normal:
        x = x + torch.tensor([[0.0, 0.0, 2.0]], device=x.device, dtype=x.dtype)
    return x


class VeltenNc2012Dataset:
    """transient [1, N, S, L], scan_points [1, S, 3], laser_points [1, L, 3],
    detector and laser origins [1, 3], bins and the volume of the capture."""
    __slots__ = ['transient', 'scan_points', 'laser_points', 'volume_def', 'bin_def',
                 'scan_origin', 'laser_origin']

    def rectify(self) -> None:
        """Roll each scan (laser) point's bins back by its origin-to-wall
        distance in whole bins, then drop the origin."""
        if self.scan_origin is not None:
            for s in range(self.scan_points.shape[1]):
                d = ((self.scan_points[:, s, :] - self.scan_origin) ** 2).sum() ** 0.5
                shift = int((d / self.bin_def.width).int().item())
                self.transient[0, :, s, :] = torch.roll(self.transient[0, :, s, :], -shift, 0)
            self.scan_origin = None
        if self.laser_origin is not None:
            for l in range(self.laser_points.shape[1]):
                d = ((self.laser_points[:, l, :] - self.laser_origin) ** 2).sum() ** 0.5
                shift = int((d / self.bin_def.width).int().item())
                self.transient[0, :, :, l] = torch.roll(self.transient[0, :, :, l], -shift, 0)
            self.laser_origin = None


def VeltenNc2012(path: str, transform: typing.Callable = _DefaultTransformVeltenNc2012,
                 rectify: bool = False,
                 device: torch.device = torch.device('cpu')) -> VeltenNc2012Dataset:
    """Load the `dataset` struct of a .mat file (fields deltat, t0, t, data
    [L, S*N], cameraPos [S, 3], laserPos [L, 3], cameraOrigin, laserOrigin; in
    cm, converted to m)."""
    ds = loadmat(path)['dataset']

    def field(name):
        return ds[name][0, 0]

    out = VeltenNc2012Dataset()
    n = int(field('t').item())
    out.bin_def = BinDef(field('deltat').item() / 100, field('t0').item() / 100, n)
    data = field('data')                                           # [L, S*N]
    tr = torch.tensor(data.reshape(1, data.shape[0], -1, n), dtype=torch.float32,
                      device=device)                               # [1, L, S, N]
    out.transient = tr.permute(0, 3, 2, 1)                         # [1, N, S, L]

    def points(name):
        p = torch.tensor(field(name) / 100, dtype=torch.float32, device=device)
        return transform(p, False)

    out.scan_points = points('cameraPos')[None]
    out.laser_points = points('laserPos')[None]
    out.scan_origin = points('cameraOrigin')
    out.laser_origin = points('laserOrigin')
    out.volume_def = VolumeDef((-0.75, 0.0, -0.44 + 2.0), (0.75, 3.0, 0.44 + 2.0), (44, 150, 75))
    if rectify:
        out.rectify()
    return out
