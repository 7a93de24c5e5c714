"""Geometry descriptions shared by the renderers and backprojections.

Same classes, attributes and methods as the reference's totri/types.py
(VolumeDef :41-116, BinDef :118-157, C0 :4) so that call sites keep working.
"""
from __future__ import annotations

import typing

import torch

C0 = 299792458.0
"""Speed of light in m/s (bin widths given in seconds are scaled by it)."""

Vec3 = typing.Tuple[float, float, float]
Size3 = typing.Tuple[int, int, int]
Vec3_ = typing.Union[Vec3, Size3, int, float]
Size3_ = typing.Union[Size3, int]


def to_vec3(value: Vec3_) -> Vec3:
    """A scalar is broadcast to three entries; a 3-sequence is converted to floats."""
    if isinstance(value, (int, float)):
        return (value, value, value)
    if isinstance(value, (tuple, list)):
        if len(value) != 3:
            raise ValueError(f'Vec3 must be tuple of length 3, got {len(value)}')
        for v in value:
            if not isinstance(v, (int, float)):
                raise ValueError(f'Vec3 entries must be int or float, got {type(v)}')
        return tuple(float(v) for v in value)
    raise ValueError(f'Vec3 must be float or tuple, got {type(value)}')


def to_size3(value: Size3_) -> Size3:
    """A scalar int is broadcast to three entries; a 3-sequence must hold ints."""
    if isinstance(value, int):
        return (value, value, value)
    if isinstance(value, (tuple, list)):
        if len(value) != 3:
            raise ValueError(f'Size3 must be tuple of length 3, got {len(value)}')
        for v in value:
            if not isinstance(v, int):
                raise ValueError(f'Size3 entries must be int, got {type(v)}')
        return (value[0], value[1], value[2])
    raise ValueError(f'Size3 must be int or tuple, got {type(value)}')


class VolumeDef:
    """Axis-aligned box [start, end] (x, y, z in metres) sampled at
    `resolution` = (D, H, W) points along (z, y, x)."""
    __slots__ = ['start', 'end', 'resolution']

    def __init__(self, start: Vec3_, end: Vec3_, resolution: Size3_):
        self.start: Vec3 = to_vec3(start)
        self.end: Vec3 = to_vec3(end)
        self.resolution: Size3 = to_size3(resolution)

    def __str__(self):
        s = ', '.join(f'{v:.3e} m' for v in self.start)
        e = ', '.join(f'{v:.3e} m' for v in self.end)
        r = ', '.join(str(v) for v in self.resolution)
        return f'VolumeDef(start={s}, end={e}, resolution={r})'

    def clone(self) -> 'VolumeDef':
        return VolumeDef(self.start, self.end, self.resolution)

    def extent(self) -> Vec3:
        """(x, y, z) edge lengths of the box."""
        return tuple(abs(a - b) for a, b in zip(self.start, self.end))

    def corner_points(self) -> torch.Tensor:
        """The eight box corners as a [3, 8] tensor; x varies slowest, then z,
        then y (the reference's column order)."""
        sx, sy, sz = self.start
        ex, ey, ez = self.end
        pts = [(x, y, z) for x in (sx, ex) for z in (sz, ez) for y in (sy, ey)]
        return torch.Tensor(pts).T

    def flat_idx_to_3d_idx(self, idx):
        """Flat index over [D, H, W] -> (x, y, z) indices."""
        w = self.resolution[2]
        h = self.resolution[1]
        return (idx % w, (idx // w) % h, (idx // w) // h)

    def index_to_coordinates(self, idx: Size3_):
        """Left-edge coordinates idx / resolution of an (x, y, z) index."""
        idx = to_size3(idx)
        d, h, w = self.resolution
        return (
            idx[0] / w * (self.end[0] - self.start[0]) + self.start[0],
            idx[1] / h * (self.end[1] - self.start[1]) + self.start[1],
            idx[2] / d * (self.end[2] - self.start[2]) + self.start[2],
        )

    def coordinates_inclusive(self, resolution: Size3_ = None,
                              device: torch.device = torch.device('cpu')) -> torch.Tensor:
        """[D, H, W, 3] (x, y, z) coordinates with the first sample at `start`
        and the last at `end` along every axis."""
        if resolution is None:
            if self.resolution is None:
                raise ValueError('No resolution specified.')
            resolution = self.resolution
        else:
            resolution = to_size3(resolution)
        d, h, w = resolution
        x = torch.linspace(self.start[0], self.end[0], w, device=device)
        y = torch.linspace(self.start[1], self.end[1], h, device=device)
        z = torch.linspace(self.start[2], self.end[2], d, device=device)
        zz, yy, xx = torch.meshgrid(z, y, x, indexing='ij')
        return torch.stack((xx, yy, zz), dim=-1)


class BinDef:
    """Time bins: `width` and `offset` in metres of path length (seconds when
    unit_is_time, converted with C0), `num` bins."""
    __slots__ = ['width', 'offset', 'num']

    def __init__(self, width: float, offset: float = 0, num: typing.Optional[int] = None,
                 unit_is_time: bool = False):
        if unit_is_time:
            width *= C0
            offset *= C0
        self.width = width
        self.offset = offset
        self.num = num

    def __str__(self):
        return (f'BinDef(width={self.width / C0:.3e}s [{self.width:.3e}m], '
                f'offset={self.offset / C0:.3e}s [{self.offset:.3e}m], num={self.num})')

    @property
    def width_s(self):
        return self.width / C0

    @width_s.setter
    def width_s(self, value):
        self.width = value * C0

    @property
    def offset_s(self):
        return self.offset / C0

    @offset_s.setter
    def offset_s(self, value):
        self.offset = value * C0
