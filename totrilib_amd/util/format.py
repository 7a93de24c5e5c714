"""Wall-point origin broadcasting, reference totri/util/format.py:8-35."""
from __future__ import annotations

import typing

import torch


def origin_of_wallpoints(origin: typing.Optional[torch.Tensor],
                         wall_points: torch.Tensor,
                         origin_name: str) -> torch.Tensor:
    """Bring the origin of each wall point to [B, N', 3].

    None gives an empty [B, 0, 3] tensor ("ignore the origin-to-wall
    distance"); a [B, 3] origin is shared by all wall points; a [B, N, 3] or
    [B, 1, 3] origin is broadcast to the wall points; [B, 0, 3] stays empty.
    """
    if origin is None:
        return torch.empty_like(wall_points[:, :0, :])
    if origin.ndim == 2:
        return origin[:, None, :].expand_as(wall_points)
    if origin.ndim == 3:
        if origin.shape[1] > 0:
            return origin.expand_as(wall_points)
        return origin
    raise ValueError(f'Shape {origin.shape} of {origin_name} is invalid.')
