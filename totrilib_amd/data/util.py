"""Scan-grid helper, reference totri/data/util.py:6-37."""
from __future__ import annotations

import torch


def make_wall_grid(x_from=-1, x_to=1, x_res=32,
                   y_from=-1, y_to=1, y_res=32,
                   z_val=0, device=torch.device('cuda'),
                   x_fastest=True, flattened=True) -> torch.Tensor:
    """Regular wall grid of x_res * y_res points at depth z_val.

    Returns [y_res * x_res, 3] (x varying fastest when x_fastest) or, with
    flattened=False, [y_res, x_res, 3] ([x_res, y_res, 3] when not x_fastest).
    """
    xs = torch.linspace(x_from, x_to, x_res, device=device)
    ys = torch.linspace(y_from, y_to, y_res, device=device)
    if x_fastest:
        gy, gx = torch.meshgrid(ys, xs, indexing='ij')
    else:
        gx, gy = torch.meshgrid(xs, ys, indexing='ij')
    gz = torch.full_like(gx, fill_value=z_val)
    pts = torch.stack((gx, gy, gz), dim=2)
    return pts.view(-1, 3) if flattened else pts
