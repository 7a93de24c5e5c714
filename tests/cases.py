"""Seeded synthetic inputs shared by the tests (geometry of the benchmark
configurations: wall z = 0, scan grid on [-1, 1]^2, volume box
x, y in [-1, 1], z in [0.5, 2.5])."""
from __future__ import annotations

import numpy as np
import torch

ORIGIN = (-1.0, -1.0, 0.5)
SIZE = (2.0, 2.0, 2.0)


def wall_grid(n: int) -> np.ndarray:
    """[3, n*n] float32 confocal scan points, x fastest (make_wall_grid)."""
    xs = torch.linspace(-1, 1, n)
    ys = torch.linspace(-1, 1, n)
    gy, gx = torch.meshgrid(ys, xs, indexing="ij")
    pts = torch.stack((gx, gy, torch.zeros_like(gx)), dim=2).view(-1, 3)
    return pts.T.contiguous().numpy().astype(np.float32)


def sparse_albedo(res, p: float, seed: int) -> np.ndarray:
    """Exactly round(p * V) occupied voxels (randperm, seed), values U(0, 1)."""
    X, Y, Z = res
    V = X * Y * Z
    g = torch.Generator().manual_seed(seed)
    k = int(round(p * V))
    idx = torch.randperm(V, generator=g)[:k]
    vals = torch.rand(k, generator=g)
    v = torch.zeros(V)
    v[idx] = vals
    return v.view(X, Y, Z).numpy().astype(np.float32)


def rel_l2(a, b) -> float:
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.linalg.norm(b.ravel())
    return float(np.linalg.norm((a - b).ravel()) / (den if den > 0 else 1.0))
