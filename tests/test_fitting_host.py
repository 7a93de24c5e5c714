"""Blob sampling of the fitting step (reference totri/fitting/rbf.py:556-563)
on the CPU: positions are the left edges of the drawn half-resolution cells
with z moved by 0.5 m."""
import torch

from totrilib_amd.fitting import sample_blob_positions
from totrilib_amd.types import VolumeDef


def test_one_cell_position():
    # full box (8, 6, 4) -> half resolution (D, H, W) = (4, 3, 2)
    vd = VolumeDef((-1.0, -1.0, 0.5), (1.0, 1.0, 2.5), (8, 6, 4))
    m = torch.zeros(1, 4, 3, 2)
    m[0, 1, 2, 1] = 1.0                     # z index 1, y index 2, x index 1
    pos = sample_blob_positions(m, vd, 1, torch.Generator().manual_seed(0))
    # x: 1/2 * 2 - 1 = 0; y: 2/3 * 2 - 1 = 1/3; z: 1/4 * 2 + 0.5 = 1.0, + 0.5
    assert torch.allclose(pos, torch.tensor([[0.0, 1.0 / 3.0, 1.5]], dtype=torch.float64))


def test_draws_without_replacement():
    vd = VolumeDef((0.0, 0.0, 0.0), (4.0, 4.0, 4.0), (8, 8, 8))
    m = torch.zeros(1, 4, 4, 4)
    cells = [(0, 0, 0), (3, 1, 2), (2, 3, 3)]  # (z, y, x)
    for c in cells:
        m[(0,) + c] = 0.5
    pos = sample_blob_positions(m, vd, 3, torch.Generator().manual_seed(1))
    # every positive cell is drawn once: left edges are the indices (1 m cells)
    want = sorted((float(x), float(y), float(z) + 0.5) for z, y, x in cells)
    assert sorted(tuple(p) for p in pos.tolist()) == want

