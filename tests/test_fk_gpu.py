"""f-k migration on the GPU (totrilib_amd.reconstruct.FkMigration and
StoltInterpolation, reference totri/reconstruct/fk_migration.py) against the
float64 oracle (oracle/fk.py) and, for the Stolt step, against torch's fp32
grid_sample, the routine the reference calls.

Tolerances: the GPU runs the FFTs and the interpolation in float32 (as the
reference does with complex64), the oracle in float64; relative L2 error of
the migrated volume <= 1e-4, and every voxel within 1e-3 of the maximum."""
import numpy as np
import pytest
import torch

from oracle import fk
from tests.test_oracle_fk import point_transient

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")


def rel_l2(a, b):
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-30))


def check_volume(got, want):
    assert got.shape == want.shape
    assert rel_l2(got, want) <= 1e-4
    assert np.abs(got - want).max() <= 1e-3 * np.abs(want).max()


@pytest.mark.parametrize("N,D,H,W,depth,height,width", [
    (1, 64, 16, 16, 0.64, 1.0, 1.0),
    (2, 96, 17, 15, 1.2, 0.8, 1.3),      # odd wall sizes; only entry 0 is returned
    (1, 1, 4, 4, 0.01, 1.0, 1.0),        # D = 1: the ramp is all zero
    (1, 200, 8, 40, 3.0, 0.5, 2.0),
    (1, 128, 1, 1, 1.0, 1.0, 1.0),
])
def test_fk_migration_matches_oracle(N, D, H, W, depth, height, width):
    from totrilib_amd.reconstruct import FkMigration
    print(f"FkMigration N={N} D={D} H={H} W={W}", flush=True)
    rng = np.random.default_rng(D * 7 + H)
    t = rng.random((N, D, H, W)).astype(np.float32)
    got = FkMigration(depth, height, width)(torch.from_numpy(t).to(DEV)).cpu().numpy()
    want = fk.fk_migration(t, depth, height, width)
    if D == 1:
        assert np.all(got == 0)
        return
    check_volume(got, want)


def test_fk_point_scatterer_on_gpu():
    from totrilib_amd.reconstruct import FkMigration
    D, H, W, bw, wall, p = 160, 27, 33, 0.01, 1.0, (-0.25, 0.15, 0.5)
    t, xs, ys = point_transient(D, H, W, wall, bw, p)
    t = t.astype(np.float32)
    got = FkMigration(D * bw, wall, wall)(torch.from_numpy(t).to(DEV)).cpu().numpy()
    check_volume(got, fk.fk_migration(t, D * bw, wall, wall))
    d, h, w = np.unravel_index(np.argmax(got[0]), got[0].shape)
    assert abs(d - 2 * p[2] / bw) <= 1.5
    assert abs(h - np.argmin(abs(ys - p[1]))) <= 1
    assert abs(w - np.argmin(abs(xs - p[0]))) <= 1


def test_fk_migration_full_size():
    """The bench's f-k workload: 1024 bins on a 64 x 64 wall."""
    from totrilib_amd.reconstruct import FkMigration
    rng = np.random.default_rng(1024)
    t = rng.random((1, 1024, 64, 64)).astype(np.float32)
    got = FkMigration(1024 * 0.0025, 2.0, 2.0)(torch.from_numpy(t).to(DEV)).cpu().numpy()
    check_volume(got, fk.fk_migration(t, 1024 * 0.0025, 2.0, 2.0))


@pytest.mark.parametrize("N,C,D2,H,W", [(2, 2, 32, 9, 12), (1, 1, 64, 16, 16)])
def test_stolt_matches_oracle_and_grid_sample(N, C, D2, H, W):
    from totrilib_amd.reconstruct.fk_migration import StoltInterpolation
    rng = np.random.default_rng(N + C + D2)
    f = (rng.standard_normal((N, C, D2, H, W))
         + 1j * rng.standard_normal((N, C, D2, H, W))).astype(np.complex64)
    s = StoltInterpolation(0.9, 0.7, 1.1)
    got = s(torch.from_numpy(f).to(DEV)).cpu().numpy()
    want = fk.stolt(f.astype(np.complex128), 0.9, 0.7, 1.1)
    assert rel_l2(got, want) <= 1e-5
    # torch fp32: grid_sample on the real and imaginary parts, then the weights
    grid, weights = s.get_buffers(D2 // 2, H, W, DEV)
    ft = torch.view_as_real(torch.from_numpy(f).to(DEV).view(N * C, D2, H, W))
    parts = [torch.nn.functional.grid_sample(ft[None, ..., i], grid, mode="bilinear",
                                             padding_mode="zeros", align_corners=False)
             for i in (0, 1)]
    ref = (torch.complex(parts[0], parts[1])[0] * weights[0]).view(N, C, D2, H, W)
    assert rel_l2(got, ref.cpu().numpy()) <= 1e-5
    assert np.all(got[:, :, :D2 // 2] == 0)


def test_stolt_rejects_bad_input():
    from totrilib_amd.reconstruct.fk_migration import StoltInterpolation
    s = StoltInterpolation(1.0, 1.0, 1.0)
    with pytest.raises(ValueError):
        s(torch.zeros((1, 1, 7, 4, 4), dtype=torch.complex64, device=DEV))
    with pytest.raises(ValueError):
        s(torch.zeros((1, 1, 8, 4, 4), device=DEV))
    with pytest.raises(ValueError):
        StoltInterpolation(1.0, 1.0, 0.0)(torch.zeros((1, 1, 8, 4, 4), dtype=torch.complex64,
                                                      device=DEV))
