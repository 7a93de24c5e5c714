"""Capture-format loaders (reference totri/data/pfm.py, otooleconfocal2018.py,
veltennc2012.py) on synthetic files written in those formats.  Expected values
come from explicit index loops and closed forms, not from the loaders' code."""
import math

import numpy as np
import pytest
import torch
from scipy.io import savemat

from totrilib_amd.data import OTooleConfocal2018, PfmTransient, VeltenNc2012
from totrilib_amd.types import C0


def test_pfm_roundtrip_and_header_checks(tmp_path):
    vals = np.arange(24, dtype=np.float32) * 0.5 - 3
    p = tmp_path / "t.pfm"
    p.write_bytes(b"Pf\n6 4\n-1\n" + vals.astype("<f4").tobytes())
    t = PfmTransient(str(p), (2, 3, 4))
    assert t.shape == (2, 3, 4) and torch.equal(t.flatten(), torch.from_numpy(vals))
    bad = tmp_path / "b.pfm"
    bad.write_bytes(b"P6\n1 1\n-1\n" + vals[:1].tobytes())
    with pytest.raises(IOError):
        PfmTransient(str(bad), (1,))
    big = tmp_path / "e.pfm"
    big.write_bytes(b"Pf\n1 1\n1\n" + vals[:1].tobytes())
    with pytest.raises(IOError):
        PfmTransient(str(big), (1,))


def test_otoole_layout_and_downsampling(tmp_path):
    rng = np.random.default_rng(0)
    W, H, N = 4, 3, 704
    rect = rng.random((W, H, N)).astype(np.float32)
    p = tmp_path / "o.mat"
    savemat(str(p), {"rect_data": rect})
    ds = OTooleConfocal2018(str(p))
    tr = ds.transient
    assert tr.shape == (1, N, H, W)
    for t in (0, 599, 600, 703):
        for h in range(H):
            for w in range(W):
                want = 0.0 if t < 600 else rect[w, H - 1 - h, t]
                assert tr[0, t, h, w].item() == want
    assert ds.scan_points.shape == (1, 64, 64, 3)
    assert ds.scan_points[0, 0, 0].tolist() == pytest.approx([-0.4, -0.4, 0.0])
    assert ds.scan_points[0, 0, 63].tolist() == pytest.approx([0.4, -0.4, 0.0])
    assert ds.bin_def.width == pytest.approx(4e-12 * C0)
    assert ds.bin_def.offset == pytest.approx(400e-12 * C0)
    before = ds.transient.clone()
    ds.downsample_temporal(2)
    assert ds.transient.shape == (1, N // 2, H, W) and ds.bin_def.num == 1024
    assert torch.equal(ds.transient[0, 310], before[0, 620] + before[0, 621])
    assert ds.bin_def.width == pytest.approx(8e-12 * C0)
    ds.downsample_spatial((2, 2))
    assert ds.transient.shape == (1, N // 2, 2, 2) and ds.scan_points.shape == (1, 2, 2, 3)
    with pytest.raises(ValueError):
        ds.downsample_temporal(3)


def _velten_mat(path, S=5, L=3, N=40, seed=1):
    rng = np.random.default_rng(seed)
    data = rng.random((L, S * N)).astype(np.float64)
    cam = np.stack([np.full(S, 375.0) + np.arange(S), np.full(S, 220.0), np.zeros(S)], 1)
    las = np.stack([np.full(L, 375.0), np.full(L, 220.0) + np.arange(L), np.zeros(L)], 1)
    savemat(str(path), {"dataset": {"deltat": 0.5, "t0": 2.0, "t": float(N), "data": data,
                                    "cameraPos": cam, "laserPos": las,
                                    "cameraOrigin": np.array([[375.0, 220.0, 100.0]]),
                                    "laserOrigin": np.array([[375.0, 220.0, 50.0]])}})
    return data, S, L, N


def test_velten_layout_units_and_transform(tmp_path):
    p = tmp_path / "v.mat"
    data, S, L, N = _velten_mat(p)
    ds = VeltenNc2012(str(p))
    assert (ds.bin_def.width, ds.bin_def.offset, ds.bin_def.num) == pytest.approx((0.005, 0.02, N))
    assert ds.transient.shape == (1, N, S, L)
    for n in (0, 17, N - 1):
        for s in range(S):
            for l in range(L):
                assert ds.transient[0, n, s, l].item() == pytest.approx(data[l, s * N + n])
    # (3.75, 2.20, 0) m is the translation's fixed point: it maps to (0, 0, 2)
    assert ds.scan_points[0, 0].tolist() == pytest.approx([0.0, 0.0, 2.0], abs=1e-6)
    # 1 cm further along x: (0.01, 0, 0) rotated by 25 deg, then (y, z, -x) + (0, 0, 2)
    c, s_ = math.cos(25 * math.pi / 180), math.sin(25 * math.pi / 180)
    assert ds.scan_points[0, 1].tolist() == pytest.approx([0.01 * s_, 0.0, 2.0 - 0.01 * c],
                                                          abs=1e-6)
    # origins: 1 m and 0.5 m above the translation's fixed point -> (0, 1, 2), (0, 0.5, 2)
    assert ds.scan_origin[0].tolist() == pytest.approx([0.0, 1.0, 2.0], abs=1e-6)
    assert ds.laser_origin[0].tolist() == pytest.approx([0.0, 0.5, 2.0], abs=1e-6)


def test_velten_rectify_rolls_by_whole_bins(tmp_path):
    p = tmp_path / "v.mat"
    data, S, L, N = _velten_mat(p)
    raw = VeltenNc2012(str(p))
    rect = VeltenNc2012(str(p), rectify=True)
    assert rect.scan_origin is None and rect.laser_origin is None
    for s in range(S):
        ds_ = float(((raw.scan_points[0, s] - raw.scan_origin[0]) ** 2).sum() ** 0.5)
        for l in range(L):
            dl = float(((raw.laser_points[0, l] - raw.laser_origin[0]) ** 2).sum() ** 0.5)
            shift = int(ds_ / 0.005) + int(dl / 0.005)
            want = torch.roll(raw.transient[0, :, s, l], -shift, 0)
            assert torch.allclose(rect.transient[0, :, s, l], want)
