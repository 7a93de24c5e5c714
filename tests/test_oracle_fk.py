"""CPU checks of the f-k migration oracle (oracle/fk.py, restating reference
totri/reconstruct/fk_migration.py) and of the host-side argument checks.

The sampler is pinned against torch's own grid_sample, the routine the
reference calls; the whole migration is pinned by its physics: the confocal
transient of a point scatterer, written here in closed form, migrates to a
volume whose maximum sits at the point (lateral grid cell, and depth bin
2 z / bin width)."""
import numpy as np
import pytest
import torch

from oracle import fk


@pytest.mark.parametrize("shape", [(3, 5, 6, 7), (2, 1, 4, 9), (1, 8, 1, 1)])
def test_grid_sample_matches_torch(shape):
    rng = np.random.default_rng(sum(shape))
    vol = rng.standard_normal(shape)
    g = rng.uniform(-1.4, 1.4, (4, 3, 5, 3))
    ref = torch.nn.functional.grid_sample(
        torch.from_numpy(vol)[None], torch.from_numpy(g)[None], mode="bilinear",
        padding_mode="zeros", align_corners=False)[0].numpy()
    got = fk.grid_sample(vol, g[..., 0], g[..., 1], g[..., 2])
    np.testing.assert_allclose(got, ref, rtol=0, atol=1e-12)
    # complex input: real and imaginary parts sampled alike
    cv = vol + 1j * vol[::-1]
    got = fk.grid_sample(cv, g[..., 0], g[..., 1], g[..., 2])
    ref_i = torch.nn.functional.grid_sample(
        torch.from_numpy(vol[::-1].copy())[None], torch.from_numpy(g)[None], mode="bilinear",
        padding_mode="zeros", align_corners=False)[0].numpy()
    np.testing.assert_allclose(got.imag, ref_i, rtol=0, atol=1e-12)


def point_transient(D, H, W, wall, bin_width, p):
    """Confocal transient [1, D, H, W] of one scatterer p = (x, y, z) seen
    from an H x W grid spanning [-wall/2, wall/2]^2 at z = 0: one bin at
    round-trip length 2r, amplitude 1/r^4."""
    xs = np.linspace(-wall / 2, wall / 2, W)
    ys = np.linspace(-wall / 2, wall / 2, H)
    t = np.zeros((1, D, H, W))
    for h in range(H):
        for w in range(W):
            r = np.sqrt((xs[w] - p[0]) ** 2 + (ys[h] - p[1]) ** 2 + p[2] ** 2)
            b = int(2 * r / bin_width)
            if b < D:
                t[0, b, h, w] += r ** -4
    return t, xs, ys


@pytest.mark.parametrize("D,H,W,p", [
    (128, 32, 32, (0.0, 0.0, 0.3)),
    (128, 32, 32, (0.2, -0.1, 0.4)),
    (160, 27, 33, (-0.25, 0.15, 0.5)),
])
def test_point_scatterer_focuses_at_the_point(D, H, W, p):
    bw, wall = 0.01, 1.0
    t, xs, ys = point_transient(D, H, W, wall, bw, p)
    v = fk.fk_migration(t, D * bw, wall, wall)
    assert v.shape == (1, D, H, W)
    d, h, w = np.unravel_index(np.argmax(v[0]), v[0].shape)
    assert abs(d - 2 * p[2] / bw) <= 1.5
    assert abs(h - np.argmin(abs(ys - p[1]))) <= 1
    assert abs(w - np.argmin(abs(xs - p[0]))) <= 1


def test_stolt_zeroes_the_negative_half_and_keeps_shape():
    rng = np.random.default_rng(5)
    f = rng.standard_normal((2, 1, 16, 6, 5)) + 1j * rng.standard_normal((2, 1, 16, 6, 5))
    g = fk.stolt(f, 0.5, 0.7, 0.9)
    assert g.shape == f.shape
    assert np.all(g[:, :, :8] == 0) and np.any(g[:, :, 8:] != 0)


def test_fk_host_checks_without_gpu():
    from totrilib_amd import _lib
    lib = _lib.load()
    rc = lib.totri_fk_stolt(None, 1, 7, 4, 4, 1.0, 1.0, 1.0, 0, 0, None, None)
    assert rc == _lib.TOTRI_EINVAL and "even" in _lib.last_error()
    rc = lib.totri_fk_stolt(None, 1, 8, 4, 4, 1.0, 1.0, 0.0, 0, 0, None, None)
    assert rc == _lib.TOTRI_EINVAL and "positive" in _lib.last_error()
    rc = lib.totri_fk_stolt(None, 1, 8, 4, 4, 1.0, 1.0, 1.0, 3, 0, None, None)
    assert rc == _lib.TOTRI_EINVAL and "layout" in _lib.last_error()
    assert lib.totri_fk_prepare(None, 0, 4, 4, None, 0, None) == _lib.TOTRI_OK
    assert lib.totri_fk_magnitude(None, 1, 4, 0, 4, None, None) == _lib.TOTRI_OK
    # the native FFT path takes powers of two up to H * W = 16384 and 2D = 4096
    assert lib.totri_fk_migration_workspace(1024, 64, 64) > 0
    assert lib.totri_fk_migration_workspace(2048, 128, 128) > 0
    for shape in [(96, 16, 16), (64, 12, 16), (4096, 4, 4), (64, 256, 128)]:
        assert lib.totri_fk_migration_workspace(*shape) == 0
    rc = lib.totri_fk_migration(None, 96, 16, 16, 1.0, 1.0, 1.0, None, None, 0, None)
    assert rc == _lib.TOTRI_EINVAL and "powers of two" in _lib.last_error()
    rc = lib.totri_fk_migration(None, 64, 16, 16, 1.0, 1.0, 1.0, None, None, 0, None)
    assert rc == _lib.TOTRI_EINVAL and "null" in _lib.last_error()
    if not torch.cuda.is_available():
        from totrilib_amd.reconstruct import FkMigration
        with pytest.raises(RuntimeError):
            FkMigration(1.0, 1.0, 1.0)(torch.zeros(1, 4, 4, 4))


def test_square_root_and_depth_ramp_of_the_prepass():
    rng = np.random.default_rng(9)
    t = rng.random((1, 24, 6, 5))
    v = fk.fk_migration(t, 0.5, 0.4, 0.4)
    # sqrt before the linear steps and |.| after: degree 1/2 homogeneity
    np.testing.assert_allclose(fk.fk_migration(4 * t, 0.5, 0.4, 0.4), 2 * v, rtol=1e-12,
                               atol=1e-14)
    # the ramp linspace(0, 1, D) is 0 at bin 0: a transient there migrates to nothing
    t0 = np.zeros_like(t)
    t0[:, 0] = rng.random((6, 5))
    assert np.abs(fk.fk_migration(t0, 0.5, 0.4, 0.4)).max() < 1e-15
