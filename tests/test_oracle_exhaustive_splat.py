"""Pins the oracle's exhaustive render (R2), its adjoint (A5) and the RBF
splat (S1, S2) independently of its C code.
Reference: src/render/volume_cuda.cu:297-377, src/util/rbf2volume_cuda.cu:15-201.
"""
import numpy as np

from oracle import oracle as O
from tests.cases import rel_l2

VS, VE = (-1.0, -1.0, 0.5), (1.0, 1.0, 2.5)


def test_r2_single_voxel_is_box_overlap_and_conserves_mass():
    """One occupied voxel spreads its occupancy over the bins covered by the
    path lengths of its lower and upper faces, proportionally to the overlap
    of [bf0, bf1] with each bin (the reference's first/inner/last weights);
    the total equals the occupancy when the span lies inside the histogram."""
    D, H, W, T = 10, 6, 7, 400
    vol = np.zeros((1, D, H, W), np.float32)
    xi, yi, zi, occ = 3, 4, 6, 0.7
    vol[0, zi, yi, xi] = occ
    scan = np.array([[[0.1], [-0.3], [0.0]]], np.float32)
    laser = np.array([[[-0.2], [0.25], [0.0]]], np.float32)
    w = 0.01
    got = O.r2_forward(vol, scan, laser, T, VS, VE, w, 0.0, mode=O.FP32)[0, :, 0, 0]
    x = VS[0] + (xi + 0.5) / W * (VE[0] - VS[0])
    y = VS[1] + (yi + 0.5) / H * (VE[1] - VS[1])
    zc = VS[2] + (zi + 0.5) / D * (VE[2] - VS[2])
    h = 0.5 / D * (VE[2] - VS[2])
    d = [np.linalg.norm(scan[0, :, 0] - [x, y, z]) + np.linalg.norm(laser[0, :, 0] - [x, y, z])
         for z in (zc - h, zc + h)]
    b0, b1 = min(d) / w, max(d) / w
    t = np.arange(T)
    want = occ * np.clip(np.minimum(t + 1, b1) - np.maximum(t, b0), 0, None) / (b1 - b0)
    assert np.count_nonzero(want) >= 2
    assert rel_l2(got, want) < 1e-5
    assert abs(got.sum() - occ) < 1e-6


def test_r2_adjoint_identity():
    rng = np.random.default_rng(41)
    B, D, H, W, T, S, L = 2, 6, 5, 4, 150, 3, 2
    vol = (rng.random((B, D, H, W)) * (rng.random((B, D, H, W)) < 0.4)).astype(np.float32)
    scan = np.concatenate([rng.uniform(-1, 1, (B, 2, S)), np.zeros((B, 1, S))], 1).astype(np.float32)
    laser = np.concatenate([rng.uniform(-1, 1, (B, 2, L)), np.zeros((B, 1, L))], 1).astype(np.float32)
    w = 0.04
    G = rng.standard_normal((B, T, S, L))
    fwd = O.r2_forward(vol, scan, laser, T, VS, VE, w, 0.0, mode=O.PINNED64)
    adj = O.r2_adjoint(G, (D, H, W), scan, laser, VS, VE, w, 0.0, mode=O.PINNED64)
    lhs = float(np.sum(fwd * G))
    rhs = float(np.sum(vol.astype(np.float64) * adj))
    # FP32 weights (occ / range vs 1 / range) limit the match to ~1e-6.
    assert abs(lhs - rhs) < 2e-6 * np.sqrt(np.sum(fwd ** 2) * np.sum(G ** 2))


def test_splat_single_rbf_closed_form():
    D, H, W = 9, 8, 7
    res = (D, H, W)
    # centre on grid node (xi, yi, zi) = (3, 5, 4)
    xs = np.float32(VS[0]) + np.float32(VE[0] - VS[0]) * np.float32(3) / np.float32(W - 1)
    ys = np.float32(VS[1]) + np.float32(VE[1] - VS[1]) * np.float32(5) / np.float32(H - 1)
    zs = np.float32(VS[2]) + np.float32(VE[2] - VS[2]) * np.float32(4) / np.float32(D - 1)
    sg = 0.37
    rbf = np.array([[[xs, ys, zs, sg, 0.6]]], np.float32)
    got = O.splat_forward(rbf, res, VS, VE, mode=O.FP32)[0]
    x = np.linspace(VS[0], VE[0], W)[None, None, :]
    y = np.linspace(VS[1], VE[1], H)[None, :, None]
    z = np.linspace(VS[2], VE[2], D)[:, None, None]
    want = np.exp(-0.5 * ((x - xs) ** 2 + (y - ys) ** 2 + (z - zs) ** 2) / sg ** 2)
    assert abs(got[0, 4, 5, 3] - 1.0) < 1e-7
    assert rel_l2(got[0], want) < 1e-6
    assert np.allclose(got[1], 0.6, rtol=1e-6)  # a single attribute reproduces itself


def test_splat_truncation_matches_reference_within_bound():
    """Truncating at k sigma changes the sum by at most the Gaussian tail
    exp(-k^2/2) per RBF; at 6 sigma it is below 1.6e-8 relative to the peak."""
    rng = np.random.default_rng(42)
    N = 30
    rbf = np.concatenate([rng.uniform(-1, 1, (1, N, 2)), rng.uniform(0.5, 2.5, (1, N, 1)),
                          rng.uniform(0.05, 0.2, (1, N, 1)), rng.random((1, N, 1))], 2).astype(np.float32)
    res = (12, 11, 10)
    full = O.splat_forward(rbf, res, VS, VE, trunc=0.0)
    t6 = O.splat_forward(rbf, res, VS, VE, trunc=6.0)
    t3 = O.splat_forward(rbf, res, VS, VE, trunc=3.0)
    assert np.max(np.abs(t6[0, 0] - full[0, 0])) <= N * np.exp(-18) + 1e-7
    assert np.max(np.abs(t3[0, 0] - full[0, 0])) <= N * np.exp(-4.5) + 1e-7
    assert np.max(np.abs(t3[0, 0] - full[0, 0])) > 0


def test_splat_gradient_matches_finite_differences():
    """S2 (rbf2volume_cuda.cu:123-201) against central differences of the
    pure FP64 splat, with attributes (A = 1) and a random upstream gradient."""
    rng = np.random.default_rng(43)
    N = 5
    rbf = np.concatenate([rng.uniform(-0.6, 0.6, (1, N, 2)), rng.uniform(1.0, 2.0, (1, N, 1)),
                          rng.uniform(0.2, 0.4, (1, N, 1)), rng.random((1, N, 1))], 2)
    rbf32 = rbf.astype(np.float32)
    res = (9, 8, 7)
    vol = O.splat_forward(rbf32, res, VS, VE, mode=O.PURE64)
    G = rng.standard_normal(vol.shape)
    ana = O.splat_grad(rbf32, vol, G, VS, VE, mode=O.PURE64)
    h = 1e-3
    for i in range(N):
        for c in range(5):
            rp = rbf32.astype(np.float64).copy()
            rm = rp.copy()
            rp[0, i, c] += h
            rm[0, i, c] -= h
            # the FP64 splat takes float32 parameters: use an exact float32 step
            rp32, rm32 = rp.astype(np.float32), rm.astype(np.float32)
            step = float(rp32[0, i, c]) - float(rm32[0, i, c])
            lp = np.sum(O.splat_forward(rp32, res, VS, VE, mode=O.PURE64) * G)
            lm = np.sum(O.splat_forward(rm32, res, VS, VE, mode=O.PURE64) * G)
            fd = (lp - lm) / step
            assert abs(fd - ana[0, i, c]) < 2e-4 * max(1.0, abs(ana[0, i, c])), (i, c, fd, ana[0, i, c])


def test_splat_voxel_subset_and_support_box_match_full_grid():
    """The subset form of the splat and the support-box walk of its gradient
    (used at full size) give the full-grid results exactly."""
    rng = np.random.default_rng(44)
    N = 40
    rbf = np.concatenate([rng.uniform(-1, 1, (1, N, 2)), rng.uniform(0.5, 2.5, (1, N, 1)),
                          rng.uniform(0.05, 0.3, (1, N, 1)), rng.random((1, N, 1))], 2).astype(np.float32)
    res = (14, 12, 11)
    for trunc in (0.0, 3.0):
        full = O.splat_forward(rbf, res, VS, VE, trunc=trunc, mode=O.PINNED64)
        vox = rng.choice(np.prod(res), 200, replace=False)
        sub = O.splat_forward(rbf, res, VS, VE, trunc=trunc, mode=O.PINNED64, vox=vox)
        flat_full = full.reshape(1, 2, -1)[:, :, vox]
        assert np.array_equal(sub.reshape(1, 2, -1)[:, :, vox], flat_full)
        assert np.count_nonzero(np.delete(sub.reshape(1, 2, -1), vox, axis=2)) == 0
    # One RBF at a time: its truncated gradient (support-box walk) equals the
    # untruncated gradient (whole-grid walk) of an upstream gradient that is
    # zero outside the truncation sphere, where the sphere test is restated
    # in float32 (zero terms add exactly).
    D, H, W = res
    def coord(i, n, a, b):
        t = np.float32(b - a) * np.float32(i)
        return np.float32(a) + t / np.float32(n - 1)
    z = coord(np.arange(D), D, VS[2], VE[2])[:, None, None]
    y = coord(np.arange(H), H, VS[1], VE[1])[None, :, None]
    x = coord(np.arange(W), W, VS[0], VE[0])[None, None, :]
    for i in range(0, N, 7):
        one = rbf[:, i:i + 1]
        vol = O.splat_forward(one, res, VS, VE, mode=O.PINNED64)
        G = rng.standard_normal(vol.shape)
        xr, yr, zr, sg = (np.float32(v) for v in one[0, 0, :4])
        r2 = (xr - x) * (xr - x)
        r2 = r2 + (yr - y) * (yr - y)
        r2 = r2 + (zr - z) * (zr - z)
        lim = np.float32(3.0) * sg
        inside = ~(r2 > lim * lim)
        assert 0 < inside.sum() < inside.size
        Gm = G * inside[None, None]
        a = O.splat_grad(one, vol, G, VS, VE, trunc=3.0, mode=O.FP32)
        b = O.splat_grad(one, vol, Gm, VS, VE, trunc=0.0, mode=O.FP32)
        assert np.array_equal(a, b), i
    # a wide truncation walks the whole grid: identical to no truncation
    full = O.splat_forward(rbf, res, VS, VE, mode=O.PINNED64)
    G = rng.standard_normal(full.shape)
    a = O.splat_grad(rbf, full, G, VS, VE, trunc=0.0, mode=O.PINNED64)
    b = O.splat_grad(rbf, full, G, VS, VE, trunc=1e4, mode=O.PINNED64)
    assert np.array_equal(a, b)
