"""Pins the oracle's backprojections (A2, A3, A4) independently of its C code:
vectorised float32 restatements of the nearest-bin formula (bit-exact), the
impulse response (a spherical/ellipsoidal shell of voxels) and, for the
Arellano splat, the closed-form sphere case.
Reference: src/reconstruct/backprojection_cuda.cu:10-287.
"""
import numpy as np

from oracle import oracle as O

f32 = np.float32
VS, VE = (-1.0, -1.0, 0.5), (1.0, 1.0, 2.5)


def inclusive32(n, a, b):
    i = np.arange(n, dtype=f32)
    t = (i / f32(n - 1)).astype(f32)
    return (f32(a) + (t * (f32(b) - f32(a))).astype(f32)).astype(f32)


def grid32(res):
    D, H, W = res
    z = inclusive32(D, VS[2], VE[2])[:, None, None]
    y = inclusive32(H, VS[1], VE[1])[None, :, None]
    x = inclusive32(W, VS[0], VE[0])[None, None, :]
    return np.broadcast_arrays(x, y, z)


def dist32(p, qx, qy, qz):
    dx = (f32(p[0]) - qx).astype(f32)
    dy = (f32(p[1]) - qy).astype(f32)
    dz = (f32(p[2]) - qz).astype(f32)
    s = (dx * dx).astype(f32)
    s = (s + (dy * dy).astype(f32)).astype(f32)
    s = (s + (dz * dz).astype(f32)).astype(f32)
    return np.sqrt(s).astype(f32)


def nearest_bin32(dist, w, off):
    tf = ((dist - f32(off)).astype(f32) / f32(w)).astype(f32)
    tf = (tf + f32(0.5)).astype(f32)
    return np.floor(tf).astype(np.int64)


def velten_confocal32(tr, scan, res, w, off):
    """Vectorised float32 restatement of A2 without origins."""
    qx, qy, qz = grid32(res)
    B, T, S = tr.shape
    out = np.zeros((B,) + tuple(res), f32)
    for b in range(B):
        acc = np.zeros(res, f32)
        for s in range(S):
            dsv = dist32(scan[b, s], qx, qy, qz)
            dist = (f32(2) * f32(0) + (f32(2) * dsv).astype(f32)).astype(f32)
            t = nearest_bin32(dist, w, off)
            ok = (t >= 0) & (t < T)
            add = np.where(ok, tr[b, np.clip(t, 0, T - 1), s].astype(f32), f32(0))
            acc = np.where(ok, (acc + add).astype(f32), acc)
        out[b] = acc
    return out


def test_velten_confocal_bit_exact_vs_float32_restatement():
    rng = np.random.default_rng(31)
    B, T, S = 2, 160, 7
    res = (9, 7, 8)
    tr = rng.random((B, T, S)).astype(f32)
    scan = np.concatenate([rng.uniform(-1, 1, (B, S, 2)), np.zeros((B, S, 1))], 2).astype(f32)
    w, off = 0.04, 0.3
    got = O.velten_confocal(tr, scan, None, res, VS, VE, w, off, mode=O.FP32)
    want = velten_confocal32(tr, scan, res, w, off)
    assert np.array_equal(got.astype(f32), want)


def test_velten_confocal_impulse_is_a_shell():
    """A single transient sample (t0, s0) backprojects to the voxels whose
    round-trip bin for s0 is t0: a spherical shell around the scan point."""
    T, S = 300, 4
    res = (16, 16, 16)
    scan = np.array([[[0.1, -0.2, 0.0], [0.5, 0.5, 0.0], [-0.5, 0.3, 0.0], [0, 0, 0]]], f32)
    tr = np.zeros((1, T, S), f32)
    t0, s0 = 150, 1
    tr[0, t0, s0] = 1.0
    w = 0.02
    got = O.velten_confocal(tr, scan, None, res, VS, VE, w, 0.0)[0]
    qx, qy, qz = grid32(res)
    want = nearest_bin32((f32(2) * dist32(scan[0, s0], qx, qy, qz)).astype(f32), w, 0.0) == t0
    assert want.sum() > 10
    assert np.array_equal(got > 0, want)
    # geometry: the shell radius is t0 * w / 2 within half a bin
    r = np.sqrt((qx - scan[0, s0, 0]) ** 2 + (qy - scan[0, s0, 1]) ** 2 + (qz - scan[0, s0, 2]) ** 2)
    assert np.all(np.abs(2 * r[want] - t0 * w) <= w / 2 + 1e-6)


def test_velten_exhaustive_matches_confocal_when_laser_equals_scan():
    """With one laser point equal to the scan point and zero origins,
    |p_s - q| + |p_l - q| = 2|p_s - q|, so A3 reduces to A2 per (s, l=s)."""
    rng = np.random.default_rng(32)
    T, S = 120, 1
    res = (6, 5, 7)
    scan = np.array([[[0.2, 0.1, 0.0]]], f32)
    tr3 = rng.random((1, T, S)).astype(f32)
    a2 = O.velten_confocal(tr3, scan, None, res, VS, VE, 0.05, 0.0)
    a3 = O.velten_exhaustive(tr3[..., None], scan, scan, None, None, res, VS, VE, 0.05, 0.0)
    # 2*d rounds identically to d + d in binary floating point.
    assert np.array_equal(a2, a3)


def test_arellano_sphere_case_closed_form():
    """Scan = laser and no origins: the ellipsoid is a sphere of radius
    tc/2 around the wall point, so each column's hit lies at
    z = cz + sqrt(r^2 - rho^2) (or the lower root); the voxel index is
    round-half-up((z - zs)/(ze - zs) * D) (the reference truncates
    d / len * D + 0.5)."""
    T, S = 240, 1
    res = (32, 9, 9)
    D = res[0]
    p = np.array([0.05, -0.1, 0.0], f32)
    scan = p.reshape(1, 3, 1)
    tr = np.zeros((1, T, S, 1), np.float64)
    tr[0, 150, 0, 0] = 1.0  # tc = 150 * 0.02 = 3.0 -> radius 1.5
    w = 0.02
    got = O.arellano(tr, scan, scan, None, None, res, VS, VE, w, 0.0)[0]
    rad = 150 * w / 2
    xs = np.linspace(VS[0], VE[0], res[2])
    ys = np.linspace(VS[1], VE[1], res[1])
    want = np.zeros(res)
    for yi, y in enumerate(ys):
        for xi, x in enumerate(xs):
            rho2 = (x - p[0]) ** 2 + (y - p[1]) ** 2
            if rad ** 2 - rho2 <= 0:
                continue
            h = np.sqrt(rad ** 2 - rho2)
            zs, ze = VS[2] - p[2], VE[2] - p[2]
            d = zs - (-h) if -h >= zs else None  # lower root along the ray
            d_lo = -h - zs
            d_hi = h - zs
            d = d_lo if d_lo >= 0 else d_hi
            b = int(d / (ze - zs) * D + 0.5)
            if 0 <= b < D:
                want[b, yi, xi] += 1.0
    mism = np.sum((got > 0) != (want > 0))
    assert np.count_nonzero(want) > 40
    assert mism <= 2, mism
