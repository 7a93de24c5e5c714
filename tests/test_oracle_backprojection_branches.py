"""Pins the oracle's backprojection branches that a scan point equal to the
laser point, or no origin, leaves open:

* A2's scan-origin term 2*|p_s - o_s| (backprojection_cuda.cu:178-187),
* A3's laser term |p_l - q| and both origin terms d_so + d_lo (:255-279),
* A4's ellipsoid transform for scan != laser and its origin subtraction
  tc = t*w + off - d_so - d_lo (:38-54, :79-101).

Each pin is independent of oracle/totri_oracle.c: a float32 numpy restatement
(bit-exact), an exact float64 bin prediction decoded from a bin-coded
transient (value t+1 in bin t, so the volume reveals the bin each voxel
took), or a brute-force line/ellipsoid root search in float64.  Each test also
shows that its check fails for a planted mistake -- laser point ignored, an
origin term dropped or halved, the ellipsoid scale skipped -- so a shared
misreading of those branches cannot pass.
"""
import numpy as np
import pytest

from oracle import oracle as O

f32 = np.float32
VS, VE = (-1.0, -1.0, 0.5), (1.0, 1.0, 2.5)


# ----------------------------------------------------------------------------
# grids and distances
# ----------------------------------------------------------------------------
def inclusive32(n, a, b):
    """backprojection_cuda.cu:168-170: start + (i / (n-1)) * (end - start)."""
    i = np.arange(n, dtype=f32)
    t = (i / f32(n - 1)).astype(f32)
    return (f32(a) + (t * (f32(b) - f32(a))).astype(f32)).astype(f32)


def grid32(res):
    D, H, W = res
    z = inclusive32(D, VS[2], VE[2])[:, None, None]
    y = inclusive32(H, VS[1], VE[1])[None, :, None]
    x = inclusive32(W, VS[0], VE[0])[None, None, :]
    return [a.astype(f32) for a in np.broadcast_arrays(x, y, z)]


def dist32(p, qx, qy, qz):
    dx = (f32(p[0]) - qx).astype(f32)
    dy = (f32(p[1]) - qy).astype(f32)
    dz = (f32(p[2]) - qz).astype(f32)
    s = (dx * dx).astype(f32)
    s = (s + (dy * dy).astype(f32)).astype(f32)
    s = (s + (dz * dz).astype(f32)).astype(f32)
    return np.sqrt(s).astype(f32)


def dist64(p, qx, qy, qz):
    p = np.asarray(p, np.float64)
    return np.sqrt((p[0] - np.float64(qx)) ** 2 + (p[1] - np.float64(qy)) ** 2
                   + (p[2] - np.float64(qz)) ** 2)


def nearest_bin32(dist, w, off):
    tf = ((dist - f32(off)).astype(f32) / f32(w)).astype(f32)
    tf = (tf + f32(0.5)).astype(f32)
    return np.floor(tf).astype(np.int64)


# ----------------------------------------------------------------------------
# A3 (and A2) as vectorised float32 restatements with every origin term
# ----------------------------------------------------------------------------
def velten_exhaustive32(tr, scan, laser, so, lo, res, w, off):
    """A3 in float32, operation for operation:
    dist = ((d_so + d_sv) + d_lv) + d_lo; t = floor((dist - off)/w + 0.5)."""
    qx, qy, qz = grid32(res)
    B, T, S, L = tr.shape
    out = np.zeros((B,) + tuple(res), f32)
    for b in range(B):
        acc = np.zeros(res, f32)
        for s in range(S):
            dsv = dist32(scan[b, s], qx, qy, qz)
            dso = f32(0) if so is None else dist32(scan[b, s], *[f32(c) for c in so[b, s]])
            for l in range(L):
                dlv = dist32(laser[b, l], qx, qy, qz)
                dlo = f32(0) if lo is None else dist32(laser[b, l], *[f32(c) for c in lo[b, l]])
                dist = (f32(dso) + dsv).astype(f32)
                dist = (dist + dlv).astype(f32)
                dist = (dist + f32(dlo)).astype(f32)
                t = nearest_bin32(dist, w, off)
                ok = (t >= 0) & (t < T)
                add = np.where(ok, tr[b, np.clip(t, 0, T - 1), s, l].astype(f32), f32(0))
                acc = np.where(ok, (acc + add).astype(f32), acc)
        out[b] = acc
    return out


def velten_confocal32(tr, scan, so, res, w, off):
    """A2 in float32: dist = 2*d_so + 2*d_sv (both doublings exact)."""
    qx, qy, qz = grid32(res)
    B, T, S = tr.shape
    out = np.zeros((B,) + tuple(res), f32)
    for b in range(B):
        acc = np.zeros(res, f32)
        for s in range(S):
            dsv = dist32(scan[b, s], qx, qy, qz)
            dso = f32(0) if so is None else dist32(scan[b, s], *[f32(c) for c in so[b, s]])
            dist = ((f32(2) * f32(dso)) + (f32(2) * dsv).astype(f32)).astype(f32)
            t = nearest_bin32(dist, w, off)
            ok = (t >= 0) & (t < T)
            add = np.where(ok, tr[b, np.clip(t, 0, T - 1), s].astype(f32), f32(0))
            acc = np.where(ok, (acc + add).astype(f32), acc)
        out[b] = acc
    return out


def _wall(rng, n):
    return np.concatenate([rng.uniform(-1, 1, (1, n, 2)), np.zeros((1, n, 1))], 2).astype(f32)


def test_velten_exhaustive_with_distinct_lasers_and_origins_bit_exact():
    rng = np.random.default_rng(71)
    B, T, S, L = 2, 260, 3, 2
    res = (9, 8, 7)
    tr = rng.random((B, T, S, L)).astype(f32)
    scan = np.concatenate([_wall(rng, S) for _ in range(B)])
    laser = np.concatenate([_wall(rng, L) for _ in range(B)])
    so = (scan + rng.uniform(-0.4, 0.4, scan.shape) - np.array([0, 0, 1.0])).astype(f32)
    lo = (laser + rng.uniform(-0.4, 0.4, laser.shape) - np.array([0, 0, 0.7])).astype(f32)
    w, off = 0.03, 0.25
    for sorg, lorg in ((so, lo), (so, None), (None, lo)):
        got = O.velten_exhaustive(tr, scan, laser, sorg, lorg, res, VS, VE, w, off, mode=O.FP32)
        want = velten_exhaustive32(tr, scan, laser, sorg, lorg, res, w, off)
        assert np.array_equal(got.astype(f32), want)
    # planted mistakes in the restatement must be seen: laser point ignored,
    # laser origin dropped
    got = O.velten_exhaustive(tr, scan, laser, so, lo, res, VS, VE, w, off, mode=O.FP32)
    no_laser = velten_exhaustive32(tr, scan, scan[:, :1].repeat(L, 1), so, lo, res, w, off)
    no_dlo = velten_exhaustive32(tr, scan, laser, so, None, res, w, off)
    assert not np.array_equal(got.astype(f32), no_laser)
    assert not np.array_equal(got.astype(f32), no_dlo)


def test_velten_confocal_with_origin_bit_exact():
    rng = np.random.default_rng(72)
    B, T, S = 2, 300, 5
    res = (8, 9, 7)
    tr = rng.random((B, T, S)).astype(f32)
    scan = np.concatenate([_wall(rng, S) for _ in range(B)])
    so = (scan + rng.uniform(-0.3, 0.3, scan.shape) - np.array([0, 0, 0.9])).astype(f32)
    w, off = 0.02, 0.1
    got = O.velten_confocal(tr, scan, so, res, VS, VE, w, off, mode=O.FP32)
    assert np.array_equal(got.astype(f32), velten_confocal32(tr, scan, so, res, w, off))


# ----------------------------------------------------------------------------
# bin-coded transients: the volume reveals each voxel's bin; compared with an
# exact float64 prediction of the published formula
# ----------------------------------------------------------------------------
TOL = 2e-4   # |u - round(u)| below which float32 rounding may flip a bin


def _predicted_bins(dist64_, w, off, T):
    u = (dist64_ - off) / w + 0.5
    t = np.floor(u).astype(np.int64)
    near = np.abs(u - np.round(u)) < TOL
    t = np.where((t >= 0) & (t < T), t, -1)
    return t, near


def _check_bins(got, want, near):
    bad = (got != want) & ~near
    return int(bad.sum())


def test_velten_confocal_origin_term_is_twice_the_origin_distance():
    T = 700
    res = (12, 10, 11)
    w, off = 0.01, 0.05
    qx, qy, qz = grid32(res)
    for p, o in (([0.1, -0.2, 0.0], [0.1, -0.2, -0.75]),
                 ([-0.35, 0.4, 0.0], [0.45, -0.3, -1.1])):
        scan = np.array([[p]], f32)
        so = np.array([[o]], f32)
        tr = (np.arange(T, dtype=f32) + 1)[None, :, None]
        vol = O.velten_confocal(tr, scan, so, res, VS, VE, w, off, mode=O.FP32)[0]
        got = vol.astype(np.int64) - 1
        dso = dist64(np.float32(p), np.float32(o[0]), np.float32(o[1]), np.float32(o[2]))
        dsv = dist64(np.float32(p), qx, qy, qz)
        want, near = _predicted_bins(2 * dso + 2 * dsv, w, off, T)
        assert np.count_nonzero(want >= 0) > 100
        assert _check_bins(got, want, near) == 0
        # planted: the origin counted once, or not at all
        for wrong in (dso + 2 * dsv, 2 * dsv):
            w_bins, w_near = _predicted_bins(wrong, w, off, T)
            assert _check_bins(got, w_bins, near | w_near) > 100


def _laser_case():
    """One scan point, two distinct laser points, origins for all three."""
    scan = np.array([[[0.15, -0.25, 0.0]]], f32)
    laser = np.array([[[-0.6, 0.5, 0.0], [0.7, 0.35, 0.0]]], f32)
    so = np.array([[[0.15, -0.25, -0.5]]], f32)            # d_so = 0.5 exactly
    lo = np.array([[[-0.1, 0.2, -0.9], [0.7, 0.35, -0.625]]], f32)
    return scan, laser, so, lo


def test_velten_exhaustive_laser_and_origin_terms():
    """tr[t, 0, l] = (t + 1) * 4096**l: the volume holds t0+1 + 4096*(t1+1),
    the bins taken for laser 0 and laser 1 (0 for out of range).  Each must be
    floor((d_so + |p_s-q| + |p_l-q| + d_lo - off)/w + 1/2) of float64."""
    T = 1200
    res = (11, 9, 10)
    w, off = 0.005, 0.3
    scan, laser, so, lo = _laser_case()
    t = np.arange(T, dtype=np.float64) + 1
    tr = np.stack([t, t * 4096.0], 1)[None, :, None, :]   # [1, T, 1, 2]
    assert tr.max() < 2 ** 24
    qx, qy, qz = grid32(res)
    dsv = dist64(scan[0, 0], qx, qy, qz)
    dso = dist64(scan[0, 0], *so[0, 0])
    for sorg, lorg in ((so, lo), (None, lo), (so, None)):
        vol = O.velten_exhaustive(tr, scan, laser, sorg, lorg, res, VS, VE, w, off, mode=O.FP32)[0]
        v = np.rint(vol).astype(np.int64)
        assert np.array_equal(v.astype(np.float64), vol)
        got = [v % 4096 - 1, v // 4096 - 1]
        d_so = dso if sorg is not None else 0.0
        for l in range(2):
            dlv = dist64(laser[0, l], qx, qy, qz)
            d_lo = dist64(laser[0, l], *lo[0, l]) if lorg is not None else 0.0
            want, near = _predicted_bins(d_so + dsv + dlv + d_lo, w, off, T)
            assert np.count_nonzero(want >= 0) > 200
            assert _check_bins(got[l], want, near) == 0, (l, sorg is None, lorg is None)
            # planted mistakes: laser point ignored, laser origin dropped,
            # scan origin dropped
            wrongs = [d_so + 2 * dsv + d_lo]
            if lorg is not None:
                wrongs.append(d_so + dsv + dlv)
            if sorg is not None:
                wrongs.append(dsv + dlv + d_lo)
            for wrong in wrongs:
                w_bins, w_near = _predicted_bins(wrong, w, off, T)
                assert _check_bins(got[l], w_bins, near | w_near) > 50


# ----------------------------------------------------------------------------
# A4: ellipsoid with foci p_s, p_l and |q - p_s| + |q - p_l| = tc
# ----------------------------------------------------------------------------
def _roots(g, lo=-3.0, hi=4.0, n=28001):
    """All sign changes of g on [lo, hi], refined by bisection (float64)."""
    f = np.linspace(lo, hi, n)
    v = g(f)
    idx = np.nonzero(np.sign(v[:-1]) * np.sign(v[1:]) < 0)[0]
    out = []
    for i in idx:
        a, b = f[i], f[i + 1]
        ga = g(np.array([a]))[0]
        for _ in range(60):
            m = 0.5 * (a + b)
            gm = g(np.array([m]))[0]
            if np.sign(gm) == np.sign(ga):
                a, ga = m, gm
            else:
                b = m
        out.append(0.5 * (a + b))
    return out


def arellano_hits64(p_s, p_l, tc, res, scale=True):
    """Per column, the voxel index the reference's ray/ellipsoid rule picks:
    the first intersection of the vertical ray (parameter f from z=start to
    z=end) with {q : |q-p_s| + |q-p_l| = tc}, or the second if the first lies
    before the ray start; index trunc(f*D + 0.5).  Returns (index or -1,
    near-boundary mask).  scale=False plants the mistake of intersecting the
    untransformed ray with the sphere of the minor radius around the centre
    (the ellipsoid scale skipped)."""
    D, H, W = res
    xs = inclusive32(W, VS[0], VE[0]).astype(np.float64)
    ys = inclusive32(H, VS[1], VE[1]).astype(np.float64)
    p_s = np.asarray(p_s, np.float64)
    p_l = np.asarray(p_l, np.float64)
    c = 0.5 * (p_s + p_l)
    r1 = np.sqrt(max((tc / 2) ** 2 - np.sum((p_s - p_l) ** 2) / 4, 0.0))
    hit = -np.ones((H, W), np.int64)
    near = np.zeros((H, W), bool)
    zs, ze = VS[2], VE[2]
    for yi, y in enumerate(ys):
        for xi, x in enumerate(xs):
            def q(f):
                return np.stack([np.full_like(f, x), np.full_like(f, y), zs + f * (ze - zs)], -1)
            if scale:
                def g(f):
                    qq = q(f)
                    return (np.linalg.norm(qq - p_s, axis=-1) + np.linalg.norm(qq - p_l, axis=-1)
                            - tc)
            else:
                def g(f):
                    return np.linalg.norm(q(f) - c, axis=-1) - r1
            r = _roots(g)
            if len(r) < 2:
                continue          # no crossing or a tangent: no hit either way
            f = r[0] if r[0] >= 0 else r[-1]
            u = f * D + 0.5
            b = int(u)            # C truncation toward zero
            near[yi, xi] = abs(u - np.round(u)) < 1e-3 or len(r) != 2
            if 0 <= b < D:
                hit[yi, xi] = b
    return hit, near


def _arellano_decode(vol):
    """Column -> the single z index holding a hit (or -1)."""
    D, H, W = vol.shape
    nz = vol != 0
    cnt = nz.sum(0)
    assert cnt.max() <= 1
    return np.where(cnt > 0, np.argmax(nz, 0), -1)


ARELLANO_CASES = [
    # (scan, laser, scan origin, laser origin, t0, w, off)
    ([-0.3, 0.2, 0.0], [0.4, -0.1, 0.0], None, None, 160, 0.02, 0.0),
    ([-0.5, -0.4, 0.0], [0.6, 0.3, 0.0], [-0.5, -0.4, -0.5], [0.6, 0.3, -0.75], 230, 0.02, 0.1),
    ([0.1, 0.6, 0.0], [-0.2, -0.7, 0.0], [0.5, 0.1, -0.8], [-0.6, -0.3, -0.4], 255, 0.02, 0.0),
]


@pytest.mark.parametrize("case", range(len(ARELLANO_CASES)))
def test_arellano_ellipsoid_and_origins_brute_force(case):
    p_s, p_l, o_s, o_l, t0, w, off = ARELLANO_CASES[case]
    res = (48, 13, 15)
    T = 300
    scan = np.array(p_s, f32).reshape(1, 3, 1)
    laser = np.array(p_l, f32).reshape(1, 3, 1)
    so = None if o_s is None else np.array(o_s, f32).reshape(1, 3, 1)
    lo = None if o_l is None else np.array(o_l, f32).reshape(1, 3, 1)
    tr = np.zeros((1, T, 1, 1))
    tr[0, t0, 0, 0] = 1.0
    vol = O.arellano(tr, scan, laser, so, lo, res, VS, VE, w, off, mode=O.FP32)[0]
    got = _arellano_decode(vol)
    d_so = 0.0 if so is None else float(np.linalg.norm(np.float64(scan[0, :, 0]) - so[0, :, 0]))
    d_lo = 0.0 if lo is None else float(np.linalg.norm(np.float64(laser[0, :, 0]) - lo[0, :, 0]))
    tc = t0 * w + off - d_so - d_lo
    want, near = arellano_hits64(scan[0, :, 0], laser[0, :, 0], tc, res)
    assert np.count_nonzero(want >= 0) > 60
    assert int(((got != want) & ~near).sum()) == 0
    # planted mistakes: ellipsoid scale skipped; laser origin dropped; scan
    # origin dropped; laser point taken as the scan point (a sphere)
    plants = [arellano_hits64(scan[0, :, 0], laser[0, :, 0], tc, res, scale=False)]
    if lo is not None:
        plants.append(arellano_hits64(scan[0, :, 0], laser[0, :, 0], tc + d_lo, res))
    if so is not None:
        plants.append(arellano_hits64(scan[0, :, 0], laser[0, :, 0], tc + d_so, res))
    plants.append(arellano_hits64(scan[0, :, 0], scan[0, :, 0], tc, res))
    for w_hit, w_near in plants:
        assert int(((got != w_hit) & ~(near | w_near)).sum()) > 10
