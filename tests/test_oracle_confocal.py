"""Pins the CPU oracle's confocal render (R1), adjoint (A1) and
measurement-point gradient (A6) against the mathematics, independently of the
oracle's own code: closed forms, an independent vectorised FP64 restatement,
a scalar float32 restatement of the decision formula, linearity, adjointness
and finite differences.  Reference: src/render/volume_cuda.cu:11-265.
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests.cases import ORIGIN, SIZE, rel_l2, sparse_albedo, wall_grid

f32 = np.float32


def cell_centre32(i, o, size, n):
    # volume_cuda.cu:35 -- product, quotient, sum, each rounded to float32
    a = f32(f32(f32(i) + f32(0.5)) * f32(size))
    a = f32(a / f32(n))
    return f32(f32(o) + a)


def r1_numpy64(vol, mp, T, origin, size, w, off):
    """Independent vectorised FP64 restatement of R1 (all decisions in FP64)."""
    X, Y, Z = vol.shape
    vol = vol.astype(np.float64)
    t = np.arange(T, dtype=np.float64) * float(w) + float(off)
    x = origin[0] + (np.arange(X) + 0.5) * size[0] / X
    y = origin[1] + (np.arange(Y) + 0.5) * size[1] / Y
    s = Z / size[2]
    out = np.zeros((mp.shape[1], T))
    for i in range(mp.shape[1]):
        wx, wy, wz = (float(v) for v in mp[:, i])
        d2 = (t[:, None, None] ** 2 - (wx - x)[None, :, None] ** 2
              - (wy - y)[None, None, :] ** 2)
        depth = np.sqrt(np.maximum(d2, 0.0))
        zh = (wz + depth - origin[2]) * s - 0.5
        fl = np.floor(zh)
        ok = (d2 > 0) & (fl >= 0) & (fl + 1 < Z)
        zi = np.clip(fl, 0, Z - 2).astype(np.int64)
        a = zh - fl
        xi = np.arange(X)[None, :, None]
        yi = np.arange(Y)[None, None, :]
        v0 = vol[xi, yi, zi]
        v1 = vol[xi, yi, zi + 1]
        f2 = (depth / t[:, None, None] ** 2) ** 4
        out[i] = np.where(ok, ((1 - a) * v0 + a * v1) * f2, 0.0).sum(axis=(1, 2))
    return out


def test_closed_form_column_under_scan_point():
    """rho = 0: depth = tc, so transient[t] = lerp(v, z_hat) / tc^4 with
    z_hat = (tc - oz) * Z / size_z - 0.5 (one occupied column, x != y index)."""
    X, Y, Z, T = 4, 5, 16, 96
    xi, yi = 1, 3
    vol = np.zeros((X, Y, Z), np.float32)
    rng = np.random.default_rng(1)
    vol[xi, yi] = rng.random(Z).astype(np.float32) + 0.5
    wx = cell_centre32(xi, ORIGIN[0], SIZE[0], X)
    wy = cell_centre32(yi, ORIGIN[1], SIZE[1], Y)
    mp = np.array([[wx], [wy], [0.0]], np.float32)
    w, off = 0.03, 0.01
    got = O.r1_forward(vol, mp, T, ORIGIN, SIZE, w, off, mode=O.PINNED64)[0]
    tc = np.arange(T) * np.float64(f32(w)) + np.float64(f32(off))
    zh = (tc - ORIGIN[2]) * (Z / SIZE[2]) - 0.5
    fl = np.floor(zh)
    ok = (fl >= 0) & (fl + 1 < Z)
    zi = np.clip(fl, 0, Z - 2).astype(int)
    a = zh - fl
    col = vol[xi, yi].astype(np.float64)
    want = np.where(ok, ((1 - a) * col[zi] + a * col[zi + 1]) / tc ** 4, 0.0)
    # Only the occupied column contributes.  The FP32 tc of the oracle
    # differs from the FP64 tc above by a few ulps.
    assert np.count_nonzero(want) > 40
    assert rel_l2(got, want) < 1e-5
    got32 = O.r1_forward(vol, mp, T, ORIGIN, SIZE, w, off, mode=O.FP32)[0]
    assert rel_l2(got32, want) < 1e-5


def test_pure_fp64_matches_independent_restatement():
    rng = np.random.default_rng(2)
    X, Y, Z, T = 5, 6, 7, 60
    vol = rng.random((X, Y, Z)).astype(np.float32)
    mp = np.stack([rng.uniform(-1, 1, 3), rng.uniform(-1, 1, 3),
                   rng.uniform(-0.2, 0.2, 3)]).astype(np.float32)
    w, off = 0.05, 0.02
    got = O.r1_forward(vol, mp, T, ORIGIN, SIZE, w, off, mode=O.PURE64)
    want = r1_numpy64(vol, mp, T, ORIGIN, SIZE, np.float64(f32(w)), np.float64(f32(off)))
    assert rel_l2(got, want) < 1e-12
    # FP32-pinned decisions agree with FP64 decisions except at rare flips.
    pinned = O.r1_forward(vol, mp, T, ORIGIN, SIZE, w, off, mode=O.PINNED64)
    assert rel_l2(pinned, want) < 1e-3


def scalar_decision32(i, t, xi, yi, mp, X, Y, Z, w, off):
    """The reference's FP32 decision sequence with numpy float32 scalars
    (volume_cuda.cu:23,35,38,40-51)."""
    tc = f32(f32(f32(t) * f32(w)) + f32(off))
    x = cell_centre32(xi, ORIGIN[0], SIZE[0], X)
    y = cell_centre32(yi, ORIGIN[1], SIZE[1], Y)
    wx, wy, wz = (f32(v) for v in mp[:, i])
    dx = f32(wx - x)
    dy = f32(wy - y)
    d2 = f32(f32(tc * tc) - f32(dx * dx))
    d2 = f32(d2 - f32(dy * dy))
    if not d2 > 0:
        return -1000000, f32(0), d2
    depth = f32(np.sqrt(d2))
    z = f32(wz + depth)
    s = f32(f32(Z) / f32(SIZE[2]))
    zh = f32(f32(f32(z - f32(ORIGIN[2])) * s) - f32(0.5))
    fl = f32(np.floor(zh))
    return int(fl), f32(zh - fl), d2


def test_fp32_decisions_match_scalar_restatement():
    rng = np.random.default_rng(3)
    X = Y = 32
    Z, T = 32, 512
    mp = wall_grid(8)
    n = 3000
    idx = np.stack([rng.integers(0, mp.shape[1], n), rng.integers(0, T, n),
                    rng.integers(0, X, n), rng.integers(0, Y, n)], axis=1)
    w, off = 0.005, 0.0
    zi, al, d2 = O.r1_decisions(idx, mp, X, Y, Z, ORIGIN, SIZE, w, off)
    for k in range(n):
        zs, as_, ds = scalar_decision32(*idx[k], mp, X, Y, Z, w, off)
        assert zi[k] == zs and d2[k] == ds, k
        assert al[k] == as_, k


def test_linearity_and_adjointness_pinned():
    rng = np.random.default_rng(4)
    X, Y, Z, T = 6, 5, 8, 80
    mp = np.stack([rng.uniform(-1, 1, 4), rng.uniform(-1, 1, 4),
                   np.zeros(4)]).astype(np.float32)
    w, off = 0.04, 0.0
    v1 = rng.random((X, Y, Z)).astype(np.float32)
    v2 = rng.random((X, Y, Z)).astype(np.float32)
    r = lambda v: O.r1_forward(v, mp, T, ORIGIN, SIZE, w, off, mode=O.PINNED64)
    a, b = 0.5, 0.25
    mix = (a * v1 + b * v2).astype(np.float32)  # one float32 rounding per voxel
    assert rel_l2(r(mix), a * r(v1) + b * r(v2)) < 1e-7
    assert rel_l2(r(2 * v1), 2 * r(v1)) < 1e-13  # exact scaling
    g = rng.standard_normal((4, T))
    at = O.a1_transpose(g, mp, X, Y, Z, ORIGIN, SIZE, w, off, mode=O.PINNED64)
    lhs = float(np.sum(r(v1) * g))
    rhs = float(np.sum(v1.astype(np.float64) * at))
    assert abs(lhs - rhs) < 1e-11 * (abs(lhs) + 1e-30)
    # FP32 flavour: same decisions, FP32 arithmetic.
    at32 = O.a1_transpose(g, mp, X, Y, Z, ORIGIN, SIZE, w, off, mode=O.FP32)
    assert rel_l2(at32, at) < 1e-5


def a1_numpy64(G, mp, X, Y, Z, origin, size, w, off, swap=False):
    """Independent vectorised FP64 transpose of r1_numpy64.  swap=True plants
    a mistake (alpha and 1 - alpha exchanged)."""
    t = np.arange(G.shape[1], dtype=np.float64) * float(w) + float(off)
    x = origin[0] + (np.arange(X) + 0.5) * size[0] / X
    y = origin[1] + (np.arange(Y) + 0.5) * size[1] / Y
    s = Z / size[2]
    out = np.zeros((X, Y, Z))
    xi = np.broadcast_to(np.arange(X)[None, :, None], (len(t), X, Y))
    yi = np.broadcast_to(np.arange(Y)[None, None, :], (len(t), X, Y))
    for i in range(mp.shape[1]):
        wx, wy, wz = (float(v) for v in mp[:, i])
        d2 = (t[:, None, None] ** 2 - (wx - x)[None, :, None] ** 2
              - (wy - y)[None, None, :] ** 2)
        depth = np.sqrt(np.maximum(d2, 0.0))
        zh = (wz + depth - origin[2]) * s - 0.5
        fl = np.floor(zh)
        ok = (d2 > 0) & (fl >= 0) & (fl + 1 < Z)
        a = zh - fl
        if swap:
            a = 1 - a
        f2g = (depth / t[:, None, None] ** 2) ** 4 * G[i][:, None, None]
        zi = fl[ok].astype(np.int64)
        np.add.at(out, (xi[ok], yi[ok], zi), ((1 - a) * f2g)[ok])
        np.add.at(out, (xi[ok], yi[ok], zi + 1), (a * f2g)[ok])
    return out


def test_adjoint_independent_restatement_and_planted_mistake():
    """The oracle's pure-FP64 adjoint equals an independent restatement; the
    adjointness identity holds for it and fails for a planted mistake
    (alpha and 1 - alpha swapped in the adjoint)."""
    rng = np.random.default_rng(5)
    X, Y, Z, T = 5, 4, 9, 70
    mp = np.stack([rng.uniform(-1, 1, 2), rng.uniform(-1, 1, 2),
                   np.zeros(2)]).astype(np.float32)
    w, off = 0.045, 0.0
    v = rng.random((X, Y, Z)).astype(np.float32)
    g = rng.standard_normal((2, T))
    wf, of = np.float64(np.float32(w)), np.float64(np.float32(off))
    ind = a1_numpy64(g, mp, X, Y, Z, ORIGIN, SIZE, wf, of)
    ora = O.a1_transpose(g, mp, X, Y, Z, ORIGIN, SIZE, w, off, mode=O.PURE64)
    assert rel_l2(ora, ind) < 1e-12
    lhs = float(np.sum(r1_numpy64(v, mp, T, ORIGIN, SIZE, wf, of) * g))
    good = float(np.sum(v.astype(np.float64) * ind))
    bad = float(np.sum(v.astype(np.float64) *
                       a1_numpy64(g, mp, X, Y, Z, ORIGIN, SIZE, wf, of, swap=True)))
    assert abs(lhs - good) < 1e-11 * abs(lhs)
    assert abs(lhs - bad) > 1e-3 * abs(lhs)


def test_albedo_gradient_matches_finite_differences():
    """L = ||R v - y||^2 is quadratic in v with FP32-pinned decisions, so
    central differences in FP64 equal the analytic 2 R^T (R v - y)."""
    rng = np.random.default_rng(6)
    X, Y, Z, T = 6, 6, 8, 64
    mp = np.stack([rng.uniform(-1, 1, 3), rng.uniform(-1, 1, 3),
                   np.zeros(3)]).astype(np.float32)
    w, off = 0.045, 0.0
    v = rng.random((X, Y, Z)).astype(np.float32)
    y = rng.random((3, T))
    R = lambda vv: O.r1_forward(vv, mp, T, ORIGIN, SIZE, w, off, mode=O.PINNED64)
    grad = 2 * O.a1_transpose(R(v) - y, mp, X, Y, Z, ORIGIN, SIZE, w, off,
                              mode=O.PINNED64)
    # R is linear: evaluate it on a float64 basis by linearity (a*e_k exactly).
    base = R(v) - y
    for _ in range(12):
        k = tuple(rng.integers(0, s) for s in (X, Y, Z))
        e = np.zeros((X, Y, Z), np.float32)
        e[k] = 1.0
        Re = R(e)
        h = 0.1  # the loss is quadratic: central differences are exact for any h
        lp = np.sum((base + h * Re) ** 2)
        lm = np.sum((base - h * Re) ** 2)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - grad[k]) <= 1e-9 * (abs(grad[k]) + 1e-6), (k, fd, grad[k])


def test_reference_window_adjoint_differs_only_on_top_slice():
    """The reference's window-form A1 (volume_cuda.cu:95-160) agrees with the
    exact transpose away from the top z slice, where it adds samples that R1
    never uses (zi = Z-1)."""
    rng = np.random.default_rng(7)
    X = Y = 8
    Z, T = 16, 256
    mp = wall_grid(4)
    w, off = 0.01, 0.0
    g = rng.random((mp.shape[1], T))
    at = O.a1_transpose(g, mp, X, Y, Z, ORIGIN, SIZE, w, off, mode=O.FP32)
    ar = O.a1_reference(g, mp, X, Y, Z, ORIGIN, SIZE, w, off, mode=O.FP32)
    assert rel_l2(ar[:, :, :-1], at[:, :, :-1]) < 1e-4
    assert rel_l2(ar[:, :, -1], at[:, :, -1]) > 1e-2


def test_measurement_point_gradient_matches_finite_differences():
    """A6 (volume_cuda.cu:188-265) against central differences of the pure
    FP64 render; the volume is zero on the end slices so the render is
    continuous in the wall point (piecewise smooth; kinks bound the FD error)."""
    rng = np.random.default_rng(8)
    X, Y, Z, T = 8, 8, 12, 200
    vol = rng.random((X, Y, Z)).astype(np.float32)
    vol[:, :, 0] = 0
    vol[:, :, -1] = 0
    vol[:, :, 1] = 0
    vol[:, :, -2] = 0
    mp = np.array([[0.13], [-0.27], [0.05]], np.float32)
    w, off = 0.012, 0.0
    G = rng.random((1, T))
    ana = O.a6(vol, mp, G, ORIGIN, SIZE, w, off, mode=O.PINNED64)[:, 0]
    h = 2e-4
    for c in range(3):
        mpp = mp.astype(np.float64).copy()
        mpm = mp.astype(np.float64).copy()
        mpp[c] += h
        mpm[c] -= h
        lp = np.sum(r1_numpy64(vol, mpp, T, ORIGIN, SIZE, w, off) * G)
        lm = np.sum(r1_numpy64(vol, mpm, T, ORIGIN, SIZE, w, off) * G)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - ana[c]) < 2e-2 * np.abs(ana).max(), (c, fd, ana[c])


def test_column_restricted_render_equals_full_render():
    """r1_forward(cols=...) -- used for full-size finite differences -- is the
    render of the volume with every other column zero, in every mode."""
    rng = np.random.default_rng(29)
    X, Y, Z, T = 9, 7, 12, 90
    mp = np.stack([rng.uniform(-1, 1, 11), rng.uniform(-1, 1, 11),
                   np.zeros(11)]).astype(np.float32)
    cols = np.array([[2, 3], [4, 6], [8, 0]])  # index order
    vol = np.zeros((X, Y, Z), np.float32)
    for xi, yi in cols:
        vol[xi, yi] = rng.random(Z)
    for mode in (O.FP32, O.PINNED64, O.PURE64):
        full = O.r1_forward(vol, mp, T, ORIGIN, SIZE, 0.03, 0.01, mode=mode)
        sub = O.r1_forward(vol, mp, T, ORIGIN, SIZE, 0.03, 0.01, mode=mode, cols=cols)
        rows = O.r1_forward(vol, mp, T, ORIGIN, SIZE, 0.03, 0.01, mode=mode, i0=3, i1=7)
        assert np.count_nonzero(full) > 0
        assert np.array_equal(full, sub) and np.array_equal(full[3:7], rows)


def test_decisions_under_nvcc_default_contraction():
    """The reference's .cu files build with nvcc's default --fmad=true, which
    fuses tc = t*w + off (volume_cuda.cu:23) and z_hat = (z - o)*s - 0.5 (:46)
    into FMAs.  This build pins the uncontracted reading.  With off = 0 the
    fused tc equals the unfused one, and z_hat's single rounding differs from
    fl(fl((z-o)s) - 0.5) only where the product sits at a binade edge, so the
    decisions agree: every sample of 64^3 / 32^2 / 512 (1.1e9 active), 64 scan
    rows of 128^3 / 64^2 / 1024 and 16 rows of 192^3 / 128^2 / 2048 (s = 96,
    not a power of two).  A nonzero offset does flip a few floors (about 1e-6
    of the samples), which shows the count is live."""
    from oracle import oracle as O
    from tests.cases import ORIGIN, SIZE, wall_grid
    cases = [(32, 512, 64, 0.005, 0.0, None), (64, 1024, 128, 0.0025, 0.0, 64),
             (128, 2048, 192, 2.56 / 2048, 0.0, 16)]
    for n, T, R, w, off, rows in cases:
        act, aflip, zflip = O.r1_contraction_flips(wall_grid(n), T, R, R, R, ORIGIN, SIZE, w,
                                                   off, 0, rows)
        print(f"{R}^3/{n}^2/{T}: {act} active samples, {aflip} activity and {zflip} floor flips")
        assert act > 4e8 and aflip == 0 and zflip == 0
    act, aflip, zflip = O.r1_contraction_flips(wall_grid(64), 1024, 128, 128, 128, ORIGIN, SIZE,
                                               0.0025, 0.013, 0, 64)
    print(f"offset 0.013: {act} active samples, {aflip} activity and {zflip} floor flips")
    assert 0 < zflip < 1e-5 * act
