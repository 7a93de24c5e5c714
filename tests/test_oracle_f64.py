"""Pins the oracle's restatement of the reference's double dispatch
(AT_DISPATCH_FLOATING_TYPES with scalar_t = double; volume_cuda.cu:85,175,
backprojection_cuda.cu:215), which keeps float geometry, a float bin radius
t*w + off and float-narrowing floorf/ceilf:

* R1: against the pure-FP64 flavour (everything double) -- they differ only
  through the float bin radius and the narrowing, far below any mistake in
  the formula -- and the closed form of a column under a scan point.
* A1: adjoint of R1 to double rounding, <R v, g> = <v, A g>; its window form
  differs from the transpose on the top slice only.
* A2: decisions against the exact float64 bin of the float distances.
"""
import numpy as np

from oracle import oracle as O
from tests.cases import ORIGIN, SIZE, rel_l2, wall_grid


def setup(seed=3, X=10, Y=9, Z=14, n=4):
    rng = np.random.default_rng(seed)
    vol = rng.random((X, Y, Z))
    mp = wall_grid(n).astype(np.float64)
    mp[:2] += rng.uniform(-0.05, 0.05, mp[:2].shape)
    return rng, vol, mp


def test_r1_f64_against_pure_fp64_and_closed_form():
    rng, vol, mp = setup()
    T, w = 300, 0.012
    got = O.r1_forward_f64(vol, mp, T, ORIGIN, SIZE, w, 0.0)
    pure = O.r1_forward(vol.astype(np.float32), mp.astype(np.float32), T, ORIGIN, SIZE, w, 0.0,
                        mode=O.PURE64)
    # the float32 volume and points of the pure flavour and the float bin
    # radius of the double dispatch: differences of float rounding only
    e = rel_l2(got, pure)
    assert e < 1e-5, e
    # a planted mistake (s = Z / size_x with a different x size) is far off
    wrong = O.r1_forward_f64(vol, mp, T, ORIGIN, (2.5, 2.0, 2.0), w, 0.0)
    assert rel_l2(wrong, got) > 1e-2
    # closed form: one scan point right above a column's centre; the column
    # alone occupied -> value = lerp(v, z_hat) / tc^4
    X, Y, Z = 6, 6, 10
    v = np.zeros((X, Y, Z))
    v[2, 3] = rng.random(Z)
    cx = ORIGIN[0] + (2 + 0.5) * SIZE[0] / X
    cy = ORIGIN[1] + (3 + 0.5) * SIZE[1] / Y
    m = np.array([[cx], [cy], [0.0]])
    out = O.r1_forward_f64(v, m, 260, ORIGIN, SIZE, 0.01, 0.0)[0]
    want = np.zeros(260)
    for t in range(260):
        tc = float(np.float32(np.float32(t) * np.float32(0.01)))
        zh = (tc - ORIGIN[2]) * Z / SIZE[2] - 0.5
        zi = int(np.floor(np.float32(zh)))
        if tc <= 0 or not (0 <= zi < Z - 1):
            continue
        a = zh - zi
        want[t] = ((1 - a) * v[2, 3, zi] + a * v[2, 3, zi + 1]) / tc ** 4
    assert np.count_nonzero(want) > 100
    assert np.allclose(out, want, rtol=1e-12, atol=0)


def test_a1_f64_is_the_adjoint_and_window_form_differs_on_top_slice():
    rng, vol, mp = setup(4)
    X, Y, Z = vol.shape
    T, w = 300, 0.012
    g = rng.standard_normal((mp.shape[1], T))
    rv = O.r1_forward_f64(vol, mp, T, ORIGIN, SIZE, w, 0.0)
    at = O.a1_transpose_f64(g, mp, X, Y, Z, ORIGIN, SIZE, w, 0.0)
    lhs, rhs = np.sum(rv * g), np.sum(vol * at)
    assert abs(lhs - rhs) < 1e-12 * np.sqrt(np.sum(rv ** 2) * np.sum(g ** 2))
    ar = O.a1_reference_f64(g, mp, X, Y, Z, ORIGIN, SIZE, w, 0.0)
    assert rel_l2(ar[:, :, :-1], at[:, :, :-1]) < 1e-3
    assert rel_l2(ar[:, :, -1], at[:, :, -1]) > 1e-2
    # the double window form against the float one (same windows but for
    # rounding): close
    ar32 = O.a1_reference(g, mp.astype(np.float32), X, Y, Z, ORIGIN, SIZE, w, 0.0,
                          mode=O.PINNED64)
    assert rel_l2(ar, ar32) < 1e-4


def test_velten_f64_bins():
    """Bin-coded transient (t + 1 in bin t): the double dispatch takes
    floor(float((2 d_so + 2 d_sv - off) / w + 1/2)) of float distances; the
    exact float64 value of that expression agrees except within rounding of
    a bin edge."""
    T, S = 700, 1
    res = (12, 10, 11)
    vs, ve = (-1.0, -1.0, 0.5), (1.0, 1.0, 2.5)
    w, off = 0.01, 0.05
    scan = np.array([[[0.1, -0.2, 0.0]]])
    so = np.array([[[0.3, 0.1, -0.7]]])
    tr = (np.arange(T) + 1.0)[None, :, None]
    vol = O.velten_confocal_f64(tr, scan, so, res, vs, ve, w, off)[0]
    got = np.rint(vol).astype(np.int64) - 1
    i = np.arange(res[2], dtype=np.float32)
    qx = (np.float32(vs[0]) + (i / np.float32(res[2] - 1)).astype(np.float32) *
          (np.float32(ve[0]) - np.float32(vs[0]))).astype(np.float64)[None, None, :]
    j = np.arange(res[1], dtype=np.float32)
    qy = (np.float32(vs[1]) + (j / np.float32(res[1] - 1)).astype(np.float32) *
          (np.float32(ve[1]) - np.float32(vs[1]))).astype(np.float64)[None, :, None]
    k = np.arange(res[0], dtype=np.float32)
    qz = (np.float32(vs[2]) + (k / np.float32(res[0] - 1)).astype(np.float32) *
          (np.float32(ve[2]) - np.float32(vs[2]))).astype(np.float64)[:, None, None]
    p = scan[0, 0]
    dsv = np.sqrt((p[0] - qx) ** 2 + (p[1] - qy) ** 2 + (p[2] - qz) ** 2)
    dso = np.sqrt(np.sum((p - so[0, 0]) ** 2))
    u = (2 * dso + 2 * dsv - off) / w + 0.5
    want = np.floor(u).astype(np.int64)
    want = np.where((want >= 0) & (want < T), want, -1)
    near = np.abs(u - np.round(u)) < 1e-4
    assert np.count_nonzero(want >= 0) > 200
    assert int(((got != want) & ~near).sum()) == 0
    # planted: the origin term counted once
    wrong = np.floor((dso + 2 * dsv - off) / w + 0.5).astype(np.int64)
    assert int(((got != wrong) & ~near).sum()) > 100
