"""GPU parity of the confocal render R1 and its exact adjoint A1 against the
CPU oracle, through the C ABI (via totrilib_amd.totri_cpp).

Bar: bin/voxel decisions bit-exact; sums within 1e-5 relative L2 of the
FP64-pinned oracle (the tolerance of the benchmark statement).
"""
import numpy as np
import pytest
import torch

from oracle import oracle as O
from tests.cases import ORIGIN, SIZE, rel_l2, sparse_albedo, wall_grid

pytestmark = pytest.mark.gpu

TOL = 1e-5


def dev():
    return torch.device("cuda:0")


def gpu_r1(vol, mp, T, w, off, dtype=torch.float32, origin=ORIGIN, size=SIZE):
    from totrilib_amd.totri_cpp.render import volume_confocal_forward
    v = torch.from_numpy(np.ascontiguousarray(vol)).to(dev())
    m = torch.from_numpy(np.ascontiguousarray(mp)).to(dev())
    out = torch.full((m.shape[1], T), float("nan"), device=dev(), dtype=dtype)
    volume_confocal_forward(v, m, out, origin, size, w, off)
    torch.cuda.synchronize()
    return out.float().cpu().numpy().astype(np.float64)


def gpu_a1(G, mp, res, w, off, origin=ORIGIN, size=SIZE):
    from totrilib_amd.totri_cpp.render import volume_confocal_grad_volume
    g = torch.as_tensor(G).to(dev())
    m = torch.from_numpy(np.ascontiguousarray(mp)).to(dev())
    out = torch.full(res, float("nan"), device=dev(), dtype=torch.float32)
    volume_confocal_grad_volume(out, m, g, origin, size, w, off)
    torch.cuda.synchronize()
    return out.cpu().numpy().astype(np.float64)


def mirror_info(mp):
    """[quads ok, n quads, x-pairs ok, n pairs] from the device detection."""
    from totrilib_amd import _lib
    lib = _lib.load()
    m = torch.from_numpy(np.ascontiguousarray(mp)).to(dev())
    N = m.shape[1]
    ws = torch.empty(lib.totri_debug_mirror_workspace(N), dtype=torch.uint8, device=dev())
    info = torch.full((4,), -7, dtype=torch.int32, device=dev())
    _lib.check(lib.totri_debug_mirror_info(m.data_ptr(), N, ws.data_ptr(), ws.numel(),
                                           info.data_ptr(),
                                           torch.cuda.current_stream().cuda_stream),
               "debug_mirror_info")
    torch.cuda.synchronize()
    return info.cpu().tolist()


def gpu_decisions(idx, mp, X, Y, Z, T, w, off):
    from totrilib_amd import _lib
    lib = _lib.load()
    m = torch.from_numpy(np.ascontiguousarray(mp)).to(dev())
    ix = torch.from_numpy(np.ascontiguousarray(idx.astype(np.int32))).to(dev())
    n = ix.shape[0]
    zi = torch.empty(n, dtype=torch.int32, device=dev())
    al = torch.empty(n, dtype=torch.float32, device=dev())
    d2 = torch.empty(n, dtype=torch.float32, device=dev())
    rc = lib.totri_debug_confocal_samples(
        m.data_ptr(), m.shape[1], ix.data_ptr(), n, X, Y, Z, T, _lib.f3(ORIGIN),
        _lib.f3(SIZE), w, off, zi.data_ptr(), al.data_ptr(), d2.data_ptr(),
        torch.cuda.current_stream().cuda_stream)
    _lib.check(rc, "debug_confocal_samples")
    torch.cuda.synchronize()
    return zi.cpu().numpy(), al.cpu().numpy(), d2.cpu().numpy()


@pytest.mark.parametrize("n_side,T,res,w", [(64, 1024, 128, 0.0025), (32, 512, 64, 0.005),
                                            (256, 2048, 256, 0.00125)])
def test_sample_decisions_bit_exact(n_side, T, res, w):
    """zi = floor(z_hat) and depth^2 of 2e6 random samples, half of them drawn
    next to voxel boundaries, equal the oracle's FP32 decisions bit for bit."""
    rng = np.random.default_rng(11)
    mp = wall_grid(n_side)
    X = Y = Z = res
    n = 1_000_000
    idx = np.stack([rng.integers(0, mp.shape[1], n), rng.integers(0, T, n),
                    rng.integers(0, X, n), rng.integers(0, Y, n)], axis=1)
    # Boundary-biased half: keep the samples whose FP32 z_hat is within 1e-3
    # of an integer, out of a larger random pool.
    pool = np.stack([rng.integers(0, mp.shape[1], 8 * n), rng.integers(0, T, 8 * n),
                     rng.integers(0, X, 8 * n), rng.integers(0, Y, 8 * n)], axis=1)
    pz, pa, _ = O.r1_decisions(pool, mp, X, Y, Z, ORIGIN, SIZE, w, 0.0)
    near = (pz > -1000000) & ((pa < 1e-3) | (pa > 1 - 1e-3))
    idx = np.concatenate([idx, pool[near][:n]])
    zo, ao, do = O.r1_decisions(idx, mp, X, Y, Z, ORIGIN, SIZE, w, 0.0)
    zg, ag, dg = gpu_decisions(idx, mp, X, Y, Z, T, w, 0.0)
    print(f"samples={len(idx)} near-boundary={near.sum()}")
    assert np.array_equal(dg.view(np.int32), do.view(np.int32))
    assert np.array_equal(zg, zo)
    act = zo > -1000000
    # alpha from the fused fast path differs from the reference by a few
    # ulps of z_hat at most.
    assert np.max(np.abs(ag[act] - ao[act])) < 64 * np.spacing(np.float32(res))


def test_r1_small_matches_oracle():
    rng = np.random.default_rng(12)
    X, Y, Z, T = 24, 20, 28, 300
    vol = rng.random((X, Y, Z)).astype(np.float32)
    mp = np.stack([rng.uniform(-1, 1, 37), rng.uniform(-1, 1, 37),
                   rng.uniform(-0.1, 0.1, 37)]).astype(np.float32)
    w, off = 0.009, 0.003
    got = gpu_r1(vol, mp, T, w, off)
    ref32 = O.r1_forward(vol, mp, T, ORIGIN, SIZE, w, off, mode=O.FP32)
    ref64 = O.r1_forward(vol, mp, T, ORIGIN, SIZE, w, off, mode=O.PINNED64)
    e32, e64 = rel_l2(got, ref32), rel_l2(got, ref64)
    print(f"rel-L2 vs FP32 oracle {e32:.2e}, vs FP64 pinned {e64:.2e}")
    assert e64 < TOL and e32 < TOL


def test_r1_cfg1_rows_match_oracle():
    """64^3 volume, 32x32 scan, 512 bins of 0.005 m (one-way): 96 seeded rows."""
    mp = wall_grid(32)
    vol = sparse_albedo((64, 64, 64), 0.05, 0)
    T, w = 512, 0.005
    got = gpu_r1(vol, mp, T, w, 0.0)
    rows = np.random.default_rng(13).choice(mp.shape[1], 96, replace=False)
    ref = np.stack([O.r1_forward(vol, mp, T, ORIGIN, SIZE, w, 0.0, mode=O.PINNED64,
                                 i0=int(r), i1=int(r) + 1)[0] for r in rows])
    e = rel_l2(got[rows], ref)
    print(f"cfg1 rows rel-L2 vs FP64 pinned {e:.2e}")
    assert e < TOL


def test_r1_bf16_fp16_storage():
    mp = wall_grid(16)
    vol = sparse_albedo((32, 32, 32), 0.05, 1)
    T, w = 256, 0.01
    ref = O.r1_forward(vol, mp, T, ORIGIN, SIZE, w, 0.0, mode=O.PINNED64)
    for dt, tol in ((torch.bfloat16, 8e-3), (torch.float16, 1e-3)):
        got = gpu_r1(vol, mp, T, w, 0.0, dtype=dt)
        e = rel_l2(got, ref)
        print(f"{dt}: rel-L2 {e:.2e}")
        assert e < tol


def test_a1_small_matches_oracle():
    rng = np.random.default_rng(14)
    X, Y, Z, T = 20, 23, 70, 260
    mp = np.stack([rng.uniform(-1, 1, 29), rng.uniform(-1, 1, 29),
                   rng.uniform(-0.1, 0.1, 29)]).astype(np.float32)
    w, off = 0.01, 0.002
    G = rng.standard_normal((29, T)).astype(np.float32)
    got = gpu_a1(G, mp, (X, Y, Z), w, off)
    ref32 = O.a1_transpose(G, mp, X, Y, Z, ORIGIN, SIZE, w, off, mode=O.FP32)
    ref64 = O.a1_transpose(G, mp, X, Y, Z, ORIGIN, SIZE, w, off, mode=O.PINNED64)
    e32, e64 = rel_l2(got, ref32), rel_l2(got, ref64)
    print(f"rel-L2 vs FP32 oracle {e32:.2e}, vs FP64 pinned {e64:.2e}")
    assert e64 < TOL and e32 < TOL


def test_a1_cfg1_columns_match_oracle():
    mp = wall_grid(32)
    T, w = 512, 0.005
    G = np.random.default_rng(15).random((mp.shape[1], T)).astype(np.float32)
    got = gpu_a1(G, mp, (64, 64, 64), w, 0.0)
    rng = np.random.default_rng(16)
    cols = np.stack([rng.integers(0, 64, 48), rng.integers(0, 64, 48)], axis=1)
    ref = O.a1_transpose(G, mp, 64, 64, 64, ORIGIN, SIZE, w, 0.0, mode=O.PINNED64,
                         cols=cols)
    e = rel_l2(got[cols[:, 0], cols[:, 1]], ref[cols[:, 0], cols[:, 1]])
    print(f"cfg1 columns rel-L2 vs FP64 pinned {e:.2e}")
    assert e < TOL


def test_adjointness_on_gpu():
    """<R v, g> = <v, A g> for the GPU pair (FP32 rounding only)."""
    mp = wall_grid(24)
    vol = sparse_albedo((48, 40, 56), 0.1, 2)
    T, w = 640, 0.004
    g = np.random.default_rng(17).standard_normal((mp.shape[1], T)).astype(np.float32)
    rv = gpu_r1(vol, mp, T, w, 0.0)
    ag = gpu_a1(g, mp, vol.shape, w, 0.0)
    lhs = float(np.sum(rv * g))
    rhs = float(np.sum(vol.astype(np.float64) * ag))
    assert abs(lhs - rhs) < 1e-5 * np.sqrt(np.sum(rv ** 2) * np.sum(g.astype(np.float64) ** 2))


@pytest.mark.parametrize("z", [0, 1, 62, 63])
def test_support_of_single_voxel_render(z):
    """A single occupied voxel at the ends of the z range: the transient is
    nonzero exactly where the FP32 decisions put a sample into that voxel."""
    mp = wall_grid(32)
    X = Y = Z = 64
    vol = np.zeros((X, Y, Z), np.float32)
    vol[21, 40, z] = 1.0
    T, w = 512, 0.005
    got = gpu_r1(vol, mp, T, w, 0.0)
    ref = O.r1_forward(vol, mp, T, ORIGIN, SIZE, w, 0.0, mode=O.FP32)
    assert np.array_equal(got != 0, ref != 0)
    assert rel_l2(got, ref) < TOL


@pytest.mark.parametrize("z", [0, 1, 62, 63])
def test_support_of_single_voxel_adjoint(z):
    mp = wall_grid(32)
    X = Y = Z = 64
    T, w = 512, 0.005
    rng = np.random.default_rng(z)
    G = np.zeros((mp.shape[1], T), np.float32)
    G[rng.integers(0, mp.shape[1], 40), rng.integers(0, T, 40)] = 1.0
    got = gpu_a1(G, mp, (X, Y, Z), w, 0.0)
    cols = np.array([[xi, yi] for xi in range(0, 64, 3) for yi in range(1, 64, 5)])
    ref = O.a1_transpose(G, mp, X, Y, Z, ORIGIN, SIZE, w, 0.0, mode=O.FP32, cols=cols)
    gs = got[cols[:, 0], cols[:, 1]]
    rs = ref[cols[:, 0], cols[:, 1]]
    assert np.array_equal(gs != 0, rs != 0)
    assert rel_l2(gs, rs) < TOL


def test_noncontiguous_output_and_errors():
    from totrilib_amd.totri_cpp.render import volume_confocal_forward
    mp = torch.from_numpy(wall_grid(4)).to(dev())
    vol = torch.rand(8, 8, 8, device=dev())
    out = torch.zeros(64, 16, device=dev()).T  # [16, 64] view, transposed
    base = torch.zeros(16, 64, device=dev())
    volume_confocal_forward(vol, mp, base, ORIGIN, SIZE, 0.05, 0.0)
    volume_confocal_forward(vol, mp, out, ORIGIN, SIZE, 0.05, 0.0)
    torch.cuda.synchronize()
    assert torch.equal(out, base)
    with pytest.raises(ValueError):
        volume_confocal_forward(vol, mp, torch.zeros(15, 64, device=dev()), ORIGIN, SIZE,
                                0.05, 0.0)
    with pytest.raises(NotImplementedError):
        volume_confocal_forward(vol.cpu(), mp, base, ORIGIN, SIZE, 0.05, 0.0)


def test_degenerate_shapes():
    from totrilib_amd.totri_cpp.render import (volume_confocal_forward,
                                               volume_confocal_grad_volume)
    mp = torch.from_numpy(wall_grid(4)).to(dev())
    out = torch.full((16, 32), 7.0, device=dev())
    volume_confocal_forward(torch.rand(4, 4, 1, device=dev()), mp, out, ORIGIN, SIZE,
                            0.05, 0.0)
    torch.cuda.synchronize()
    assert torch.count_nonzero(out) == 0  # Z = 1: no sample has zi + 1 < Z
    vol = torch.full((4, 4, 6), 7.0, device=dev())
    volume_confocal_grad_volume(vol, mp[:, :0], torch.zeros(0, 32, device=dev()), ORIGIN,
                                SIZE, 0.05, 0.0)
    torch.cuda.synchronize()
    assert torch.count_nonzero(vol) == 0  # no scan points: zero gradient


@pytest.mark.parametrize("shape,n_scan,scan", [
    ((5, 7, 2), 9, "random"),      # Z = 2: one slab, no boundary row
    ((9, 6, 3), 1, "random"),      # one scan point
    ((7, 9, 9), 5, "random"),      # Z = 9: a one-row top slab takes the boundary row
    ((6, 8, 17), 16, "random"),    # odd slab count
    ((10, 8, 17), 16, "quads"),    # the xy-mirror walk with Z = 17
    ((12, 10, 8), 64, "quads"),    # Z a multiple of the slab height: no boundary rows
    ((10, 9, 17), 16, "quads"),    # odd Y: the x-pair walk with Z = 17
    ((6, 7, 8), 64, "quads"),      # the x-pair walk, one tile row, no boundary rows
])
def test_a1_row_walks_edge_shapes(shape, n_scan, scan):
    """The row walks' slabs (8 voxel rows), boundary rows and scan ranges on
    small, odd and slab-sized z extents, few scan points, against the
    oracle's exact transpose."""
    rng = np.random.default_rng(sum(shape) + n_scan)
    X, Y, Z = shape
    T, w = 300, 0.01
    if scan == "quads":
        mp = wall_grid(int(round(n_scan ** 0.5)))
    else:
        mp = np.stack([rng.uniform(-1, 1, n_scan), rng.uniform(-1, 1, n_scan),
                       np.zeros(n_scan)]).astype(np.float32)
    info = mirror_info(mp)
    assert info[0] == (1 if scan == "quads" else 0)
    G = rng.standard_normal((mp.shape[1], T)).astype(np.float32)
    got = gpu_a1(G, mp, shape, w, 0.0)
    ref = O.a1_transpose(G, mp, X, Y, Z, ORIGIN, SIZE, w, 0.0, mode=O.PINNED64)
    e = rel_l2(got, ref)
    print(f"{shape} n={n_scan} {scan}: rel-L2 vs FP64 pinned {e:.2e}")
    assert np.all(np.isfinite(got)) and e < TOL


@pytest.mark.parametrize("scan", ["quads", "pairs", "random"])
@pytest.mark.parametrize("dt", [-6, -2, 0, 2, 6])
def test_a1_row_walks_near_last_bin(scan, dt):
    """Bin counts around the one where the farthest voxel centre's radius
    meets the last bin: waves whose bin ranges end at least two bins before
    the last walk without a per-step stop test, the others stop at their
    range end; both against the oracle's exact transpose."""
    X, Y, Z = 16, 15 if scan == "pairs" else 16, 24  # odd Y: x-pairs only
    w = 0.01
    rng = np.random.default_rng(40 + dt + (0 if scan == "quads" else 100))
    if scan in ("quads", "pairs"):
        mp = wall_grid(8)
    else:
        mp = np.stack([rng.uniform(-1, 1, 40), rng.uniform(-1, 1, 40),
                       np.zeros(40)]).astype(np.float32)
    cx = ORIGIN[0] + (np.arange(X) + 0.5) * SIZE[0] / X
    cy = ORIGIN[1] + (np.arange(Y) + 0.5) * SIZE[1] / Y
    cz = ORIGIN[2] + (np.arange(Z) + 0.5) * SIZE[2] / Z
    far = max(np.sqrt((np.abs(mp[0, i] - cx).max()) ** 2 + (np.abs(mp[1, i] - cy).max()) ** 2 +
                      (cz.max() - mp[2, i]) ** 2) for i in range(mp.shape[1]))
    T = int(np.ceil(far / w)) + dt
    G = rng.standard_normal((mp.shape[1], T)).astype(np.float32)
    got = gpu_a1(G, mp, (X, Y, Z), w, 0.0)
    ref = O.a1_transpose(G, mp, X, Y, Z, ORIGIN, SIZE, w, 0.0, mode=O.PINNED64)
    e = rel_l2(got, ref)
    print(f"{scan} T={T}: rel-L2 vs FP64 pinned {e:.2e}")
    assert np.all(np.isfinite(got)) and e < TOL


def gpu_a6(vol, mp, G, w, off, origin=ORIGIN, size=SIZE):
    from totrilib_amd.totri_cpp.render import volume_confocal_grad_measurement_points
    v = torch.from_numpy(np.ascontiguousarray(vol)).to(dev())
    m = torch.from_numpy(np.ascontiguousarray(mp)).to(dev())
    g = torch.as_tensor(G).to(dev())
    out = torch.full((3, m.shape[1]), float("nan"), device=dev())
    volume_confocal_grad_measurement_points(out, v, m, g, origin, size, w, off)
    torch.cuda.synchronize()
    return out.cpu().numpy().astype(np.float64)


def test_a6_matches_oracle():
    rng = np.random.default_rng(21)
    X, Y, Z, T = 20, 18, 24, 700
    vol = rng.random((X, Y, Z)).astype(np.float32)
    mp = np.stack([rng.uniform(-1, 1, 21), rng.uniform(-1, 1, 21),
                   rng.uniform(-0.1, 0.1, 21)]).astype(np.float32)
    w, off = 0.004, 0.001
    G = rng.standard_normal((21, T)).astype(np.float32)
    got = gpu_a6(vol, mp, G, w, off)
    ref64 = O.a6(vol, mp, G, ORIGIN, SIZE, w, off, mode=O.PINNED64)
    ref32 = O.a6(vol, mp, G, ORIGIN, SIZE, w, off, mode=O.FP32)
    e64, e32 = rel_l2(got, ref64), rel_l2(got, ref32)
    print(f"A6 rel-L2 vs FP64 pinned {e64:.2e}, vs FP32 {e32:.2e}")
    assert e64 < TOL


@pytest.mark.parametrize("scan", ["mirror", "xpairs", "odd_y", "perturbed"])
def test_a6_mirror_quads_and_general_path(scan):
    """Mirror quads of scan points on a box centred at x = y = 0 take the
    symmetric A6 (one decision per four scan points, gradients with the
    mirror signs); a box off-centre in y ("xpairs") or an odd Y ("odd_y")
    the x-pair A6 (one decision per two); one point moved by an ulp the
    general kernel.  All match the oracle.  (16 cells over [-1, 1] have
    antisymmetric centres in float; 18 or 20 do not.)"""
    rng = np.random.default_rng(23)
    X, Y, Z, T = 16, 15 if scan == "odd_y" else 16, 24, 700
    origin = (ORIGIN[0], ORIGIN[1] + 0.1, ORIGIN[2]) if scan == "xpairs" else ORIGIN
    vol = rng.random((X, Y, Z)).astype(np.float32)
    mp = wall_grid(6)
    if scan == "perturbed":
        mp[0, 7] = np.nextafter(mp[0, 7], np.float32(2))
    w, off = 0.004, 0.001
    G = rng.standard_normal((mp.shape[1], T)).astype(np.float32)
    info = mirror_info(mp)
    assert info[:2] == ([0, info[1]] if scan == "perturbed" else [1, 9])
    assert r1_paths((X, Y, Z), mp.shape[1], T, origin) == (2 if scan in ("xpairs", "odd_y")
                                                           else 3)
    got = gpu_a6(vol, mp, G, w, off, origin=origin)
    ref64 = O.a6(vol, mp, G, origin, SIZE, w, off, mode=O.PINNED64)
    e64 = rel_l2(got, ref64)
    print(f"{scan}: A6 rel-L2 vs FP64 pinned {e64:.2e}")
    assert e64 < TOL


@pytest.mark.parametrize("layout", ["inside", "two_walls"])
def test_a6_general_wall_inside_the_box(layout):
    """The general A6 with a wall inside the box's depth range, alone or next
    to a wall below it.  Checked against the oracle's FP32 mode: its FP64
    mode (decisions pinned from FP32, weights in FP64) is not defined here,
    since a near-wall sample can have d2 > 0 in FP32 and d2 <= 0 in FP64,
    and its 1/depth term is then NaN."""
    rng = np.random.default_rng(24)
    X, Y, Z, T = 20, 18, 24, 700
    vol = rng.random((X, Y, Z)).astype(np.float32)
    mp = wall_grid(6)
    mp[0, 7] = np.nextafter(mp[0, 7], np.float32(2))   # general path
    origin = ORIGIN
    if layout == "inside":
        origin = (-1.0, -1.0, -0.5)
    else:
        upper = mp.copy()
        upper[2] = 1.2
        mp = np.ascontiguousarray(np.concatenate([mp, upper], axis=1))
    w, off = 0.004, 0.001
    G = rng.standard_normal((mp.shape[1], T)).astype(np.float32)
    assert mirror_info(mp)[0] == 0
    got = gpu_a6(vol, mp, G, w, off, origin=origin)
    ref32 = O.a6(vol, mp, G, origin, SIZE, w, off, mode=O.FP32)
    e = rel_l2(got, ref32)
    print(f"{layout}: A6 rel-L2 vs FP32 oracle {e:.2e}")
    assert np.all(np.isfinite(got))
    assert e < TOL


def test_autograd_volume_and_points():
    """VolumeRenderConfocal.apply: forward R1, backward A1 (albedo) and A6
    (wall points), as the reference's autograd Function."""
    from totrilib_amd.render import VolumeRenderConfocal
    rng = np.random.default_rng(22)
    vol_np = rng.random((12, 10, 16)).astype(np.float32)
    mp_np = np.stack([rng.uniform(-1, 1, 9), rng.uniform(-1, 1, 9),
                      np.zeros(9)]).astype(np.float32)
    T, w = 400, 0.006
    v = torch.from_numpy(vol_np).to(dev()).requires_grad_(True)
    m = torch.from_numpy(mp_np).to(dev()).requires_grad_(True)
    tr = VolumeRenderConfocal.apply(v, m, ORIGIN, SIZE, T, w, 0.0)
    G = torch.from_numpy(rng.standard_normal((9, T)).astype(np.float32)).to(dev())
    (tr * G).sum().backward()
    Gn = G.cpu().numpy()
    ev = rel_l2(v.grad.cpu().numpy(),
                O.a1_transpose(Gn, mp_np, 12, 10, 16, ORIGIN, SIZE, w, 0.0, mode=O.PINNED64))
    em = rel_l2(m.grad.cpu().numpy(), O.a6(vol_np, mp_np, Gn, ORIGIN, SIZE, w, 0.0,
                                           mode=O.PINNED64))
    assert ev < TOL and em < TOL, (ev, em)


def r1_paths(shape, n, T, origin=ORIGIN):
    from totrilib_amd import _lib
    lib = _lib.load()
    return lib.totri_debug_r1_paths(*shape, n, T, _lib.f3(origin), _lib.f3(SIZE))


@pytest.mark.parametrize("scan", ["mirror", "xpairs", "odd_y", "xpairs_only", "perturbed",
                                  "random"])
def test_r1_mirror_quads_and_general_path(scan):
    """A scan of mirror quads on a box centred at x = y = 0 takes the
    symmetric kernel (one decision per four samples).  The x-pair kernel (one
    decision per two samples) takes a box off-centre in y ("xpairs"), an odd
    Y ("odd_y") or a scan of x-pairs without y-mirrors ("xpairs_only"); one
    moved point or a random scan takes the general kernel.  All match the
    oracle."""
    rng = np.random.default_rng(31)
    # (Y = 32: cell centres of [-1, 1] antisymmetric in float; Y = 30's are not)
    X, Y, Z, T = 32, 31 if scan == "odd_y" else 32, 40, 700
    origin = (ORIGIN[0], ORIGIN[1] + 0.1, ORIGIN[2]) if scan == "xpairs" else ORIGIN
    vol = sparse_albedo((X, Y, Z), 0.2, 32)
    mp = wall_grid(8)
    if scan == "perturbed":
        mp[0, 5] = np.nextafter(mp[0, 5], np.float32(2))
    elif scan == "random":
        mp = np.stack([rng.uniform(-1, 1, 64), rng.uniform(-1, 1, 64),
                       np.zeros(64)]).astype(np.float32)
    elif scan == "xpairs_only":
        # x-pairs at random y: no point has a y-mirror
        h = np.stack([rng.uniform(0.05, 1, 32), rng.uniform(-1, 1, 32)]).astype(np.float32)
        mp = np.concatenate([np.concatenate([h, h * [[-1], [1]]], 1), np.zeros((1, 64))],
                            0).astype(np.float32)
    w = 0.0045
    info = mirror_info(mp)
    want = {"mirror": [1, 16, 1, 32], "xpairs": [1, 16, 1, 32], "odd_y": [1, 16, 1, 32],
            "xpairs_only": [0, info[1], 1, 32]}.get(scan, [0, info[1], 0, info[3]])
    assert info == want, info
    paths = r1_paths((X, Y, Z), mp.shape[1], T, origin)
    assert paths == (3 if scan in ("mirror", "xpairs_only", "perturbed", "random") else 2)
    got = gpu_r1(vol, mp, T, w, 0.0, origin=origin)
    ref = O.r1_forward(vol, mp, T, origin, SIZE, w, 0.0, mode=O.PINNED64)
    e = rel_l2(got, ref)
    print(f"{scan}: mirror info {info}, R1 paths {paths}, rel-L2 vs FP64 pinned {e:.2e}")
    assert e < TOL
    assert np.count_nonzero(got) > 0


@pytest.mark.parametrize("scan", ["mirror", "xpairs", "random"])
@pytest.mark.parametrize("layout", ["inside", "two_walls"])
def test_r1_mirror_wall_inside_the_box(layout, scan):
    """The mirror and the general R1 take a shorter activity test in
    workgroups whose wall points all lie below the box (a sample with
    d2 <= 0 then decides a negative row).  A wall inside the box's depth range, alone or next to a
    wall below it (so that workgroups take both tests), matches the oracle.
    With the wall inside the box, near-wall samples make d2 a difference of
    nearly equal squares, and the reference's own FP32 arithmetic (the
    oracle's FP32 mode) is about 2e-5 from the FP64 oracle: the check is
    against the FP32 oracle, and the FP64 error may not exceed 1.5x the FP32 oracle's."""
    X, Y, Z, T = 32, 32, 40, 700
    vol = sparse_albedo((X, Y, Z), 0.2, 33)
    rng = np.random.default_rng(34)

    def wall(z):
        if scan in ("mirror", "xpairs"):
            m = wall_grid(8)
        else:
            m = np.stack([rng.uniform(-1, 1, 64), rng.uniform(-1, 1, 64),
                          np.zeros(64)]).astype(np.float32)
        m[2] = z
        return m

    origin = ORIGIN
    if layout == "inside":
        origin = (-1.0, -1.0, -0.5)
        mp = wall(0.0)
    else:
        mp = np.ascontiguousarray(np.concatenate([wall(0.0), wall(1.2)], axis=1))
    if scan == "xpairs":  # the box off-centre in y: the x-pair kernel
        origin = (origin[0], origin[1] + 0.1, origin[2])
    w = 0.0045
    info = mirror_info(mp)
    if scan != "random":
        assert info[0] == 1 and info[1] == mp.shape[1] // 4
        assert r1_paths((X, Y, Z), mp.shape[1], T, origin) == (2 if scan == "xpairs" else 3)
    else:
        assert info[0] == 0
    got = gpu_r1(vol, mp, T, w, 0.0, origin=origin)
    ref = O.r1_forward(vol, mp, T, origin, SIZE, w, 0.0, mode=O.PINNED64)
    r32 = O.r1_forward(vol, mp, T, origin, SIZE, w, 0.0, mode=O.FP32)
    e, e32, eref = rel_l2(got, r32), rel_l2(got, ref), rel_l2(r32, ref)
    print(f"{layout}/{scan}: rel-L2 vs FP32 oracle {e:.2e}, vs FP64 {e32:.2e} "
          f"(FP32 oracle {eref:.2e})")
    assert e < TOL
    assert e32 <= max(TOL, 1.5 * eref)
    assert np.count_nonzero(got) > 0


@pytest.mark.parametrize("scan", ["quads", "quads_safe", "quads_negoff", "xpairs", "odd_y",
                                  "odd_y_safe", "odd_y_negoff", "perturbed", "random"])
def test_a1_mirror_pairs_and_general_path(scan):
    """Path selection of A1, each checked against the oracle's exact adjoint.

    - Mirror quads of scan points on a box centred at x = y = 0, with w > 0
      and offset >= 0, take the xy-mirror walk (one decision per four
      samples): "quads", and "quads_safe" whose offset 0.5 mm makes 1/tc^8
      overflow-prone, so the table takes its safe form.
    - x-pairs without y-mirrors or an odd Y allow only the x-mirror paths
      (one decision per two samples): the x-pair row walk ("xpairs",
      "odd_y", and "odd_y_safe" with the safe table form), or, where a
      negative offset (tc < 0 for the first bins) fails the row walks'
      monotonicity condition, the x-mirror window walk ("odd_y_negoff").
      "quads_negoff" takes the window walk too.
    - A broken pair or a random scan takes the general kernel.
    The grid is 32 x 32 (31 for an odd Y): cell counts over [-1, 1] that are
    powers of two give centres that are antisymmetric in float, so the
    mirror walks apply (the walks the planner allows are asserted).
    """
    rng = np.random.default_rng(33)
    X, Y, Z, T = 32, 32, 44, 600
    mp = wall_grid(8)
    w, off = 0.0045, 0.0
    if scan == "quads_safe":
        off = 0.0005
    elif scan == "quads_negoff":
        off = -0.05
    elif scan == "xpairs":
        # move one x-pair in y: x-pairs stay, the y mirror of that pair is lost
        k = 9
        j = int(np.flatnonzero((mp[0] == -mp[0, k]) & (mp[1] == mp[1, k]))[0])
        mp[1, [k, j]] = np.nextafter(mp[1, k], np.float32(2))
    elif scan in ("odd_y", "odd_y_safe", "odd_y_negoff"):
        Y = 31
        off = {"odd_y": 0.0, "odd_y_safe": 0.0005, "odd_y_negoff": -0.05}[scan]
    elif scan == "perturbed":
        mp[1, 9] = np.nextafter(mp[1, 9], np.float32(2))   # breaks one x-pair
        mp[0, 9] = np.nextafter(mp[0, 9], np.float32(2))
    elif scan == "random":
        mp = np.stack([rng.uniform(-1, 1, 64), rng.uniform(-1, 1, 64),
                       np.zeros(64)]).astype(np.float32)
    G = rng.standard_normal((mp.shape[1], T)).astype(np.float32)
    info = mirror_info(mp)
    expect = {"quads": [1, 16, 1, 32], "quads_safe": [1, 16, 1, 32],
              "quads_negoff": [1, 16, 1, 32], "odd_y": [1, 16, 1, 32],
              "odd_y_safe": [1, 16, 1, 32], "odd_y_negoff": [1, 16, 1, 32],
              "xpairs": [0, info[1], 1, 32]}.get(scan, [0, info[1], 0, info[3]])
    assert info == expect
    from totrilib_amd import _lib
    walks = _lib.load().totri_debug_a1_walks(X, Y, Z, mp.shape[1], T, _lib.f3(ORIGIN),
                                             _lib.f3(SIZE), w, off)
    # bits: 1 x-mirror paths, 2 xy-mirror row walk, 4 general row walk, 8 x-pair row walk
    assert walks == {"quads_negoff": 1, "odd_y": 13, "odd_y_safe": 13,
                     "odd_y_negoff": 1}.get(scan, 15), walks
    got = gpu_a1(G, mp, (X, Y, Z), w, off)
    ref = O.a1_transpose(G, mp, X, Y, Z, ORIGIN, SIZE, w, off, mode=O.PINNED64)
    e = rel_l2(got, ref)
    print(f"{scan}: mirror info {info}, walks {walks}, rel-L2 vs FP64 pinned {e:.2e}")
    assert e < TOL
    assert np.count_nonzero(got) > 0


@pytest.mark.parametrize("scan", ["quads", "random"])
def test_a1_transient_longer_than_float_bin_counter(scan):
    """T = 2^24 + 64 bins: the row walks count bins in a float, which stops
    advancing at 2^24, so such transients take the window walks (x-mirror for
    the quad scan, general for the random one).  The gradient is nonzero only
    in bins near the top of the range, where the geometry puts samples in the
    volume (bin width 2.6 m / 2^24), and matches the oracle's exact
    transpose."""
    from totrilib_amd import _lib
    X, Y, Z = 4, 4, 8
    T = (1 << 24) + 64
    w = 2.6 / (1 << 24)
    rng = np.random.default_rng(77)
    if scan == "quads":
        mp = wall_grid(2) * np.float32(0.5)
    else:
        mp = np.stack([rng.uniform(-0.5, 0.5, 3), rng.uniform(-0.5, 0.5, 3),
                       np.zeros(3)]).astype(np.float32)
    lib = _lib.load()
    walks = lib.totri_debug_a1_walks(X, Y, Z, mp.shape[1], T, _lib.f3(ORIGIN), _lib.f3(SIZE),
                                     w, 0.0)
    assert walks == (1 if scan == "quads" else 0)
    G = np.zeros((mp.shape[1], T), np.float32)
    for i in range(mp.shape[1]):
        hot = rng.integers(T - 3_000_000, T, 3000)
        G[i, hot] = rng.random(3000).astype(np.float32)
    print(f"{scan}: N={mp.shape[1]} T={T}", flush=True)
    got = gpu_a1(torch.from_numpy(G), mp, (X, Y, Z), w, 0.0)
    ref = O.a1_transpose(G, mp, X, Y, Z, ORIGIN, SIZE, w, 0.0, mode=O.PINNED64)
    e = rel_l2(got, ref)
    print(f"{scan}: rel-L2 vs FP64 pinned {e:.2e}, nonzero {np.count_nonzero(ref)}")
    assert np.count_nonzero(ref) > 20
    assert e < TOL


@pytest.mark.parametrize("shape,n,off", [((16, 14, 24), 6, 0.0), ((10, 12, 100), 5, 0.03)])
def test_a1_reference_window_form_opt_in(shape, n, off):
    """apply_transposed(..., reference_window=True) computes the reference's
    own window form (volume_cuda.cu:95-160): bit for bit the FP32 oracle's
    restatement (correctly rounded roots and quotients in the same order),
    including the top slice, where it differs from the exact transpose.
    Z = 100 puts a column's voxels in two waves."""
    from totrilib_amd.render import VolumeRenderConfocal
    X, Y, Z = shape
    T, w = 500, 0.006
    rng = np.random.default_rng(sum(shape))
    mp = wall_grid(n)
    mp[0] += rng.uniform(-0.01, 0.01, mp.shape[1]).astype(np.float32)
    G = rng.standard_normal((mp.shape[1], T)).astype(np.float32)
    m = torch.from_numpy(mp).to(dev())
    g = torch.from_numpy(G).to(dev())
    got = VolumeRenderConfocal.apply_transposed(g, m, ORIGIN, SIZE, shape, w, off,
                                                reference_window=True).cpu().numpy()
    ref32 = O.a1_reference(G, mp, X, Y, Z, ORIGIN, SIZE, w, off, mode=O.FP32)
    assert np.array_equal(got.astype(np.float64), ref32)
    ref64 = O.a1_reference(G, mp, X, Y, Z, ORIGIN, SIZE, w, off, mode=O.PINNED64)
    assert rel_l2(got, ref64) < TOL
    exact = VolumeRenderConfocal.apply_transposed(g, m, ORIGIN, SIZE, shape, w, off).cpu().numpy()
    top = rel_l2(got[:, :, -1], exact[:, :, -1])
    rest = rel_l2(got[:, :, :-1], exact[:, :, :-1])
    print(f"{shape}: window form vs exact transpose: top slice {top:.2e}, rest {rest:.2e}")
    assert top > 1e-2 and rest < 1e-3
