"""GPU parity of the backprojections (A2, A3, A4), the exhaustive render and
its adjoint (R2, A5) and the RBF splat (S1, S2) against the CPU oracle,
through the drop-in Python layer (totri_cpp bindings and autograd Functions).
"""
import numpy as np
import pytest
import torch

from oracle import oracle as O
from tests.cases import rel_l2, sparse_albedo, wall_grid

pytestmark = pytest.mark.gpu

VS, VE = (-1.0, -1.0, 0.5), (1.0, 1.0, 2.5)


def dev():
    return torch.device("cuda:0")


def t(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev())


def test_velten_confocal_cfg1_matches_oracle():
    """Velten backprojection of a noisy cfg1 transient into 64^3 (round-trip
    bins of 0.01 m), with and without a scan origin."""
    from totrilib_amd.reconstruct import BackprojectionConfocalVelten
    from totrilib_amd.types import BinDef, VolumeDef
    mp = wall_grid(32)                         # [3, S]
    scan = mp.T[None].copy()                   # [1, S, 3]
    rng = np.random.default_rng(51)
    T = 512
    tr = rng.random((1, T, mp.shape[1])).astype(np.float32)
    vd = VolumeDef(VS, VE, (64, 64, 64))
    bd = BinDef(0.01, 0.0, T)
    got = BackprojectionConfocalVelten.apply(t(tr), t(scan), vd, bd).cpu().numpy()
    ref = O.velten_confocal(tr, scan, None, (64, 64, 64), VS, VE, 0.01, 0.0, mode=O.FP32)
    assert np.array_equal(got.astype(np.float64), ref), rel_l2(got, ref)
    org = np.array([[0.0, 0.0, -1.0]], np.float32)
    got = BackprojectionConfocalVelten.apply(t(tr), t(scan), vd, bd, t(org)).cpu().numpy()
    ref = O.velten_confocal(tr, scan, np.broadcast_to(org[:, None], scan.shape), (64, 64, 64),
                            VS, VE, 0.01, 0.0, mode=O.FP32)
    assert np.array_equal(got.astype(np.float64), ref), rel_l2(got, ref)


@pytest.mark.parametrize("case", ["exact_formula", "mixed", "batch_origins"])
def test_velten_confocal_decision_paths(case):
    """A2's two per-scan-point decision paths, bit for bit against the FP32
    oracle.  "exact_formula": 0.05 mm bins 2.0 m out put every scan point
    beyond the estimate's 2^16-bin bound, so each pair takes the reference's
    formula.  "mixed": per-point origins 0.2 m or 14 m away put half the
    points on the estimate and half on the formula, all within 70,000 bins.  "batch_origins": two batches with
    their own scan points and origins (each batch its own threshold rows)."""
    from totrilib_amd.reconstruct import BackprojectionConfocalVelten
    from totrilib_amd.types import BinDef, VolumeDef
    rng = np.random.default_rng(57)
    res = (32, 28, 36)
    if case == "exact_formula":
        B, n, T, w, off = 1, 16, 4096, 0.00005, 2.0
    elif case == "mixed":
        B, n, T, w, off = 1, 8, 70000, 0.0005, 1.0
    else:
        B, n, T, w, off = 2, 12, 700, 0.004, 0.3
    S = n * n
    scan = np.stack([wall_grid(n).T + rng.uniform(-0.01, 0.01, (S, 3)).astype(np.float32)
                     for _ in range(B)])
    if case == "exact_formula":
        org = None
    elif case == "mixed":
        far = rng.uniform(0, 1, S) < 0.5
        org = scan + np.where(far[:, None], [0.0, 0.0, -14.0], [0.0, 0.0, -0.2])
        org = org.astype(np.float32)
    else:
        org = (scan + rng.uniform(-0.5, 0.5, (B, S, 3))).astype(np.float32)
    tr = rng.random((B, T, S)).astype(np.float32)
    vd = VolumeDef(VS, VE, res)
    bd = BinDef(w, off, T)
    args = (t(tr), t(scan), vd, bd) + ((t(org),) if org is not None else ())
    got = BackprojectionConfocalVelten.apply(*args).cpu().numpy()
    ref = O.velten_confocal(tr, scan, org, res, VS, VE, w, off, mode=O.FP32)
    assert np.count_nonzero(ref) > 1000
    assert np.array_equal(got.astype(np.float64), ref), rel_l2(got, ref)


def test_velten_exhaustive_matches_oracle():
    from totrilib_amd.reconstruct import BackprojectionExhaustiveVelten
    from totrilib_amd.types import BinDef, VolumeDef
    rng = np.random.default_rng(52)
    B, T, S, L = 2, 300, 7, 3
    res = (20, 18, 16)
    tr = rng.random((B, T, S, L)).astype(np.float32)
    scan = np.concatenate([rng.uniform(-1, 1, (B, S, 2)), np.zeros((B, S, 1))], 2).astype(np.float32)
    laser = np.concatenate([rng.uniform(-1, 1, (B, L, 2)), np.zeros((B, L, 1))], 2).astype(np.float32)
    so = np.array([[0.0, 0.1, -1.0], [0.1, 0.0, -1.2]], np.float32)
    vd, bd = VolumeDef(VS, VE, res), BinDef(0.02, 0.1, T)
    got = BackprojectionExhaustiveVelten.apply(t(tr), t(scan), t(laser), vd, bd, t(so)).cpu().numpy()
    ref = O.velten_exhaustive(tr, scan, laser, np.broadcast_to(so[:, None], scan.shape), None,
                              res, VS, VE, 0.02, 0.1, mode=O.FP32)
    assert np.array_equal(got.astype(np.float64), ref), rel_l2(got, ref)


def test_velten_exhaustive_laser_origin_matches_oracle():
    """Both origin terms of A3 (backprojection_cuda.cu:268-279): a per-point
    scan origin and a per-point laser origin, then each alone."""
    from totrilib_amd.reconstruct import BackprojectionExhaustiveVelten
    from totrilib_amd.types import BinDef, VolumeDef
    rng = np.random.default_rng(56)
    B, T, S, L = 2, 400, 5, 3
    res = (18, 16, 20)
    tr = rng.random((B, T, S, L)).astype(np.float32)
    scan = np.concatenate([rng.uniform(-1, 1, (B, S, 2)), np.zeros((B, S, 1))], 2).astype(np.float32)
    laser = np.concatenate([rng.uniform(-1, 1, (B, L, 2)), np.zeros((B, L, 1))], 2).astype(np.float32)
    so = (scan + rng.uniform(-0.3, 0.3, scan.shape) - [0, 0, 1.0]).astype(np.float32)
    lo = (laser + rng.uniform(-0.3, 0.3, laser.shape) - [0, 0, 0.6]).astype(np.float32)
    vd, bd = VolumeDef(VS, VE, res), BinDef(0.015, 0.2, T)
    for sorg, lorg in ((so, lo), (None, lo), (so, None)):
        got = BackprojectionExhaustiveVelten.apply(
            t(tr), t(scan), t(laser), vd, bd, None if sorg is None else t(sorg),
            None if lorg is None else t(lorg)).cpu().numpy()
        ref = O.velten_exhaustive(tr, scan, laser, sorg, lorg, res, VS, VE, 0.015, 0.2,
                                  mode=O.FP32)
        assert np.count_nonzero(ref) > 1000
        assert np.array_equal(got.astype(np.float64), ref), rel_l2(got, ref)


@pytest.mark.parametrize("case", ["exact_formula", "mixed", "long_rows"])
def test_velten_exhaustive_decision_paths(case):
    """A3's two per-pair decision paths, bit for bit against the FP32 oracle.
    "exact_formula": 0.03 mm bins put every pair beyond the estimate's
    2^16-bin bound, so each (voxel, pair) divides as the reference does.
    "mixed": laser origins 0.3 m or 28 m away put half the pairs on each
    path within 72,000 bins.  "long_rows": 2,500 bins, rows too long for the
    LDS stage (read from global memory)."""
    from totrilib_amd.reconstruct import BackprojectionExhaustiveVelten
    from totrilib_amd.types import BinDef, VolumeDef
    rng = np.random.default_rng(58)
    B, S, L = 1, 4, 4
    res = (20, 18, 16)
    T, w, off = {"exact_formula": (4000, 0.00003, 2.0), "mixed": (72000, 0.0005, 0.0),
                 "long_rows": (2500, 0.002, 1.0)}[case]
    scan = np.concatenate([rng.uniform(-1, 1, (B, S, 2)), np.zeros((B, S, 1))], 2).astype(np.float32)
    laser = np.concatenate([rng.uniform(-1, 1, (B, L, 2)), np.zeros((B, L, 1))], 2).astype(np.float32)
    lorg = None
    if case == "mixed":
        far = np.array([True, False, True, False])[None, :, None]
        lorg = (laser + np.where(far, [0.0, 0.0, -28.0], [0.0, 0.0, -0.3])).astype(np.float32)
    tr = rng.random((B, T, S, L)).astype(np.float32)
    vd, bd = VolumeDef(VS, VE, res), BinDef(w, off, T)
    got = BackprojectionExhaustiveVelten.apply(
        t(tr), t(scan), t(laser), vd, bd, None, None if lorg is None else t(lorg)).cpu().numpy()
    ref = O.velten_exhaustive(tr, scan, laser, None, lorg, res, VS, VE, w, off, mode=O.FP32)
    assert np.count_nonzero(ref) > 500
    assert np.array_equal(got.astype(np.float64), ref), rel_l2(got, ref)


def test_arellano_with_origins_matches_oracle():
    """Arellano with scan != laser and both origins (tc -= d_so + d_lo,
    backprojection_cuda.cu:38-54,79-80), then each origin alone and an origin
    given per batch ([B, 3]).  Hit counts bit for bit with a transient of
    ones; sums to rounding."""
    from totrilib_amd.reconstruct import BackprojectionExhaustiveArellano
    from totrilib_amd.types import BinDef, VolumeDef
    rng = np.random.default_rng(57)
    B, T, S, L = 2, 256, 4, 3
    res = (40, 18, 20)
    tr = (rng.random((B, T, S, L)) * (rng.random((B, T, S, L)) < 0.3)).astype(np.float32)
    scan = np.concatenate([rng.uniform(-1, 1, (B, 2, S)), np.zeros((B, 1, S))], 1).astype(np.float32)
    laser = np.concatenate([rng.uniform(-1, 1, (B, 2, L)), np.zeros((B, 1, L))], 1).astype(np.float32)
    so = (scan + rng.uniform(-0.2, 0.2, scan.shape) - np.array([0, 0, 0.4])[None, :, None]).astype(np.float32)
    lo = (laser + rng.uniform(-0.2, 0.2, laser.shape) - np.array([0, 0, 0.3])[None, :, None]).astype(np.float32)
    so_b = np.array([[0.0, 0.0, -0.5], [0.1, -0.1, -0.45]], np.float32)
    vd, bd = VolumeDef(VS, VE, res), BinDef(0.025, 0.8, T)
    ones = np.ones_like(tr)
    cases = ((so, lo, so, lo), (None, lo, None, lo), (so, None, so, None),
             (so_b, lo, np.broadcast_to(so_b[:, :, None], scan.shape), lo))
    for sorg, lorg, sref, lref in cases:
        args = (None if sorg is None else t(sorg), None if lorg is None else t(lorg))
        got1 = BackprojectionExhaustiveArellano.apply(t(ones), t(scan), t(laser), vd, bd,
                                                      *args).cpu().numpy()
        ref1 = O.arellano(ones, scan, laser, sref, lref, res, VS, VE, 0.025, 0.8, mode=O.FP32)
        assert np.count_nonzero(ref1) > 100
        assert np.array_equal(got1.astype(np.float64), ref1)
        got = BackprojectionExhaustiveArellano.apply(t(tr), t(scan), t(laser), vd, bd,
                                                     *args).cpu().numpy()
        ref = O.arellano(tr, scan, laser, sref, lref, res, VS, VE, 0.025, 0.8, mode=O.FP32)
        assert np.array_equal(got != 0, ref != 0)
        assert rel_l2(got, ref) < 1e-6


@pytest.mark.parametrize("B,res,S,L", [(1, (40, 20, 24), 4, 3), (2, (300, 4, 6), 5, 2)])
def test_arellano_matches_oracle(B, res, S, L):
    """Voxel decisions bit for bit: with a transient of ones every voxel holds
    its exact hit count, whatever the summation order.  Sums of a sparse
    random transient (positive bins are listed and split over waves, then
    partial volumes are added in order) agree with the sequential FP32 oracle
    to rounding.  D = 300 exercises the global-memory column path."""
    from totrilib_amd.reconstruct import BackprojectionExhaustiveArellano
    from totrilib_amd.types import BinDef, VolumeDef
    rng = np.random.default_rng(53)
    T = 256
    tr = (rng.random((B, T, S, L)) * (rng.random((B, T, S, L)) < 0.3)).astype(np.float32)
    scan = np.concatenate([rng.uniform(-1, 1, (B, 2, S)), np.zeros((B, 1, S))], 1).astype(np.float32)
    laser = np.concatenate([rng.uniform(-1, 1, (B, 2, L)), np.zeros((B, 1, L))], 1).astype(np.float32)
    vd, bd = VolumeDef(VS, VE, res), BinDef(0.025, 0.0, T)
    ones = np.ones_like(tr)
    got1 = BackprojectionExhaustiveArellano.apply(t(ones), t(scan), t(laser), vd, bd).cpu().numpy()
    ref1 = O.arellano(ones, scan, laser, None, None, res, VS, VE, 0.025, 0.0, mode=O.FP32)
    assert np.count_nonzero(ref1) > 100
    assert np.array_equal(got1.astype(np.float64), ref1)
    got = BackprojectionExhaustiveArellano.apply(t(tr), t(scan), t(laser), vd, bd).cpu().numpy()
    ref = O.arellano(tr, scan, laser, None, None, res, VS, VE, 0.025, 0.0, mode=O.FP32)
    assert np.array_equal(got != 0, ref != 0)
    assert rel_l2(got, ref) < 1e-6


def test_exhaustive_render_and_adjoint_match_oracle():
    from totrilib_amd.render.volume import VolumeRenderExhaustive
    from totrilib_amd.types import BinDef, VolumeDef
    rng = np.random.default_rng(54)
    B, T, S, L = 2, 400, 6, 2
    res = (24, 20, 22)
    vol = np.stack([sparse_albedo(res, 0.05, 55 + b) for b in range(B)])
    scan = np.concatenate([rng.uniform(-1, 1, (B, 2, S)), np.zeros((B, 1, S))], 1).astype(np.float32)
    laser = np.concatenate([rng.uniform(-1, 1, (B, 2, L)), np.zeros((B, 1, L))], 1).astype(np.float32)
    vd, bd = VolumeDef(VS, VE, res), BinDef(0.01, 0.2, T)
    v = t(vol).requires_grad_(True)
    tr = VolumeRenderExhaustive.apply(v, t(scan), t(laser), vd, bd)
    ref = O.r2_forward(vol, scan, laser, T, VS, VE, 0.01, 0.2, mode=O.PINNED64)
    got = tr.detach().cpu().numpy()
    assert np.array_equal(got != 0, ref != 0)
    assert rel_l2(got, ref) < 1e-6
    G = rng.standard_normal((B, T, S, L)).astype(np.float32)
    (tr * t(G)).sum().backward()
    adj = O.r2_adjoint(G, res, scan, laser, VS, VE, 0.01, 0.2, mode=O.PINNED64)
    assert rel_l2(v.grad.cpu().numpy(), adj) < 1e-6


@pytest.mark.parametrize("trunc,A", [(0.0, 1), (3.0, 1), (0.0, 0), (0.0, 3), (3.0, 3)])
def test_splat_forward_and_grad_match_oracle(trunc, A, report):
    """Forward and RBF gradient against the oracle with 0, 1 and 3 attributes
    per RBF, untruncated (shared voxel tiles) and truncated at 3 sigma (one
    wave per RBF over its support box)."""
    from totrilib_amd.types import VolumeDef
    from totrilib_amd.util import Rbf2VolumeGaussian
    rng = np.random.default_rng(56 + A)
    N = 300
    res = (40, 36, 32)
    rbf = np.concatenate([rng.uniform(-1, 1, (1, N, 2)), rng.uniform(0.5, 2.5, (1, N, 1)),
                          rng.uniform(0.02, 0.1, (1, N, 1)), rng.random((1, N, A))],
                         2).astype(np.float32)
    r = t(rbf).requires_grad_(True)
    vol = Rbf2VolumeGaussian.apply(r, VolumeDef(VS, VE, res), trunc)
    got = vol.detach().cpu().numpy()
    ref = O.splat_forward(rbf, res, VS, VE, trunc=trunc, mode=O.FP32)
    assert rel_l2(got, ref) < 1e-5
    G = rng.standard_normal(got.shape).astype(np.float32)
    (vol * t(G)).sum().backward()
    gref = O.splat_grad(rbf, ref, G, VS, VE, trunc=trunc, mode=O.PINNED64)
    e = rel_l2(r.grad.cpu().numpy(), gref)
    report(f"splat trunc={trunc}, {A} attributes: splat rel-L2 vs FP32 oracle {rel_l2(got, ref):.2e}, "
           f"RBF gradient rel-L2 vs FP64 {e:.2e}")
    assert e < 1e-5


def test_splat_render_chain_gradient_matches_fd(report):
    """RBF weights, centres and sigma through the untruncated splat (the
    reference's form; albedo = V0 * V1), the confocal render and an L2 loss:
    the GPU gradient against central differences of the FP64 chain.

    The render is linear in the albedo, so the chain's derivative is
    2 <R alb - y, R (d alb / d theta)> exactly; only the splat is
    differenced, in FP64 (oracle mode 2), with steps of 2e-3 sigma.  The bar
    is tied to the differences' own error, estimated from steps h and h/2,
    and to FP32 rounding of the GPU chain (1e-5 of the largest gradient)."""
    from totrilib_amd.render import VolumeRenderConfocal
    from totrilib_amd.types import VolumeDef
    from totrilib_amd.util import Rbf2VolumeGaussian
    rng = np.random.default_rng(57)
    N, R, T, w = 12, 24, 192, 0.01
    box_o, box_s = (-1.0, -1.0, 0.5), (2.0, 2.0, 2.0)
    dv = 2.0 / R
    vd = VolumeDef((-1 + dv / 2, -1 + dv / 2, 0.5 + dv / 2), (1 - dv / 2, 1 - dv / 2, 2.5 - dv / 2),
                   (R, R, R))
    rbf = np.concatenate([rng.uniform(-0.7, 0.7, (1, N, 2)), rng.uniform(1.0, 2.0, (1, N, 1)),
                          rng.uniform(0.1, 0.2, (1, N, 1)), rng.random((1, N, 1))], 2).astype(np.float32)
    mp = wall_grid(6)
    target = rng.random((mp.shape[1], T)).astype(np.float32) * 1e-3

    def albedo64(params):
        v = O.splat_forward(params, (R, R, R), vd.start, vd.end, trunc=0.0, mode=O.PURE64)[0]
        return (v[0] * v[1]).transpose(2, 1, 0)  # [D,H,W] -> [X,Y,Z]

    def render64(alb):
        return O.r1_forward(alb.astype(np.float32), mp, T, box_o, box_s, w, 0.0,
                            mode=O.PINNED64)

    resid = render64(albedo64(rbf)) - target

    r = t(rbf).requires_grad_(True)
    vol = Rbf2VolumeGaussian.apply(r, vd, 0.0)[0]
    alb = (vol[0] * vol[1]).permute(2, 1, 0).contiguous()
    tr = VolumeRenderConfocal.apply(alb, t(mp), box_o, box_s, T, w, 0.0)
    loss = ((tr - t(target)) ** 2).sum()
    loss.backward()
    g = r.grad.cpu().numpy()[0]
    scale = np.abs(g).max()

    def fd(i, c, h):
        p = rbf.astype(np.float64)
        rp, rm = rbf.copy(), rbf.copy()
        rp[0, i, c] = np.float32(p[0, i, c] + h)
        rm[0, i, c] = np.float32(p[0, i, c] - h)
        step = float(rp[0, i, c]) - float(rm[0, i, c])
        dalb = (albedo64(rp) - albedo64(rm)) / step
        # linear render: dL = 2 <resid, R dalb>; R of a small direction is
        # formed at unit scale so its FP32 input rounding stays relative
        m = np.abs(dalb).max()
        return 2.0 * np.sum(resid * render64(dalb / m)) * m if m > 0 else 0.0

    worst = 0.0
    for i in range(0, N, 3):
        for c in range(5):
            h = 2e-3 * float(rbf[0, i, 3])
            f1, f2 = fd(i, c, h), fd(i, c, h / 2)
            err_fd = abs(f1 - f2)
            err = abs(f2 - g[i, c])
            worst = max(worst, err / scale)
            assert err <= 1e-5 * scale + 4 * err_fd, (i, c, f2, g[i, c], err_fd)
    report(f"splat -> render -> L2 chain, 12 RBFs: GPU gradient vs FP64 central differences: "
           f"max error {worst:.2e} of max |dL/dtheta|")


def test_exhaustive_render_repeats_bit_for_bit(report):
    """R2 adds each wave's voxels into private LDS histograms of 64-bit
    fixed-point sums; integer addition does not depend on the order of the
    atomics, so every run gives the same bits."""
    from totrilib_amd.render.volume import VolumeRenderExhaustive
    from totrilib_amd.types import BinDef, VolumeDef
    rng = np.random.default_rng(58)
    B, T, S, L = 1, 512, 16, 4
    res = (40, 40, 40)
    vol = sparse_albedo(res, 0.2, 59)[None]
    scan = np.concatenate([rng.uniform(-1, 1, (B, 2, S)), np.zeros((B, 1, S))], 1).astype(np.float32)
    laser = np.concatenate([rng.uniform(-1, 1, (B, 2, L)), np.zeros((B, 1, L))], 1).astype(np.float32)
    vd, bd = VolumeDef(VS, VE, res), BinDef(0.01, 0.0, T)
    outs = [VolumeRenderExhaustive.apply(t(vol), t(scan), t(laser), vd, bd) for _ in range(4)]
    same = all(torch.equal(outs[0], o) for o in outs[1:])
    report(f"R2 exhaustive render, 4 runs at 40^3 (20% occupied), 16 x 4 points, 512 bins: "
           f"{'bit-identical' if same else 'differ in the last bits'}")
    assert same


def test_exhaustive_render_with_infinite_occupancy():
    """A batch whose occupancy holds an infinity takes float histograms, so
    inf and NaN reach the same bins as in the reference; the other batch
    keeps the fixed-point sums."""
    from totrilib_amd.render.volume import VolumeRenderExhaustive
    from totrilib_amd.types import BinDef, VolumeDef
    rng = np.random.default_rng(61)
    B, T, S, L = 2, 300, 4, 2
    res = (16, 14, 18)
    vol = np.stack([sparse_albedo(res, 0.1, 62 + b) for b in range(B)])
    vol[0, 5, 6, 7] = np.inf
    scan = np.concatenate([rng.uniform(-1, 1, (B, 2, S)), np.zeros((B, 1, S))], 1).astype(np.float32)
    laser = np.concatenate([rng.uniform(-1, 1, (B, 2, L)), np.zeros((B, 1, L))], 1).astype(np.float32)
    vd, bd = VolumeDef(VS, VE, res), BinDef(0.01, 0.1, T)
    got = VolumeRenderExhaustive.apply(t(vol), t(scan), t(laser), vd, bd).cpu().numpy()
    ref = O.r2_forward(vol, scan, laser, T, VS, VE, 0.01, 0.1, mode=O.FP32)
    assert np.array_equal(np.isnan(got), np.isnan(ref))
    assert np.array_equal(np.isinf(got), np.isinf(ref)) and np.isinf(ref[0]).any()
    fin = np.isfinite(ref)
    assert np.array_equal(np.sign(got[fin]), np.sign(ref[fin]))
    assert rel_l2(got[fin], ref[fin]) < 1e-6


def test_long_transients_beyond_the_lds_stage():
    """Rows longer than the 8192-float LDS stage (T = 9000) are read from
    global memory: A2 and A3 bit for bit against the FP32 oracle, A5 against
    the FP64-pinned adjoint.  R2's histograms cover 4096 bins, so T = 9000
    takes three bin tiles (4096, 4096, 808)."""
    from totrilib_amd.reconstruct import (BackprojectionConfocalVelten,
                                          BackprojectionExhaustiveVelten)
    from totrilib_amd.types import BinDef, VolumeDef
    rng = np.random.default_rng(60)
    T, S, L = 9000, 5, 2
    res = (12, 10, 14)
    w = 5.0 / T
    vd, bd = VolumeDef(VS, VE, res), BinDef(w, 0.0, T)
    scan = np.concatenate([rng.uniform(-1, 1, (1, S, 2)), np.zeros((1, S, 1))], 2).astype(np.float32)
    laser = np.concatenate([rng.uniform(-1, 1, (1, L, 2)), np.zeros((1, L, 1))], 2).astype(np.float32)
    tr2 = rng.random((1, T, S)).astype(np.float32)
    got = BackprojectionConfocalVelten.apply(t(tr2), t(scan), vd, bd).cpu().numpy()
    ref = O.velten_confocal(tr2, scan, None, res, VS, VE, w, 0.0, mode=O.FP32)
    assert np.count_nonzero(ref) > 100 and np.array_equal(got.astype(np.float64), ref)
    tr3 = rng.random((1, T, S, L)).astype(np.float32)
    got = BackprojectionExhaustiveVelten.apply(t(tr3), t(scan), t(laser), vd, bd).cpu().numpy()
    ref = O.velten_exhaustive(tr3, scan, laser, None, None, res, VS, VE, w, 0.0, mode=O.FP32)
    assert np.count_nonzero(ref) > 100 and np.array_equal(got.astype(np.float64), ref)
    from totrilib_amd.totri_cpp.render import volume_render_exhaustive_grad_volume
    sc3, la3 = scan.transpose(0, 2, 1).copy(), laser.transpose(0, 2, 1).copy()
    G = rng.standard_normal((1, T, S, L)).astype(np.float32)
    grad = torch.empty((1,) + res, device=dev())
    none = torch.empty((1, 3, 0), device=dev())
    volume_render_exhaustive_grad_volume(t(G), t(sc3), t(la3), none, none, grad, vd.start,
                                         vd.end, w, 0.0)
    adj = O.r2_adjoint(G, res, sc3, la3, VS, VE, w, 0.0, mode=O.PINNED64)
    assert np.count_nonzero(adj) > 100
    # each voxel sums spans of dozens of 0.56 mm bins over 10 point pairs in FP32
    assert rel_l2(grad.cpu().numpy(), adj) < 1e-5
    from totrilib_amd.render.volume import VolumeRenderExhaustive
    vol = sparse_albedo(res, 0.3, 61)[None]
    tr = VolumeRenderExhaustive.apply(t(vol), t(sc3), t(la3), vd, bd).cpu().numpy()
    ref = O.r2_forward(vol, sc3, la3, T, VS, VE, w, 0.0, mode=O.PINNED64)
    # nonzero bins in every tile, and the same support as the oracle's
    assert all(np.count_nonzero(ref[:, a:b]) > 20 for a, b in ((0, 4096), (4096, 8192), (8192, T)))
    assert np.array_equal(tr != 0, ref != 0)
    assert rel_l2(tr, ref) < 1e-6


@pytest.mark.parametrize("case", ["binned", "overflow"])
def test_truncated_splat_brick_lists(case, report):
    """Truncated splats list each brick's RBFs first (one pass into the
    brick's list slots, then an index-order sort per brick).  "binned": small
    supports, every brick takes its list.  "overflow": wide supports put more
    RBFs into most bricks than the bricks have slots (512 here), and these
    bricks scan all RBFs; a few narrow RBFs keep some bricks on their lists.
    Both against the FP32 oracle, zeros included."""
    from totrilib_amd.types import VolumeDef
    from totrilib_amd.util import Rbf2VolumeGaussian
    rng = np.random.default_rng(62 if case == "binned" else 63)
    res = (36, 40, 28)
    if case == "binned":
        N = 600
        rbf = np.concatenate([rng.uniform(-1, 1, (1, N, 2)), rng.uniform(0.5, 2.5, (1, N, 1)),
                              rng.uniform(0.01, 0.06, (1, N, 1)), rng.random((1, N, 1))], 2)
    else:
        # 3000 wide RBFs near x = -1 overfill the bricks of that side (68 of
        # 180 hold more than 2048); 300 narrow ones spread over the box
        wide = np.concatenate([rng.uniform(-1, -0.8, (1, 3000, 1)),
                               rng.uniform(-1, 1, (1, 3000, 1)),
                               rng.uniform(0.5, 2.5, (1, 3000, 1)),
                               rng.uniform(0.4, 0.5, (1, 3000, 1)), rng.random((1, 3000, 1))], 2)
        narrow = np.concatenate([rng.uniform(-1, 1, (1, 300, 2)), rng.uniform(0.5, 2.5, (1, 300, 1)),
                                 rng.uniform(0.01, 0.04, (1, 300, 1)), rng.random((1, 300, 1))], 2)
        rbf = np.concatenate([wide, narrow], 1)
        N = rbf.shape[1]
    rbf = rbf.astype(np.float32)
    vol = Rbf2VolumeGaussian.apply(t(rbf), VolumeDef(VS, VE, res), 3.0).cpu().numpy()
    ref = O.splat_forward(rbf, res, VS, VE, trunc=3.0, mode=O.FP32)
    e = rel_l2(vol, ref)
    report(f"truncated splat ({case}, {N} RBFs -> {res}): rel-L2 vs FP32 oracle {e:.2e}")
    assert np.array_equal(vol[:, 0] == 0, ref[:, 0] == 0)
    assert e < 1e-6
