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


@pytest.mark.parametrize("trunc", [0.0, 3.0])
def test_splat_forward_and_grad_match_oracle(trunc):
    from totrilib_amd.types import VolumeDef
    from totrilib_amd.util import Rbf2VolumeGaussian
    rng = np.random.default_rng(56)
    N = 300
    res = (40, 36, 32)
    rbf = np.concatenate([rng.uniform(-1, 1, (1, N, 2)), rng.uniform(0.5, 2.5, (1, N, 1)),
                          rng.uniform(0.02, 0.1, (1, N, 1)), rng.random((1, N, 1))], 2).astype(np.float32)
    r = t(rbf).requires_grad_(True)
    vol = Rbf2VolumeGaussian.apply(r, VolumeDef(VS, VE, res), trunc)
    got = vol.detach().cpu().numpy()
    ref = O.splat_forward(rbf, res, VS, VE, trunc=trunc, mode=O.FP32)
    assert rel_l2(got, ref) < 1e-5
    G = rng.standard_normal(got.shape).astype(np.float32)
    (vol * t(G)).sum().backward()
    gref = O.splat_grad(rbf, ref, G, VS, VE, trunc=trunc, mode=O.PINNED64)
    e = rel_l2(r.grad.cpu().numpy(), gref)
    print(f"trunc={trunc}: splat rel-L2 {rel_l2(got, ref):.2e}, grad rel-L2 {e:.2e}")
    assert e < 1e-4


def test_splat_render_chain_gradient_matches_fd():
    """RBF weights, centres and sigma through splat (albedo = V0 * V1), the
    confocal render and an L2 loss: the GPU gradient against central finite
    differences of the FP64 oracle chain (6-sigma truncation keeps the splat
    smooth)."""
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

    def chain_np(r32):
        v = O.splat_forward(r32, (R, R, R), vd.start, vd.end, trunc=6.0, mode=O.PURE64)[0]
        alb = (v[0] * v[1]).transpose(2, 1, 0)  # [D,H,W] -> [X,Y,Z]
        tr = O.r1_forward(alb.astype(np.float32), mp, T, box_o, box_s, w, 0.0, mode=O.PINNED64)
        return tr

    r = t(rbf).requires_grad_(True)
    vol = Rbf2VolumeGaussian.apply(r, vd, 6.0)[0]
    alb = (vol[0] * vol[1]).permute(2, 1, 0).contiguous()
    tr = VolumeRenderConfocal.apply(alb, t(mp), box_o, box_s, T, w, 0.0)
    loss = ((tr - t(target)) ** 2).sum()
    loss.backward()
    g = r.grad.cpu().numpy()[0]
    picks = [(i, c) for i in range(0, N, 3) for c in range(5)]
    for i, c in picks:
        h = 1e-3 if c != 3 else 2e-4
        rp, rm = rbf.copy(), rbf.copy()
        rp[0, i, c] += h
        rm[0, i, c] -= h
        step = float(rp[0, i, c]) - float(rm[0, i, c])
        lp = np.sum((chain_np(rp) - target) ** 2)
        lm = np.sum((chain_np(rm) - target) ** 2)
        fd = (lp - lm) / step
        scale = np.abs(g).max()
        assert abs(fd - g[i, c]) < 2e-2 * scale, (i, c, fd, g[i, c])
