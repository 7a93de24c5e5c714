"""Full-size GPU checks of the benchmark configurations against the oracle.

  64^3 / 32x32 scan / 512 bins     render, noise, backproject (A1 and Velten)
                                   with the oracle at full size
  128^3 / 64x64 scan / 1024 bins   albedo gradient of an L2 loss against FP64
                                   central differences (h = 1e-4) on 64 voxels
  256^3 / 256x256 scan / 2048 bins FP32, BF16 and FP16 transient storage: error
                                   of render and backprojection against the
                                   FP64 oracle on seeded subsets
  20,000 RBFs -> 128^3 (3 sigma)   splat, render, L2 loss, gradient to the RBF
                                   parameters against the FP64 analytic
                                   gradient of the same truncated splat

Every error is printed; run with -s to see the table.
"""
import os

import numpy as np
import pytest
import torch

from oracle import oracle as O
from tests.cases import ORIGIN, SIZE, rel_l2, sparse_albedo, wall_grid

pytestmark = pytest.mark.gpu

TOL = 1e-5


def dev():
    return torch.device("cuda:0")


def t(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to(dev())


@pytest.fixture(autouse=True, scope="module")
def oracle_threads():
    O.set_threads(min(os.cpu_count() or 1, 16))


def render(vol, mp, T, w, dtype=None):
    from totrilib_amd.render import VolumeRenderConfocal
    args = (vol, mp, ORIGIN, SIZE, T, w, 0.0)
    with torch.no_grad():
        out = VolumeRenderConfocal.apply(*args) if dtype is None else \
            VolumeRenderConfocal.apply(*args, dtype)
    torch.cuda.synchronize()
    return out


def test_64_cubed_render_noise_backproject(report):
    from totrilib_amd.reconstruct import BackprojectionConfocalVelten
    from totrilib_amd.render import VolumeRenderConfocal
    from totrilib_amd.types import BinDef, VolumeDef
    n, T, R, w = 32, 512, 64, 0.005
    mp_np = wall_grid(n)
    vol_np = sparse_albedo((R, R, R), 0.05, 0)
    tr = render(t(vol_np), t(mp_np), T, w).cpu().numpy().astype(np.float64)
    ref = O.r1_forward(vol_np, mp_np, T, ORIGIN, SIZE, w, 0.0, mode=O.PINNED64)
    e_r = rel_l2(tr, ref)
    g = torch.Generator().manual_seed(100)
    noisy = (tr + 1e-3 * tr.max() * torch.randn(tr.shape, generator=g, dtype=torch.float64)
             .numpy()).astype(np.float32)
    bp = VolumeRenderConfocal.apply_transposed(t(noisy), t(mp_np), ORIGIN, SIZE, (R, R, R),
                                               w, 0.0).cpu().numpy()
    bref = O.a1_transpose(noisy, mp_np, R, R, R, ORIGIN, SIZE, w, 0.0, mode=O.PINNED64)
    e_a = rel_l2(bp, bref)
    # Velten: inclusive grid over the same box, round-trip bins of 0.01 m
    vd, bd = VolumeDef((-1, -1, 0.5), (1, 1, 2.5), (R, R, R)), BinDef(2 * w, 0.0, T)
    trT = noisy.T[None].copy()          # [1, T, S]
    scan = mp_np.T[None].copy()         # [1, S, 3]
    vel = BackprojectionConfocalVelten.apply(t(trT), t(scan), vd, bd).cpu().numpy()
    vref = O.velten_confocal(trT, scan, None, (R, R, R), vd.start, vd.end, 2 * w, 0.0,
                             mode=O.FP32)
    v64 = O.velten_confocal(trT, scan, None, (R, R, R), vd.start, vd.end, 2 * w, 0.0,
                            mode=O.PINNED64)
    report(f"64^3 render rel-L2 vs FP64 {e_r:.2e}; A1 backprojection {e_a:.2e}; "
          f"Velten bit-exact vs FP32 oracle: {np.array_equal(vel, vref)}, "
          f"rel-L2 vs FP64 {rel_l2(vel, v64):.2e}")
    assert e_r < TOL and e_a < TOL
    assert np.array_equal(vel.astype(np.float64), vref)


def test_128_cubed_albedo_gradient_matches_central_differences(report):
    """L(v) = ||R1 v - R1 v_target||^2.  R1 is linear, so with r = R1 v - y
    and a_j = R1 e_j (the FP64 oracle's render of voxel j alone),
    L(v +- h e_j) = ||r +- h a_j||^2 exactly; the central difference of
    those two FP64 losses is compared with the GPU gradient."""
    from totrilib_amd.render import VolumeRenderConfocal
    n, T, R, w, h = 64, 1024, 128, 0.0025, 1e-4
    mp_np = wall_grid(n)
    v_np = sparse_albedo((R, R, R), 0.03, 1)
    y = render(t(sparse_albedo((R, R, R), 0.03, 2)), t(mp_np), T, w)
    v = t(v_np).requires_grad_(True)
    tr = VolumeRenderConfocal.apply(v, t(mp_np), ORIGIN, SIZE, T, w, 0.0)
    ((tr - y) ** 2).sum().backward()
    g = v.grad.cpu().numpy()
    r = (tr.detach() - y).cpu().numpy().astype(np.float64)
    rng = np.random.default_rng(3)
    idx = rng.choice(R ** 3, 64, replace=False)
    fd = np.zeros(64)
    gg = np.zeros(64)
    for k, flat in enumerate(idx):
        xi, yi, zi = np.unravel_index(flat, (R, R, R))
        e = np.zeros((R, R, R), np.float32)
        e[xi, yi, zi] = 1.0
        a = O.r1_forward(e, mp_np, T, ORIGIN, SIZE, w, 0.0, mode=O.PINNED64,
                         cols=[[xi, yi]])
        fd[k] = (np.sum((r + h * a) ** 2) - np.sum((r - h * a) ** 2)) / (2 * h)
        gg[k] = g[xi, yi, zi]
    scale = np.abs(fd).max()
    err = np.abs(gg - fd).max() / scale
    report(f"128^3 gradient vs FP64 central differences on 64 voxels: max error "
          f"{err:.2e} of max |dL/dv| = {scale:.3e} ({np.count_nonzero(fd)} nonzero)")
    assert np.count_nonzero(fd) > 32
    assert err < 1e-5


CFG4 = dict(n=256, T=2048, R=256, w=0.00125)


@pytest.fixture(scope="module")
def cfg4_inputs():
    mp_np = wall_grid(CFG4["n"])
    vol_np = sparse_albedo((CFG4["R"],) * 3, 0.01, 5)
    return mp_np, vol_np


def test_256_cubed_render_error_per_storage_dtype(cfg4_inputs, report):
    mp_np, vol_np = cfg4_inputs
    T, w = CFG4["T"], CFG4["w"]
    rows = np.sort(np.random.default_rng(51).choice(mp_np.shape[1], 24, replace=False))
    ref = np.concatenate([O.r1_forward(vol_np, mp_np, T, ORIGIN, SIZE, w, 0.0,
                                       mode=O.PINNED64, i0=int(i), i1=int(i) + 1)
                          for i in rows])
    errs = {}
    vol, mp = t(vol_np), t(mp_np)
    for name, dt in (("fp32", None), ("bf16", torch.bfloat16), ("fp16", torch.float16)):
        out = render(vol, mp, T, w, dt)
        errs[name] = rel_l2(out[torch.from_numpy(rows).to(dev())].float().cpu().numpy(), ref)
        del out
    report("256^3 render rel-L2 vs FP64 (24 scan rows): " +
          ", ".join(f"{k} {v:.2e}" for k, v in errs.items()))
    assert errs["fp32"] < TOL and errs["fp16"] < 1e-3 and errs["bf16"] < 8e-3


def test_256_cubed_backprojection_error_per_storage_dtype(cfg4_inputs, report):
    from totrilib_amd.render import VolumeRenderConfocal
    mp_np, vol_np = cfg4_inputs
    T, w, R = CFG4["T"], CFG4["w"], CFG4["R"]
    mp = t(mp_np)
    tr = render(t(vol_np), mp, T, w)
    g = torch.Generator(device=dev()).manual_seed(52)
    tr += 1e-3 * tr.max() * torch.randn(tr.shape, device=dev(), generator=g)
    rng = np.random.default_rng(53)
    cols = np.stack([rng.integers(0, R, 16), rng.integers(0, R, 16)], axis=1)
    ref = O.a1_transpose(tr.cpu().numpy(), mp_np, R, R, R, ORIGIN, SIZE, w, 0.0,
                         mode=O.PINNED64, cols=cols)[cols[:, 0], cols[:, 1]]
    errs = {}
    for name, dt in (("fp32", torch.float32), ("bf16", torch.bfloat16),
                     ("fp16", torch.float16)):
        vol = VolumeRenderConfocal.apply_transposed(tr.to(dt), mp, ORIGIN, SIZE, (R, R, R),
                                                    w, 0.0)
        ci = torch.from_numpy(cols).to(dev())
        errs[name] = rel_l2(vol[ci[:, 0], ci[:, 1]].cpu().numpy(), ref)
        del vol
    report("256^3 backprojection rel-L2 vs FP64 (16 columns, 4096 voxels): " +
          ", ".join(f"{k} {v:.2e}" for k, v in errs.items()))
    assert errs["fp32"] < TOL and errs["fp16"] < 1e-3 and errs["bf16"] < 8e-3


def test_20000_rbf_splat_render_gradient(report):
    """Splat (3 sigma) -> albedo V0 * V1 -> render -> L2 loss.  The splat is
    checked on 4096 voxels against the FP64 oracle; the RBF gradient against
    the oracle's FP64 analytic gradient of the same truncated splat, given
    the GPU's dL/dV (the albedo gradient, itself checked above)."""
    from totrilib_amd.render import VolumeRenderConfocal
    from totrilib_amd.types import VolumeDef
    from totrilib_amd.util import Rbf2VolumeGaussian
    n, T, R, w = 64, 1024, 128, 0.0025
    rng = np.random.default_rng(4)
    M = 20000
    rbf = np.concatenate([rng.uniform(-1, 1, (1, M, 2)), rng.uniform(0.5, 2.5, (1, M, 1)),
                          rng.uniform(0.01, 0.05, (1, M, 1)), rng.uniform(0, 1, (1, M, 1))],
                         2).astype(np.float32)
    d = 2.0 / R
    vd = VolumeDef((-1 + d / 2, -1 + d / 2, 0.5 + d / 2), (1 - d / 2, 1 - d / 2, 2.5 - d / 2),
                   (R, R, R))
    mp = t(wall_grid(n))
    y = render(t(sparse_albedo((R, R, R), 0.03, 7)), mp, T, w)
    r = t(rbf).requires_grad_(True)
    V = Rbf2VolumeGaussian.apply(r, vd, 3.0)
    V.retain_grad()
    alb = (V[0, 0] * V[0, 1]).permute(2, 1, 0).contiguous()   # [D,H,W] -> [X,Y,Z]
    tr = VolumeRenderConfocal.apply(alb, mp, ORIGIN, SIZE, T, w, 0.0)
    ((tr - y) ** 2).sum().backward()
    torch.cuda.synchronize()
    Vn = V.detach().cpu().numpy().astype(np.float64)
    vox = rng.choice(R ** 3, 4096, replace=False)
    sref = O.splat_forward(rbf, (R, R, R), vd.start, vd.end, trunc=3.0, mode=O.PINNED64,
                           vox=vox)
    e_s = rel_l2(Vn.reshape(1, 2, -1)[:, :, vox], sref.reshape(1, 2, -1)[:, :, vox])
    gref = O.splat_grad(rbf, Vn, V.grad.cpu().numpy().astype(np.float64), vd.start, vd.end,
                        trunc=3.0, mode=O.PINNED64)
    got = r.grad.cpu().numpy()
    e_g = rel_l2(got, gref)
    per = [rel_l2(got[..., c], gref[..., c]) for c in range(5)]
    report(f"20,000 RBFs: splat rel-L2 vs FP64 (4096 voxels) {e_s:.2e}; gradient rel-L2 "
          f"{e_g:.2e} (x {per[0]:.1e}, y {per[1]:.1e}, z {per[2]:.1e}, sigma {per[3]:.1e}, "
          f"weight {per[4]:.1e})")
    assert e_s < TOL
    assert e_g < 1e-5
