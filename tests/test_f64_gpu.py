"""float64 tensors: the render R1, its adjoint A1 (exact transpose and the
reference's window form) and the confocal Velten backprojection A2 in the
reference's double semantics, bit for bit against the oracle's restatement
of that dispatch (same double operations in the same order), through the
drop-in bindings and autograd Functions."""
import numpy as np
import pytest
import torch

from oracle import oracle as O
from tests.cases import ORIGIN, SIZE, wall_grid

pytestmark = pytest.mark.gpu


def dev():
    return torch.device("cuda:0")


def t64(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).to(dev())


def setup(seed, shape=(12, 10, 16), n=4):
    rng = np.random.default_rng(seed)
    vol = rng.random(shape)
    mp = wall_grid(n).astype(np.float64)
    mp[:2] += rng.uniform(-0.05, 0.05, mp[:2].shape)
    return rng, vol, mp


@pytest.mark.parametrize("off", [0.0, 0.02])
def test_render_and_adjoints_f64_bit_exact(off, report):
    from totrilib_amd.render import VolumeRenderConfocal
    from totrilib_amd.totri_cpp.render import volume_confocal_grad_volume
    rng, vol, mp = setup(81)
    X, Y, Z = vol.shape
    T, w = 320, 0.011
    v = t64(vol).requires_grad_(True)
    m = t64(mp)
    tr = VolumeRenderConfocal.apply(v, m, ORIGIN, SIZE, T, w, off)
    assert tr.dtype == torch.float64
    ref = O.r1_forward_f64(vol, mp, T, ORIGIN, SIZE, w, off)
    assert np.count_nonzero(ref) > 1000
    assert np.array_equal(tr.detach().cpu().numpy(), ref)
    G = rng.standard_normal((mp.shape[1], T))
    (tr * t64(G)).sum().backward()
    at = O.a1_transpose_f64(G, mp, X, Y, Z, ORIGIN, SIZE, w, off)
    assert np.array_equal(v.grad.cpu().numpy(), at)
    win = torch.empty(vol.shape, dtype=torch.float64, device=dev())
    volume_confocal_grad_volume(win, m, t64(G), ORIGIN, SIZE, w, off, reference_window=True)
    aw = O.a1_reference_f64(G, mp, X, Y, Z, ORIGIN, SIZE, w, off)
    assert np.array_equal(win.cpu().numpy(), aw)
    lhs, rhs = float(np.sum(ref * G)), float(np.sum(vol * at))
    report(f"float64 R1/A1 (off {off}): bit-exact vs the oracle's double dispatch; "
           f"<Rv,g> - <v,A g> = {abs(lhs - rhs) / abs(lhs):.1e} relative")


def test_velten_f64_bit_exact():
    from totrilib_amd.reconstruct import BackprojectionConfocalVelten
    from totrilib_amd.types import BinDef, VolumeDef
    rng = np.random.default_rng(82)
    B, T, S = 2, 400, 9
    res = (14, 12, 13)
    vs, ve = (-1.0, -1.0, 0.5), (1.0, 1.0, 2.5)
    tr = rng.random((B, T, S))
    scan = np.concatenate([rng.uniform(-1, 1, (B, S, 2)), np.zeros((B, S, 1))], 2)
    so = scan + rng.uniform(-0.3, 0.3, scan.shape) - [0, 0, 0.8]
    for org in (None, so):
        got = BackprojectionConfocalVelten.apply(
            t64(tr), t64(scan), VolumeDef(vs, ve, res), BinDef(0.012, 0.05, T),
            None if org is None else t64(org))
        assert got.dtype == torch.float64
        ref = O.velten_confocal_f64(tr, scan, org, res, vs, ve, 0.012, 0.05)
        assert np.count_nonzero(ref) > 1000
        assert np.array_equal(got.cpu().numpy(), ref)


def test_f64_mixed_dtypes_and_scan_point_gradient_raise():
    from totrilib_amd.render import VolumeRenderConfocal
    from totrilib_amd.totri_cpp.render import volume_confocal_forward
    _, vol, mp = setup(83, (6, 6, 8), 2)
    with pytest.raises(TypeError):
        volume_confocal_forward(t64(vol), t64(mp).float(),
                                torch.empty(4, 50, dtype=torch.float64, device=dev()),
                                ORIGIN, SIZE, 0.05, 0.0)
    m = t64(mp).requires_grad_(True)
    tr = VolumeRenderConfocal.apply(t64(vol), m, ORIGIN, SIZE, 50, 0.05, 0.0)
    with pytest.raises(TypeError):
        tr.sum().backward()
