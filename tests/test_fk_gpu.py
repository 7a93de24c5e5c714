"""f-k migration on the GPU (totrilib_amd.reconstruct.FkMigration and
StoltInterpolation, reference totri/reconstruct/fk_migration.py) against the
float64 oracle (oracle/fk.py) and a plain PyTorch fp32 composition of the
reference's own operations (fftn, fftshift, grid_sample, weights, ifftn).

Tolerance: the reference computes its sample grid in float32, so its own
result differs from the float64 oracle by an amount that grows with the
depth size (about 8e-5 relative L2 at 2D = 2048, 2.4e-4 at 4096).  The HIP
result must lie within 1e-5 + 2x that float32 error of the oracle, in
relative L2 and in max-abs relative to the volume's maximum."""
import numpy as np
import pytest
import torch

from oracle import fk
from tests.test_oracle_fk import point_transient

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")


def rel_l2(a, b):
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-30))


def torch_fp32_fk(t, depth, height, width):
    """FkMigration.forward as the reference writes it, fp32 torch on the GPU,
    for one transient t [D, H, W] -> [1, D, H, W]."""
    from totrilib_amd.reconstruct.fk_migration import StoltInterpolation
    t = torch.as_tensor(t, device=DEV)
    D, H, W = t.shape
    gz = torch.linspace(0, 1, D, device=DEV).view(D, 1, 1)
    p = torch.cat((torch.sqrt(t) * gz, torch.zeros_like(t)), 0).to(torch.complex64)
    f = torch.fft.fftshift(torch.fft.fftn(p))
    grid, weights = StoltInterpolation(depth, height, width).get_buffers(D, H, W, DEV)
    fr = torch.view_as_real(f)
    parts = [torch.nn.functional.grid_sample(fr[None, None, ..., i].contiguous(), grid,
                                             mode="bilinear", padding_mode="zeros",
                                             align_corners=False) for i in (0, 1)]
    g = torch.complex(parts[0], parts[1])[0, 0] * weights[0]
    v = torch.fft.ifftn(torch.fft.ifftshift(g))
    return v[:D].abs()[None].cpu().numpy()


def check_volume(got, want, ref32):
    assert got.shape == want.shape == ref32.shape
    scale = np.abs(want).max()
    tol_l2 = 1e-5 + 2 * rel_l2(ref32, want)
    tol_max = 1e-5 + 2 * np.abs(ref32 - want).max() / scale
    assert rel_l2(got, want) <= tol_l2
    assert np.abs(got - want).max() / scale <= tol_max


def check_fk(t, got, depth, height, width):
    check_volume(got, fk.fk_migration(t, depth, height, width),
                 torch_fp32_fk(t[0], depth, height, width))


@pytest.mark.parametrize("N,D,H,W,depth,height,width", [
    (1, 64, 16, 16, 0.64, 1.0, 1.0),
    (2, 96, 17, 15, 1.2, 0.8, 1.3),      # odd wall sizes; only entry 0 is returned
    (1, 1, 4, 4, 0.01, 1.0, 1.0),        # D = 1: the ramp is all zero
    (1, 200, 8, 40, 3.0, 0.5, 2.0),
    (1, 128, 1, 1, 1.0, 1.0, 1.0),
])
def test_fk_migration_matches_oracle(N, D, H, W, depth, height, width):
    from totrilib_amd.reconstruct import FkMigration
    print(f"FkMigration N={N} D={D} H={H} W={W}", flush=True)
    rng = np.random.default_rng(D * 7 + H)
    t = rng.random((N, D, H, W)).astype(np.float32)
    got = FkMigration(depth, height, width)(torch.from_numpy(t).to(DEV)).cpu().numpy()
    if D == 1:
        assert np.all(got == 0)
        return
    check_fk(t, got, depth, height, width)


@pytest.mark.parametrize("D,H,W", [(2048, 8, 16), (32, 128, 128), (16, 2, 64), (8, 64, 2),
                                   (64, 256, 64)])
def test_fk_native_fft_shapes_match_oracle(D, H, W):
    """Power-of-two sizes at the limits of the native FFT path: 2D = 4096
    (four columns per depth workgroup), H * W = 16384 slices, flat slices."""
    from totrilib_amd import _lib
    from totrilib_amd.reconstruct import FkMigration
    print(f"native FkMigration D={D} H={H} W={W}", flush=True)
    assert _lib.load().totri_fk_migration_workspace(D, H, W) > 0
    rng = np.random.default_rng(D + H + W)
    t = rng.random((1, D, H, W)).astype(np.float32)
    got = FkMigration(D * 0.004, 1.5, 1.2)(torch.from_numpy(t).to(DEV)).cpu().numpy()
    check_fk(t, got, D * 0.004, 1.5, 1.2)


def test_fk_native_and_hipfft_paths_agree():
    from totrilib_amd import _lib
    from totrilib_amd.reconstruct import FkMigration
    rng = np.random.default_rng(77)
    t = torch.from_numpy(rng.random((256, 32, 32)).astype(np.float32)).to(DEV)
    m = FkMigration(256 * 0.005, 1.0, 1.0)
    native = m(t[None])
    lib = _lib.load()
    via_hipfft = torch.empty_like(native)
    with torch.cuda.device(DEV):
        m._torch_fft(lib, t, via_hipfft, DEV, torch.cuda.current_stream(DEV).cuda_stream)
    a, b = native.cpu().numpy(), via_hipfft.cpu().numpy()
    assert rel_l2(a, b) <= 1e-5


def test_fk_point_scatterer_on_gpu():
    from totrilib_amd.reconstruct import FkMigration
    D, H, W, bw, wall, p = 160, 27, 33, 0.01, 1.0, (-0.25, 0.15, 0.5)
    t, xs, ys = point_transient(D, H, W, wall, bw, p)
    t = t.astype(np.float32)
    got = FkMigration(D * bw, wall, wall)(torch.from_numpy(t).to(DEV)).cpu().numpy()
    check_fk(t, got, D * bw, wall, wall)
    d, h, w = np.unravel_index(np.argmax(got[0]), got[0].shape)
    assert abs(d - 2 * p[2] / bw) <= 1.5
    assert abs(h - np.argmin(abs(ys - p[1]))) <= 1
    assert abs(w - np.argmin(abs(xs - p[0]))) <= 1


def test_fk_migration_full_size():
    """The bench's f-k workload: 1024 bins on a 64 x 64 wall."""
    from totrilib_amd.reconstruct import FkMigration
    rng = np.random.default_rng(1024)
    t = rng.random((1, 1024, 64, 64)).astype(np.float32)
    got = FkMigration(1024 * 0.0025, 2.0, 2.0)(torch.from_numpy(t).to(DEV)).cpu().numpy()
    check_fk(t, got, 1024 * 0.0025, 2.0, 2.0)


@pytest.mark.parametrize("N,C,D2,H,W", [(2, 2, 32, 9, 12), (1, 1, 64, 16, 16)])
def test_stolt_matches_oracle_and_grid_sample(N, C, D2, H, W):
    from totrilib_amd.reconstruct.fk_migration import StoltInterpolation
    rng = np.random.default_rng(N + C + D2)
    f = (rng.standard_normal((N, C, D2, H, W))
         + 1j * rng.standard_normal((N, C, D2, H, W))).astype(np.complex64)
    s = StoltInterpolation(0.9, 0.7, 1.1)
    got = s(torch.from_numpy(f).to(DEV)).cpu().numpy()
    want = fk.stolt(f.astype(np.complex128), 0.9, 0.7, 1.1)
    assert rel_l2(got, want) <= 1e-5
    # torch fp32: grid_sample on the real and imaginary parts, then the weights
    grid, weights = s.get_buffers(D2 // 2, H, W, DEV)
    ft = torch.view_as_real(torch.from_numpy(f).to(DEV).view(N * C, D2, H, W))
    parts = [torch.nn.functional.grid_sample(ft[None, ..., i], grid, mode="bilinear",
                                             padding_mode="zeros", align_corners=False)
             for i in (0, 1)]
    ref = (torch.complex(parts[0], parts[1])[0] * weights[0]).view(N, C, D2, H, W)
    assert rel_l2(got, ref.cpu().numpy()) <= 1e-5
    assert np.all(got[:, :, :D2 // 2] == 0)


def test_stolt_rejects_bad_input():
    from totrilib_amd.reconstruct.fk_migration import StoltInterpolation
    s = StoltInterpolation(1.0, 1.0, 1.0)
    with pytest.raises(ValueError):
        s(torch.zeros((1, 1, 7, 4, 4), dtype=torch.complex64, device=DEV))
    with pytest.raises(ValueError):
        s(torch.zeros((1, 1, 8, 4, 4), device=DEV))
    with pytest.raises(ValueError):
        StoltInterpolation(1.0, 1.0, 0.0)(torch.zeros((1, 1, 8, 4, 4), dtype=torch.complex64,
                                                      device=DEV))


@pytest.mark.parametrize("D,H,W", [(16, 6, 8), (24, 5, 7)])
def test_stolt_spectrum_layouts_agree(D, H, W):
    """The three spectrum layouts of the Stolt entry point: fftshift order,
    raw fftn order and one-sided rfftn order give the same interpolation."""
    from totrilib_amd import _lib
    from totrilib_amd.reconstruct.fk_migration import _stolt
    rng = np.random.default_rng(D + W)
    p = rng.standard_normal((2 * D, H, W)).astype(np.float32)
    full = np.fft.fftn(p.astype(np.float64))
    want = np.fft.ifftshift(fk.stolt(np.fft.fftshift(full)[None, None], 0.8, 0.6, 0.7)[0, 0])
    pt = torch.from_numpy(p).to(DEV)
    outs = []
    for layout, spec in ((_lib.TOTRI_FK_UNSHIFTED, torch.fft.fftn(pt)),
                         (_lib.TOTRI_FK_UNSHIFTED_ONESIDED, torch.fft.rfftn(pt))):
        for zeroed in (False, True):
            out = torch.zeros((2 * D, H, W), dtype=torch.complex64, device=DEV)
            _stolt(spec.contiguous()[None], out, 0.8, 0.6, 0.7, layout, zeroed)
            outs.append(out.cpu().numpy())
    for o in outs:
        assert rel_l2(o, want) <= 1e-5
    assert np.array_equal(outs[0], outs[1]) and np.array_equal(outs[2], outs[3])


def test_fk_repeated_calls_and_shape_changes():
    """Buffers kept between calls: results repeat exactly, also after a call
    of another shape in between."""
    from totrilib_amd.reconstruct import FkMigration
    rng = np.random.default_rng(3)
    a = torch.from_numpy(rng.random((1, 48, 12, 10)).astype(np.float32)).to(DEV)
    b = torch.from_numpy(rng.random((1, 32, 8, 8)).astype(np.float32)).to(DEV)
    m = FkMigration(0.5, 0.9, 0.9)
    r1 = m(a).cpu().numpy()
    r2 = m(a).cpu().numpy()
    m(b)
    r3 = m(a).cpu().numpy()
    assert np.array_equal(r1, r2) and np.array_equal(r1, r3)
    check_fk(a.cpu().numpy(), r1, 0.5, 0.9, 0.9)
