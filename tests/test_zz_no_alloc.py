"""The allocation-free Velten backprojection against the entry point that
takes a workspace, through the C ABI: same kernels, so identical bits, on two
calls in a row (the second smaller, so it reuses the scratch the first one
left), with and without a scan origin; and a transient beyond the scratch is
rejected with a clear error before any device work."""
import ctypes

import numpy as np
import pytest
import torch

from tests.cases import wall_grid

pytestmark = pytest.mark.gpu

VS, VE = (-1.0, -1.0, 0.5), (1.0, 1.0, 2.5)


def dev():
    return torch.device("cuda:0")


def run(lib, fn, tr, scan, so, res, w, off, ws=None):
    from totrilib_amd import _lib
    B, T, S = tr.shape
    D, H, W = res
    out = torch.full((B, D, H, W), float("nan"), device=dev())
    st = torch.cuda.current_stream().cuda_stream
    so_ptr = so.data_ptr() if so is not None else None
    So = so.shape[1] if so is not None else 0
    args = [tr.data_ptr(), B, T, S, scan.data_ptr(), so_ptr, So, out.data_ptr(), D, H, W,
            _lib.f3(VS), _lib.f3(VE), ctypes.c_float(w), ctypes.c_float(off)]
    if ws is not None:
        args += [ws.data_ptr(), ws.numel()]
    rc = getattr(lib, fn)(*args, st)
    _lib.check(rc, fn)
    torch.cuda.synchronize()
    return out


def case(seed, B, n, T, res, origin):
    rng = np.random.default_rng(seed)
    S = n * n
    scan = torch.from_numpy(np.ascontiguousarray(
        np.broadcast_to(wall_grid(n).T[None], (B, S, 3)))).to(dev())
    tr = torch.from_numpy(rng.random((B, T, S)).astype(np.float32)).to(dev())
    so = None
    if origin:
        so = torch.from_numpy(np.ascontiguousarray(np.broadcast_to(
            np.array([0.1, -0.2, -1.0], np.float32), (B, S, 3)))).to(dev())
    return tr, scan, so, res


@pytest.mark.parametrize("origin", [False, True])
def test_noalloc_matches_workspace_entry_point(origin):
    from totrilib_amd import _lib
    lib = _lib.load()
    assert lib.totri_backprojection_confocal_velten_noalloc_max_bytes() == 64 << 20
    # a 32x32 scan with 512 bins (2 MiB of scratch) into 64^3, then a smaller
    # 2-batch 16x16 scan with 300 bins into 40x36x44
    for k, (B, n, T, res, w, off) in enumerate([(1, 32, 512, (64, 64, 64), 0.01, 0.0),
                                                (2, 16, 300, (40, 36, 44), 0.015, 0.2)]):
        tr, scan, so, res = case(70 + k, B, n, T, res, origin)
        print(f"call {k}: B={B} S={n * n} T={T} res={res} origin={origin}", flush=True)
        need = lib.totri_backprojection_confocal_velten_workspace(B, T, n * n)
        ws = torch.empty(need, dtype=torch.uint8, device=dev())
        want = run(lib, "totri_backprojection_confocal_velten_forward", tr, scan, so, res, w,
                   off, ws)
        got = run(lib, "totri_backprojection_confocal_velten_forward_noalloc", tr, scan, so,
                  res, w, off)
        assert torch.count_nonzero(want) > 1000
        assert torch.equal(got, want), float((got - want).abs().max())


def test_noalloc_rejects_transients_beyond_its_scratch():
    from totrilib_amd import _lib
    lib = _lib.load()
    # 1 x 4096 bins x 4097 points = 67 Mbytes > 64 MiB: rejected before any launch
    o = _lib.f3(VS)
    rc = lib.totri_backprojection_confocal_velten_forward_noalloc(
        None, 1, 4096, 4097, None, None, 0, None, 8, 8, 8, o, o, ctypes.c_float(0.01),
        ctypes.c_float(0.0), None)
    assert rc == _lib.TOTRI_EINVAL
    assert "exceeds" in _lib.last_error() and "scratch" in _lib.last_error()
