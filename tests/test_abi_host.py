"""CPU checks of the C ABI library and the host-side Python layer (no GPU
calls): exported symbols, geometry types, wall grid, origin broadcasting,
the second-derivative filter and argument validation."""
import ctypes
import os
import re

import numpy as np
import pytest
import torch

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "totri_hip.h")


def header_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(totri_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    from totrilib_amd import _lib
    lib = ctypes.CDLL(_lib.LIB_PATH)
    names = header_functions()
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n
    assert sorted(_lib.SIGNATURES) == names
    assert _lib.load().totri_abi_version() >= 100


def test_error_reporting_without_gpu():
    from totrilib_amd import _lib
    lib = _lib.load()
    o = _lib.f3((0, 0, 0))
    rc = lib.totri_volume_confocal_forward(None, -1, 1, 1, None, 1, None, 1, 0, o, o, 1.0, 0.0,
                                           None, 0, None)
    assert rc == _lib.TOTRI_EINVAL
    assert "negative" in _lib.last_error()
    with pytest.raises(ValueError):
        _lib.check(rc, "probe")
    rc = lib.totri_rbf2volume_gaussian_forward(None, 1, 1, 9, None, 1, 1, 1, o, o, 0.0, None, 0,
                                               None)
    assert rc == _lib.TOTRI_EINVAL and "attributes" in _lib.last_error()
    # empty problems succeed without touching the device
    rc = lib.totri_volume_confocal_forward(None, 4, 4, 4, None, 0, None, 8, 0, o, o, 1.0, 0.0,
                                           None, 0, None)
    assert rc == _lib.TOTRI_OK


def test_volume_and_bin_defs():
    from totrilib_amd.types import C0, BinDef, VolumeDef, to_size3, to_vec3
    v = VolumeDef((-1, -1, 0.5), (1, 1, 2.5), (4, 3, 2))
    assert v.extent() == (2.0, 2.0, 2.0)
    c = v.corner_points()
    assert tuple(c.shape) == (3, 8)
    assert c[:, 0].tolist() == [-1, -1, 0.5] and c[:, 7].tolist() == [1, 1, 2.5]
    assert c[:, 1].tolist() == [-1, 1, 0.5] and c[:, 2].tolist() == [-1, -1, 2.5]
    assert v.flat_idx_to_3d_idx(2 * 3 * 2 + 1 * 2 + 1) == (1, 1, 2)
    assert v.index_to_coordinates((1, 1, 2)) == (0.0, -1 + 2 / 3, 1.5)
    g = v.coordinates_inclusive()
    assert tuple(g.shape) == (4, 3, 2, 3)
    assert g[0, 0, 0].tolist() == [-1, -1, 0.5] and g[-1, -1, -1].tolist() == [1, 1, 2.5]
    assert to_vec3(2) == (2, 2, 2) and to_size3(5) == (5, 5, 5)
    with pytest.raises(ValueError):
        to_vec3((1, 2))
    with pytest.raises(ValueError):
        to_size3((1.0, 2, 3))
    b = BinDef(1e-11, 2e-11, 100, unit_is_time=True)
    assert abs(b.width - 1e-11 * C0) < 1e-12 and b.num == 100
    assert abs(b.width_s - 1e-11) < 1e-20


def test_make_wall_grid_orders_x_fastest():
    from totrilib_amd.data import make_wall_grid
    g = make_wall_grid(-1, 1, 4, -2, 2, 3, z_val=0.5, device=torch.device("cpu"))
    assert tuple(g.shape) == (12, 3)
    assert g[:4, 0].tolist() == pytest.approx([-1, -1 / 3, 1 / 3, 1])
    assert g[:4, 1].tolist() == [-2, -2, -2, -2]
    assert torch.all(g[:, 2] == 0.5)
    g2 = make_wall_grid(-1, 1, 4, -2, 2, 3, device=torch.device("cpu"), flattened=False)
    assert tuple(g2.shape) == (3, 4, 3)


def test_origin_of_wallpoints():
    from totrilib_amd.util import origin_of_wallpoints
    w = torch.zeros(2, 5, 3)
    assert tuple(origin_of_wallpoints(None, w, "o").shape) == (2, 0, 3)
    o = torch.arange(6.0).view(2, 3)
    e = origin_of_wallpoints(o, w, "o")
    assert tuple(e.shape) == (2, 5, 3) and torch.equal(e[1, 4], o[1])
    assert tuple(origin_of_wallpoints(torch.zeros(2, 1, 3), w, "o").shape) == (2, 5, 3)
    with pytest.raises(ValueError):
        origin_of_wallpoints(torch.zeros(3), w, "o")


def test_second_derivative_filter():
    from totrilib_amd.reconstruct import second_derivative_filter
    z = torch.arange(8.0)
    # -(z - 3)^2 has second difference -2 everywhere away from the end.
    vol = (-(z - 3) ** 2).view(1, 8, 1, 1).repeat(1, 1, 2, 3)
    out = second_derivative_filter(vol.clone())
    assert torch.allclose(out[0, :6, 0, 0], torch.full((6,), 2.0))
    # convex profile: positive second differences are clipped to 0 (the
    # replicated last slice makes the last two differences concave)
    out2 = second_derivative_filter(((z - 3) ** 2).view(1, 8, 1, 1).clone())
    assert torch.all(out2[0, :6] == 0) and out2[0, 6].item() == 7.0


def test_bindings_reject_cpu_tensors_and_bad_shapes():
    from totrilib_amd.totri_cpp import render, reconstruct, util
    vol = torch.zeros(4, 4, 4)
    mp = torch.zeros(3, 5)
    tr = torch.zeros(5, 16)
    with pytest.raises(NotImplementedError):
        render.volume_confocal_forward(vol, mp, tr, (0, 0, 0), (1, 1, 1), 0.1, 0.0)
    with pytest.raises(NotImplementedError):
        reconstruct.backprojection_confocal_velten_forward(
            torch.zeros(1, 8, 5), torch.zeros(1, 5, 3), torch.zeros(1, 0, 3),
            torch.zeros(1, 2, 2, 2), (0, 0, 0), (1, 1, 1), 0.1, 0.0)
    with pytest.raises(NotImplementedError):
        util.rbf2volume_gaussian_forward(torch.zeros(1, 2, 5), torch.zeros(1, 2, 2, 2, 2),
                                         (0, 0, 0), (1, 1, 1))


def test_install_as_totri_aliases_reference_modules():
    import sys
    import totrilib_amd
    totrilib_amd.install_as_totri()
    from totri.render import VolumeRenderConfocal  # noqa: F401
    from totri.reconstruct import BackprojectionConfocalVelten  # noqa: F401
    from totri.types import BinDef, VolumeDef  # noqa: F401
    from totri.util import Rbf2VolumeGaussian  # noqa: F401
    from totri_cpp.render import volume_confocal_forward  # noqa: F401
    from totri_cpp.util import rbf2volume_gaussian_grad  # noqa: F401
    assert sys.modules["totri_cpp.reconstruct"].backprojection_confocal_velten_forward


def test_r1_mirror_paths_on_the_host():
    """The mirror R1 kernels a geometry allows (host-side plan, no GPU): the
    xy-mirror kernel needs a grid antisymmetric in x and y and N a multiple
    of 4; the x-pair kernel only x antisymmetry and an even N."""
    from totrilib_amd import _lib
    lib = _lib.load()
    o, sz = _lib.f3((-1, -1, 0.5)), _lib.f3((2, 2, 2))

    def paths(X=8, Y=8, N=16, origin=o):
        return lib.totri_debug_r1_paths(X, Y, 8, N, 64, origin, sz)

    assert paths() == 1 | 2
    assert paths(Y=7) == 2                                # odd Y: x-pairs only
    assert paths(origin=_lib.f3((-1, -0.9, 0.5))) == 2    # box off-centre in y
    assert paths(origin=_lib.f3((-0.9, -1, 0.5))) == 0    # off-centre in x
    assert paths(N=6) == 2                                # pairs, no quads
    assert paths(N=5) == 0
    assert paths(X=7) == 0
    assert paths(X=-1) < 0


def test_a1_row_walks_are_gated_by_the_float_bin_counter():
    """The A1 row walks count bins in a float, exact below 2^24; longer
    transients must take the window walks (host-side plan, no GPU)."""
    from totrilib_amd import _lib
    lib = _lib.load()
    o, sz = _lib.f3((-1, -1, 0.5)), _lib.f3((2, 2, 2))

    def walks(T, N=4, w=0.0025, off=0.0, X=8, Y=8, origin=o):
        return lib.totri_debug_a1_walks(X, Y, 8, N, T, origin, sz, w, off)

    assert walks(1024) == 1 | 2 | 4 | 8
    big = (1 << 24) + 16
    assert walks(big) == 1                 # x-mirror window walk only
    assert walks((1 << 24) - 2) == 15
    assert walks((1 << 24) - 1) == 1
    assert walks(1024, off=-0.1) == 1      # rows not monotone
    assert walks(1024, N=3) == 4           # no pairs: general row walk
    assert walks(big, N=3) == 0            # general window walk
    assert walks(1024, origin=_lib.f3((-1, -0.9, 0.5))) == 1 | 4 | 8  # y not antisymmetric
    assert walks(1024, X=7) == 4
