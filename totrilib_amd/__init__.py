"""MI355X-native volumetric transient library (drop-in for totrilib's
volume renderer, backprojections and RBF splat).

Module layout mirrors the reference package `totri`:
    totrilib_amd.types        VolumeDef, BinDef, C0
    totrilib_amd.render       VolumeRenderConfocal, VolumeRenderExhaustive
    totrilib_amd.reconstruct  Backprojection* and second_derivative_filter
    totrilib_amd.util         Rbf2VolumeGaussian, origin_of_wallpoints
    totrilib_amd.data         make_wall_grid
    totrilib_amd.totri_cpp    the compiled-module API (render / reconstruct / util)
The kernels live in libtotri_hip.so (C ABI: include/totri_hip.h).
"""
from __future__ import annotations

import importlib
import sys

__version__ = "0.1.0"

_ALIASES = {
    "totri": "totrilib_amd",
    "totri.types": "totrilib_amd.types",
    "totri.render": "totrilib_amd.render",
    "totri.render.volume": "totrilib_amd.render.volume",
    "totri.reconstruct": "totrilib_amd.reconstruct",
    "totri.reconstruct.backprojection": "totrilib_amd.reconstruct.backprojection",
    "totri.reconstruct.fk_migration": "totrilib_amd.reconstruct.fk_migration",
    "totri.util": "totrilib_amd.util",
    "totri.util.rbf2volume": "totrilib_amd.util.rbf2volume",
    "totri.util.format": "totrilib_amd.util.format",
    "totri.data": "totrilib_amd.data",
    "totri.data.util": "totrilib_amd.data.util",
    "totri.data.pfm": "totrilib_amd.data.pfm",
    "totri.data.otooleconfocal2018": "totrilib_amd.data.otooleconfocal2018",
    "totri.data.veltennc2012": "totrilib_amd.data.veltennc2012",
    "totri_cpp": "totrilib_amd.totri_cpp",
    "totri_cpp.render": "totrilib_amd.totri_cpp.render",
    "totri_cpp.reconstruct": "totrilib_amd.totri_cpp.reconstruct",
    "totri_cpp.util": "totrilib_amd.totri_cpp.util",
}


def install_as_totri(extension_only: bool = False) -> None:
    """Register this package under the reference's module names (`totri`,
    `totri.render`, ..., `totri_cpp`, `totri_cpp.render`, ...) so that code
    written against the reference imports it unchanged.  With
    `extension_only=True` only the compiled-module names (`totri_cpp.*`) are
    registered, so the reference's own Python layer runs on these kernels."""
    for alias, target in _ALIASES.items():
        if extension_only and not alias.startswith("totri_cpp"):
            continue
        if alias not in sys.modules:
            sys.modules[alias] = importlib.import_module(target)
