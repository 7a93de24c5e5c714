"""ctypes wrapper of oracle/liboracle.so -- TEST INFRASTRUCTURE ONLY.

Each function restates the reference arithmetic (see totri_oracle.c for the
file:line citations) and returns float64 numpy arrays.  Modes:
  0 FP32 op-for-op, 1 FP64 weights with FP32-pinned decisions, 2 pure FP64.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")

FP32, PINNED64, PURE64 = 0, 1, 2

_lib = None


def build() -> str:
    subprocess.run(["make", "-s", "-C", HERE, "liboracle.so"], check=True)
    return LIB


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB) or os.path.getmtime(LIB) < os.path.getmtime(
                os.path.join(HERE, "totri_oracle.c")):
            build()
        _lib = ctypes.CDLL(LIB)
    return _lib


def _f32(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.float32))


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64))


def _p(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


def _f3(v):
    return (ctypes.c_float * 3)(*[float(x) for x in v])


def set_threads(n: int) -> None:
    lib().oracle_set_num_threads(ctypes.c_int(n))


def num_threads() -> int:
    return int(lib().oracle_num_threads())


def r1_forward(volume, mp, T, origin, size, w, off, mode=FP32, i0=0, i1=None, cols=None):
    """R1 transient [i1-i0, T] (volume [X,Y,Z], mp [3,N]).  With `cols`
    [n, 2] only those (xi, yi) columns are summed, in the listed order -- in
    (xi, yi) index order that is the exact render of a volume that is zero
    outside them."""
    vol = _f32(volume)
    m = _f32(mp)
    X, Y, Z = vol.shape
    N = m.shape[1]
    i1 = N if i1 is None else i1
    out = np.zeros((max(i1 - i0, 0), T), dtype=np.float64)
    if cols is None:
        cp, nc = None, 0
    else:
        c = np.ascontiguousarray(np.asarray(cols, dtype=np.int32))
        cp, nc = _p(c), c.shape[0]
    lib().oracle_r1_forward(_p(vol), X, Y, Z, _p(m), N, T, _f3(origin), _f3(size),
                            ctypes.c_float(w), ctypes.c_float(off), mode, _p(out),
                            i0, i1, cp, nc)
    return out


def r1_decisions(idx, mp, X, Y, Z, origin, size, w, off):
    """Per-sample FP32 decisions for rows (i, t, xi, yi) of idx [n, 4]:
    returns (zi, alpha, d2); zi = -1000000 where d2 <= 0."""
    ix = np.ascontiguousarray(np.asarray(idx, dtype=np.int32))
    m = _f32(mp)
    n = ix.shape[0]
    zi = np.zeros(n, dtype=np.int32)
    al = np.zeros(n, dtype=np.float32)
    d2 = np.zeros(n, dtype=np.float32)
    lib().oracle_r1_decisions(_p(ix), n, _p(m), m.shape[1], X, Y, Z, _f3(origin),
                              _f3(size), ctypes.c_float(w), ctypes.c_float(off),
                              _p(zi), _p(al), _p(d2))
    return zi, al, d2


def a1_transpose(G, mp, X, Y, Z, origin, size, w, off, mode=FP32, cols=None):
    """Exact transpose of R1 applied to G [N, T] -> volume [X, Y, Z]; only the
    (xi, yi) columns listed in `cols` [n, 2] are computed (others are 0)."""
    g = _f64(G)
    m = _f32(mp)
    N, T = g.shape
    out = np.zeros((X, Y, Z), dtype=np.float64)
    if cols is None:
        cp, nc = None, 0
    else:
        c = np.ascontiguousarray(np.asarray(cols, dtype=np.int32))
        cp, nc = _p(c), c.shape[0]
    lib().oracle_a1_transpose(_p(g), _p(m), N, T, X, Y, Z, _f3(origin), _f3(size),
                              ctypes.c_float(w), ctypes.c_float(off), mode, cp, nc,
                              _p(out))
    return out


def a1_reference(G, mp, X, Y, Z, origin, size, w, off, mode=FP32, cols=None):
    """The reference's own window-form A1 (volume_cuda.cu:95-160)."""
    g = _f64(G)
    m = _f32(mp)
    N, T = g.shape
    out = np.zeros((X, Y, Z), dtype=np.float64)
    if cols is None:
        cp, nc = None, 0
    else:
        c = np.ascontiguousarray(np.asarray(cols, dtype=np.int32))
        cp, nc = _p(c), c.shape[0]
    lib().oracle_a1_reference(_p(g), _p(m), N, T, X, Y, Z, _f3(origin), _f3(size),
                              ctypes.c_float(w), ctypes.c_float(off), mode, cp, nc,
                              _p(out))
    return out


def a6(volume, mp, G, origin, size, w, off, mode=FP32, i0=0, i1=None):
    """Gradient w.r.t. measurement points [3, N] (volume_cuda.cu:188-265)."""
    vol = _f32(volume)
    m = _f32(mp)
    g = _f64(G)
    X, Y, Z = vol.shape
    N, T = g.shape
    i1 = N if i1 is None else i1
    out = np.zeros((3, N), dtype=np.float64)
    lib().oracle_a6(_p(vol), X, Y, Z, _p(m), N, _p(g), T, _f3(origin), _f3(size),
                    ctypes.c_float(w), ctypes.c_float(off), mode, _p(out), i0, i1)
    return out


def velten_confocal(tr, scan, sorigin, res, vs, ve, w, off, mode=FP32):
    """A2 (backprojection_cuda.cu:152-194): tr [B,T,S], scan [B,S,3],
    sorigin [B,So,3] (So in {0,S}) -> volume [B,D,H,W]."""
    t = _f64(tr)
    s = _f32(scan)
    so = _f32(sorigin) if sorigin is not None and np.asarray(sorigin).size else np.zeros((t.shape[0], 0, 3), np.float32)
    B, T, S = t.shape
    D, H, W = res
    out = np.zeros((B, D, H, W))
    lib().oracle_velten_confocal(_p(t), B, T, S, _p(s), _p(so), so.shape[1], D, H, W,
                                 _f3(vs), _f3(ve), ctypes.c_float(w), ctypes.c_float(off),
                                 mode, _p(out))
    return out


def velten_exhaustive(tr, scan, laser, sorigin, lorigin, res, vs, ve, w, off, mode=FP32):
    """A3 (backprojection_cuda.cu:228-287): tr [B,T,S,L], scan [B,S,3],
    laser [B,L,3], origins [B,So,3] / [B,Lo,3] -> volume [B,D,H,W]."""
    t = _f64(tr)
    s = _f32(scan)
    l = _f32(laser)
    B, T, S, L = t.shape
    so = _f32(sorigin) if sorigin is not None else np.zeros((B, 0, 3), np.float32)
    lo = _f32(lorigin) if lorigin is not None else np.zeros((B, 0, 3), np.float32)
    D, H, W = res
    out = np.zeros((B, D, H, W))
    lib().oracle_velten_exhaustive(_p(t), B, T, S, L, _p(s), _p(l), _p(so), so.shape[1],
                                   _p(lo), lo.shape[1], D, H, W, _f3(vs), _f3(ve),
                                   ctypes.c_float(w), ctypes.c_float(off), mode, _p(out))
    return out


def arellano(tr, scan, laser, sorigin, lorigin, res, vs, ve, w, off, mode=FP32):
    """A4 (backprojection_cuda.cu:10-115): tr [B,T,S,L], scan [B,3,S],
    laser [B,3,L], origins [B,3,So] / [B,3,Lo] -> volume [B,D,H,W]."""
    t = _f64(tr)
    s = _f32(scan)
    l = _f32(laser)
    B, T, S, L = t.shape
    so = _f32(sorigin) if sorigin is not None else np.zeros((B, 3, 0), np.float32)
    lo = _f32(lorigin) if lorigin is not None else np.zeros((B, 3, 0), np.float32)
    D, H, W = res
    out = np.zeros((B, D, H, W))
    lib().oracle_arellano(_p(t), B, T, S, L, _p(s), _p(l), _p(so), so.shape[2], _p(lo),
                          lo.shape[2], D, H, W, _f3(vs), _f3(ve), ctypes.c_float(w),
                          ctypes.c_float(off), mode, _p(out))
    return out


def r2_forward(vol, scan, laser, T, vs, ve, w, off, mode=FP32):
    """R2 (volume_cuda.cu:297-377): vol [B,D,H,W], scan [B,3,S], laser [B,3,L]
    -> transient [B,T,S,L].  mode 0 FP32; 1 FP32 weights, FP64 sums;
    2 FP32 decisions, FP64 weights and sums."""
    v = _f32(vol)
    s = _f32(scan)
    l = _f32(laser)
    B, D, H, W = v.shape
    S, L = s.shape[2], l.shape[2]
    out = np.zeros((B, T, S, L))
    lib().oracle_r2_forward(_p(v), B, D, H, W, _p(s), S, _p(l), L, T, _f3(vs), _f3(ve),
                            ctypes.c_float(w), ctypes.c_float(off), mode, _p(out))
    return out


def r2_adjoint(G, res, scan, laser, vs, ve, w, off, mode=FP32):
    """Exact adjoint of R2 applied to G [B,T,S,L] -> [B,D,H,W]."""
    g = _f64(G)
    s = _f32(scan)
    l = _f32(laser)
    B, T, S, L = g.shape
    D, H, W = res
    out = np.zeros((B, D, H, W))
    lib().oracle_r2_adjoint(_p(g), B, D, H, W, _p(s), S, _p(l), L, T, _f3(vs), _f3(ve),
                            ctypes.c_float(w), ctypes.c_float(off), mode, _p(out))
    return out


def splat_forward(rbf, res, vs, ve, trunc=0.0, mode=FP32, vox=None):
    """S1 (rbf2volume_cuda.cu:49-93): rbf [B,N,4+A] -> [B,1+A,D,H,W].  With
    `vox` (flat indices b*D*H*W + (z*H + y)*W + x) only those voxels are
    computed; the others are left 0."""
    r = _f32(rbf)
    B, N, C = r.shape
    A = C - 4
    D, H, W = res
    out = np.zeros((B, 1 + A, D, H, W))
    if vox is None:
        vp, nv = None, 0
    else:
        v = np.ascontiguousarray(np.asarray(vox, dtype=np.int64))
        vp, nv = _p(v), v.shape[0]
    lib().oracle_splat_forward(_p(r), B, N, A, D, H, W, _f3(vs), _f3(ve),
                               ctypes.c_float(trunc), mode, _p(out), vp,
                               ctypes.c_longlong(nv))
    return out


def splat_grad(rbf, vol, vgrad, vs, ve, trunc=0.0, mode=FP32):
    """S2 (rbf2volume_cuda.cu:123-201): gradient [B,N,4+A]."""
    r = _f32(rbf)
    v = _f64(vol)
    g = _f64(vgrad)
    B, N, C = r.shape
    A = C - 4
    D, H, W = v.shape[2:]
    out = np.zeros((B, N, C))
    lib().oracle_splat_grad(_p(r), B, N, A, _p(v), _p(g), D, H, W, _f3(vs), _f3(ve),
                            ctypes.c_float(trunc), mode, _p(out))
    return out


def r1_contraction_flips(mp, T, X, Y, Z, origin, size, w, off, i0=0, i1=None):
    """Compare R1's decisions in the uncontracted FP32 reading with those under
    nvcc's default FMA contraction, over scan rows [i0, i1), every bin and
    column: (active samples, activity flips, floor-index flips)."""
    m = _f32(mp)
    N = m.shape[1]
    i1 = N if i1 is None else i1
    out = np.zeros(3, dtype=np.int64)
    lib().oracle_r1_contraction_flips(_p(m), N, T, X, Y, Z, _f3(origin), _f3(size),
                                      ctypes.c_float(w), ctypes.c_float(off), i0, i1, _p(out))
    return tuple(int(v) for v in out)


# -- the reference's double dispatch (see totri_oracle.c) ---------------------
def r1_forward_f64(volume, mp, T, origin, size, w, off, i0=0, i1=None):
    """R1 with scalar_t = double: transient rows [i1-i0, T]."""
    vol = _f64(volume)
    m = _f64(mp)
    X, Y, Z = vol.shape
    N = m.shape[1]
    i1 = N if i1 is None else i1
    out = np.zeros((max(i1 - i0, 0), T), dtype=np.float64)
    lib().oracle_r1_forward_f64(_p(vol), X, Y, Z, _p(m), N, T, _f3(origin), _f3(size),
                                ctypes.c_float(w), ctypes.c_float(off), _p(out), i0, i1)
    return out


def _cols(cols):
    if cols is None:
        return None, None, 0
    c = np.ascontiguousarray(np.asarray(cols, dtype=np.int32))
    return c, _p(c), c.shape[0]


def a1_transpose_f64(G, mp, X, Y, Z, origin, size, w, off, cols=None):
    """Exact transpose of r1_forward_f64 applied to G [N, T] -> [X, Y, Z]."""
    g = _f64(G)
    m = _f64(mp)
    N, T = g.shape
    out = np.zeros((X, Y, Z), dtype=np.float64)
    keep, cp, nc = _cols(cols)
    lib().oracle_a1_transpose_f64(_p(g), _p(m), N, T, X, Y, Z, _f3(origin), _f3(size),
                                  ctypes.c_float(w), ctypes.c_float(off), cp, nc, _p(out))
    return out


def a1_reference_f64(G, mp, X, Y, Z, origin, size, w, off, cols=None):
    """The reference's window-form A1 with scalar_t = double."""
    g = _f64(G)
    m = _f64(mp)
    N, T = g.shape
    out = np.zeros((X, Y, Z), dtype=np.float64)
    keep, cp, nc = _cols(cols)
    lib().oracle_a1_reference_f64(_p(g), _p(m), N, T, X, Y, Z, _f3(origin), _f3(size),
                                  ctypes.c_float(w), ctypes.c_float(off), cp, nc, _p(out))
    return out


def velten_confocal_f64(tr, scan, sorigin, res, vs, ve, w, off):
    """A2 with scalar_t = double: tr [B,T,S], scan [B,S,3], sorigin [B,So,3]."""
    t = _f64(tr)
    s = _f64(scan)
    B, T, S = t.shape
    so = _f64(sorigin) if sorigin is not None else np.zeros((B, 0, 3))
    D, H, W = res
    out = np.zeros((B, D, H, W))
    lib().oracle_velten_confocal_f64(_p(t), B, T, S, _p(s), _p(so), so.shape[1], D, H, W,
                                     _f3(vs), _f3(ve), ctypes.c_float(w), ctypes.c_float(off),
                                     _p(out))
    return out
