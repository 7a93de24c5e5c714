"""Bin decisions of the production R1 and A1 walks, bit for bit.

The L2 tests cannot see a wrong floor: lerp is continuous in z_hat, so a
sample whose z_hat sits next to an integer b moves only a tiny weight when
its floor flips between b-1 and b.  These tests make such a flip visible as a
change of support:

* A1: the upstream gradient is one-hot at one (scan point, bin) whose samples
  include near-boundary ones.  Each column's result may be nonzero only at
  the two voxels (zi, zi+1) that the FP32 decision of the oracle gives; a
  flipped floor writes its small weight one voxel outside that pair.
* R1: the volume is one-hot at voxel (c, b+1) for a near-boundary sample of
  column c at boundary b.  A sample that floors to b reads (v[b], v[b+1]) and
  is nonzero; one that floors to b-1 reads (v[b-1], v[b]) and is exactly 0.
  The whole transient's support must equal the FP32 oracle's.

They run the bench workload's shape (128^3 voxels, 64x64 scan, 1024 bins of
2.5 mm one-way) through the entry points, so the mirror-quad walks (R1
r1_sym_kernel, A1 a1_quad_kernel), the x-pair walks (a box shifted in y: R1
r1_pair_kernel, A1 a1_pair_kernel), the
general walks (one scan point moved by an ulp) and the window walks (a
negative bin offset) each take their own decisions (fast path behind the
guard, exact path within it).  Near-boundary samples are picked from the
oracle's FP32 decisions: |z_hat - round(z_hat)| < 2e-3, which spans both the
guarded exact path and the fast path just outside it.
Reference formula: src/render/volume_cuda.cu:23,35-49.
"""
import numpy as np
import pytest
import torch

from oracle import oracle as O
from tests.cases import ORIGIN, SIZE, wall_grid

pytestmark = pytest.mark.gpu

RES = (128, 128, 128)
T, W = 1024, 0.0025
WIN = 2e-3
# bound on |alpha_gpu - alpha_oracle|: outside the guard the kernels take
# alpha from the fast z_hat, which the guard's proven bound delta keeps within
# 2^-22 * B of the exact one (about 3.4e-4 at this geometry); typical
# differences are an ulp or two of z_hat
ALPHA_TOL = 4e-4


def dev():
    return torch.device("cuda:0")


def scenario(name):
    """(measurement points [3, N], origin, bin offset) of each path."""
    mp = wall_grid(64)
    origin, off = ORIGIN, 0.0
    if name == "general":
        mp[0, 777] = np.nextafter(mp[0, 777], np.float32(2))
    elif name == "xpairs":
        origin = (ORIGIN[0], ORIGIN[1] + 0.1, ORIGIN[2])
    elif name == "window":
        off = -0.01
    return np.ascontiguousarray(mp), origin, off


def near_boundary(mp, origin, off, n, seed):
    """n near-boundary samples (s, t, xi, yi, zi, b), b the integer next to
    z_hat, from random candidates; one per (s, t) and per column."""
    X, Y, Z = RES
    rng = np.random.default_rng(seed)
    s = rng.integers(0, mp.shape[1], 24)
    cols = np.stack([rng.integers(0, X, 96), rng.integers(0, Y, 96)], 1)
    S_, T_, C_ = np.meshgrid(s, np.arange(T), np.arange(len(cols)), indexing="ij")
    idx = np.stack([S_.ravel(), T_.ravel(), cols[C_.ravel(), 0], cols[C_.ravel(), 1]], 1)
    zi, al, d2 = O.r1_decisions(idx, mp, X, Y, Z, origin, SIZE, W, off)
    act = (d2 > 0) & (zi >= 0) & (zi <= Z - 2)
    lo = act & (al < WIN)
    hi = act & (al > 1 - WIN)
    cand = np.flatnonzero(lo | hi)
    rng.shuffle(cand)
    out, seen_st, seen_c = [], set(), set()
    for k in cand:
        st = (idx[k, 0], idx[k, 1])
        c = (idx[k, 2], idx[k, 3])
        b = zi[k] + (1 if hi[k] else 0)
        if st in seen_st or c in seen_c or not (1 <= b <= Z - 3):
            continue
        seen_st.add(st)
        seen_c.add(c)
        out.append((*idx[k], zi[k], b))
        if len(out) == n:
            break
    assert len(out) == n, len(out)
    return out


def mirror_info(mp):
    from totrilib_amd import _lib
    lib = _lib.load()
    m = torch.from_numpy(mp).to(dev())
    N = m.shape[1]
    ws = torch.empty(lib.totri_debug_mirror_workspace(N), dtype=torch.uint8, device=dev())
    info = torch.full((4,), -7, dtype=torch.int32, device=dev())
    _lib.check(lib.totri_debug_mirror_info(m.data_ptr(), N, ws.data_ptr(), ws.numel(),
                                           info.data_ptr(),
                                           torch.cuda.current_stream().cuda_stream),
               "debug_mirror_info")
    torch.cuda.synchronize()
    return info.cpu().tolist()


EXPECT_INFO = {"quads": [1, 1024, 1, 2048], "xpairs": [1, 1024, 1, 2048],
               "window": [1, 1024, 1, 2048], "general": [0, None, 0, None]}


def check_info(name, mp):
    info = mirror_info(mp)
    want = EXPECT_INFO[name]
    assert all(w is None or w == i for w, i in zip(want, info)), info


@pytest.mark.parametrize("name", ["quads", "xpairs", "general", "window"])
def test_a1_walk_decisions_one_hot(name):
    from totrilib_amd.totri_cpp.render import volume_confocal_grad_volume
    X, Y, Z = RES
    mp, origin, off = scenario(name)
    check_info(name, mp)
    m = torch.from_numpy(mp).to(dev())
    samples = near_boundary(mp, origin, off, 10, 11 + len(name))
    allc = np.stack(np.meshgrid(np.arange(X), np.arange(Y), indexing="ij"), -1).reshape(-1, 2)
    flips = 0
    for (s, t, xi, yi, zi0, b) in samples:
        print(f"{name}: one-hot G[{s}, {t}] (near-boundary column ({xi}, {yi}), boundary {b})",
              flush=True)
        G = torch.zeros(mp.shape[1], T, device=dev())
        G[s, t] = 1.0
        out = torch.full(RES, float("nan"), device=dev())
        volume_confocal_grad_volume(out, m, G, origin, SIZE, W, off)
        got = out.cpu().numpy().astype(np.float64)
        idx = np.concatenate([np.full((len(allc), 1), s), np.full((len(allc), 1), t), allc], 1)
        zi, al, d2 = O.r1_decisions(idx, mp, X, Y, Z, origin, SIZE, W, off)
        act = (d2 > 0) & (zi >= 0) & (zi <= Z - 2)
        tc = float(np.float32(np.float32(t) * np.float32(W)) + np.float32(off))
        depth = np.sqrt(d2[act].astype(np.float64))
        wgt = (depth / (tc * tc)) ** 4
        a = al[act].astype(np.float64)
        cx, cy, z = allc[act, 0], allc[act, 1], zi[act]
        want = np.zeros(RES)
        touched = np.zeros(RES, bool)
        np.add.at(want, (cx, cy, z), (1 - a) * wgt)
        np.add.at(want, (cx, cy, z + 1), a * wgt)
        touched[cx, cy, z] = True
        touched[cx, cy, z + 1] = True
        assert np.all(np.isfinite(got))
        outside = (got != 0) & ~touched
        missing = (want > 1e-3 * want.max()) & (got == 0)
        flips += int(outside.sum()) + int(missing.sum())
        assert not outside.any(), np.argwhere(outside)[:5]
        assert not missing.any(), np.argwhere(missing)[:5]
        # one sample per voxel: the weights carry that sample's alpha, whose
        # fast-path rounding is not averaged over scan points as in a sum
        assert np.abs(got - want).max() <= ALPHA_TOL * want.max()
        # the near-boundary column itself: exactly the oracle's pair
        nz = np.flatnonzero(got[xi, yi] != 0)
        assert len(nz) >= 1 and set(nz) <= {zi0, zi0 + 1}, (nz, zi0)
    assert flips == 0


@pytest.mark.parametrize("name", ["quads", "xpairs", "general", "window"])
def test_r1_walk_decisions_one_hot(name):
    from totrilib_amd.totri_cpp.render import volume_confocal_forward
    mp, origin, off = scenario(name)
    check_info(name, mp)
    m = torch.from_numpy(mp).to(dev())
    samples = near_boundary(mp, origin, off, 8, 23 + len(name))
    for (s, t, xi, yi, zi0, b) in samples:
        print(f"{name}: one-hot voxel ({xi}, {yi}, {b + 1}) (sample s={s} t={t} "
              f"boundary {b})", flush=True)
        vol = np.zeros(RES, np.float32)
        vol[xi, yi, b + 1] = 1.0
        out = torch.full((mp.shape[1], T), float("nan"), device=dev())
        volume_confocal_forward(torch.from_numpy(vol).to(dev()), m, out, origin, SIZE, W, off)
        got = out.cpu().numpy().astype(np.float64)
        ref = O.r1_forward(vol, mp, T, origin, SIZE, W, off, mode=O.FP32,
                           cols=np.array([[xi, yi]]))
        assert np.count_nonzero(ref) > 100
        diff = (got != 0) != (ref != 0)
        assert not diff.any(), np.argwhere(diff)[:5]
        # one occupied voxel: each bin is one or a few samples' weights (see
        # the A1 test for the alpha rounding)
        assert np.abs(got - ref).max() <= ALPHA_TOL * np.abs(ref).max()
