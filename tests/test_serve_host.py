"""CPU checks of the serving engine's host side: the arena allocator that
admission relies on, and the batched-step ABI's validation and scratch sizes
(no device calls)."""
import ctypes

import pytest

from totrilib_amd import _lib
from totrilib_amd.serve.arena import Arena


def test_arena_first_fit_alignment_and_coalescing():
    a = Arena(4096, align=256)
    x = a.alloc(1)
    y = a.alloc(300)
    z = a.alloc(256)
    assert (x, y, z) == (0, 256, 768)
    assert a.used == 256 + 512 + 256
    a.free(y)
    assert a.alloc(600) == 1024          # the 512-byte hole is too small
    assert a.alloc(512) == 256           # first fit reuses the hole
    a.free(256)
    a.free(x)
    a.free(z)
    a.free(1024)
    assert a.used == 0 and a.largest_free == 4096   # fully coalesced
    assert a.alloc(4097) is None


def test_arena_alloc_many_is_all_or_nothing():
    a = Arena(1024, align=256)
    keep = a.alloc(256)
    assert a.alloc_many([256, 256, 512]) is None
    assert a.used == 256 and a.largest_free == 768
    got = a.alloc_many([256, 512])
    assert got is not None and a.used == 1024
    assert a.fits([512, 512]) and not a.fits([512, 513])
    for o in got + [keep]:
        a.free(o)
    assert a.blocks() == []


def test_arena_random_ops_never_overlap():
    import random
    rng = random.Random(7)
    a = Arena(1 << 16, align=64)
    live = {}
    for _ in range(2000):
        if live and rng.random() < 0.45:
            off = rng.choice(list(live))
            a.free(off)
            del live[off]
        else:
            n = rng.randint(1, 3000)
            off = a.alloc(n)
            if off is not None:
                live[off] = -(-n // 64) * 64
    spans = sorted(live.items())
    for (o1, n1), (o2, _) in zip(spans, spans[1:]):
        assert o1 + n1 <= o2
    assert a.used == sum(live.values())


def job(kind, X, Y, Z, N, T, w=0.005, off=0.0, ws=None, ws_bytes=0):
    j = _lib.ConfocalJob(kind=kind, X=X, Y=Y, Z=Z, N=N, T=T,
                         volume_origin=_lib.f3((-1, -1, 0.5)), volume_size=_lib.f3((2, 2, 2)),
                         bin_width=w, bin_offset=off)
    j.volume = 0x1000
    j.measurement_points = 0x2000
    j.transient = 0x3000
    j.workspace = ws
    j.workspace_bytes = ws_bytes
    return j


def test_job_workspace_matches_single_request_entry_points():
    lib = _lib.load()
    # few scan points: R1 splits columns into groups, A1 splits scan ranges
    r = job(_lib.TOTRI_JOB_RENDER, 64, 64, 64, 1024, 512)
    want = lib.totri_volume_confocal_forward_workspace(64, 64, 64, 1024, 512)
    assert want > 0 and lib.totri_confocal_job_workspace(ctypes.byref(r)) == want
    b = job(_lib.TOTRI_JOB_BACKPROJECT, 64, 64, 64, 1024, 512)
    want = lib.totri_volume_confocal_grad_volume_workspace(64, 64, 64, 1024, 512)
    assert lib.totri_confocal_job_workspace(ctypes.byref(b)) == want
    # the query needs no buffers (it is made before they are allocated)
    b.volume = b.measurement_points = b.transient = None
    assert lib.totri_confocal_job_workspace(ctypes.byref(b)) == want
    # enough scan points for one column group and a box off the x = y = 0
    # mirror (no mirror path): an FP32 render needs none
    r = job(_lib.TOTRI_JOB_RENDER, 16, 16, 16, 8 * 4096, 64)
    r.volume_origin = _lib.f3((-0.9, -1, 0.5))
    assert lib.totri_confocal_job_workspace(ctypes.byref(r)) == 0
    r.volume_origin = _lib.f3((-1, -1, 0.5))   # centred: scratch for mirror detection
    assert lib.totri_confocal_job_workspace(ctypes.byref(r)) == \
        lib.totri_volume_confocal_forward_workspace(16, 16, 16, 8 * 4096, 64) > 0


def test_batch_prepare_validates_jobs():
    lib = _lib.load()
    nb = lib.totri_confocal_batch_table_bytes()
    table = (ctypes.c_uint8 * nb)()
    ok = job(_lib.TOTRI_JOB_RENDER, 8, 8, 8, 4096, 64)
    ok.volume_origin = _lib.f3((-0.9, -1, 0.5))   # no mirror path: needs no scratch
    arr = (_lib.ConfocalJob * 1)(ok)
    assert lib.totri_confocal_batch_prepare(arr, 1, table, nb) == _lib.TOTRI_OK
    assert lib.totri_confocal_batch_prepare(arr, 1, table, nb - 1) == _lib.TOTRI_EINVAL
    assert lib.totri_confocal_batch_prepare(arr, 65, table, nb) == _lib.TOTRI_EINVAL
    for bad in (job(7, 8, 8, 8, 64, 64), job(_lib.TOTRI_JOB_RENDER, 8, 8, 1, 64, 64),
                job(_lib.TOTRI_JOB_BACKPROJECT, 8, 8, 8, 0, 64)):
        arr = (_lib.ConfocalJob * 1)(bad)
        assert lib.totri_confocal_batch_prepare(arr, 1, table, nb) == _lib.TOTRI_EINVAL
    small = job(_lib.TOTRI_JOB_BACKPROJECT, 8, 8, 8, 64, 64, ws=0x4000, ws_bytes=16)
    arr = (_lib.ConfocalJob * 1)(small)
    assert lib.totri_confocal_batch_prepare(arr, 1, table, nb) == _lib.TOTRI_EWORKSPACE
    assert "workspace" in _lib.last_error()
    # a table that batch_prepare did not write is refused before any device call
    blank = (ctypes.c_uint8 * nb)()
    assert lib.totri_confocal_batch_run(blank, ctypes.c_void_p(0x10000), nb, None) == \
        _lib.TOTRI_EINVAL


def test_engine_requires_gpu_without_one():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    from totrilib_amd.serve import ServingEngine
    with pytest.raises(RuntimeError):
        ServingEngine(device_budget=1 << 20, host_budget=1 << 20)
