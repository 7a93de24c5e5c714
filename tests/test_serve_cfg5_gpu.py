"""The serving mix at its own sizes: scans of 32^2, 64^2 and 128^2 points,
512, 1024 and 2048 bins over 5.12 m of round trip, volumes of 64^3, 128^3 and
192^3, seeded sparse albedo (1-5 % occupied), 50/50 renders and
backprojections.  A seeded mini-mix always holds the largest corner (a
128^2-scan, 2048-bin, 192^3 render and backprojection).

* Each request's result is the same alone (its own step) and in the mix
  (Poisson arrivals over two lanes), bit for bit.
* The largest render matches the FP64-pinned oracle on sampled transient
  rows, and the largest backprojection on sampled volume columns, at 1e-5
  relative L2.
"""
import numpy as np
import pytest
import torch

from oracle import oracle as O
from tests.cases import ORIGIN, SIZE, rel_l2, sparse_albedo, wall_grid

pytestmark = pytest.mark.gpu

SIDES, BINS, RES = (32, 64, 128), (512, 1024, 2048), (64, 128, 192)
PATH = 5.12   # round-trip metres covered by the bins: one-way width 2.56 / T


def dev():
    return torch.device("cuda:0")


def mix(seed=6, draws=6):
    """(kind, side, T, res, occupancy) specs: the two largest corners, then
    seeded uniform draws from the sets."""
    rng = np.random.default_rng(seed)
    specs = [("render", 128, 2048, 192, 0.03), ("backproject", 128, 2048, 192, 0.02)]
    for _ in range(draws):
        specs.append(("render" if rng.random() < 0.5 else "backproject",
                      int(rng.choice(SIDES)), int(rng.choice(BINS)), int(rng.choice(RES)),
                      float(rng.uniform(0.01, 0.05))))
    return specs


def build_requests(specs):
    """Render inputs in pinned host memory; a backprojection's input is the
    render of its own sparse albedo, resident on the device (produced on the
    current stream right before it is submitted)."""
    from totrilib_amd.serve import BackprojectRequest, RenderRequest
    from totrilib_amd.totri_cpp.render import volume_confocal_forward
    reqs, albedo = [], []
    for k, (kind, n, T, r, p) in enumerate(specs):
        w = PATH / 2 / T
        mp = torch.from_numpy(wall_grid(n))
        vol = sparse_albedo((r, r, r), p, 600 + k)
        albedo.append(vol)
        if kind == "render":
            reqs.append(RenderRequest(torch.from_numpy(vol).pin_memory(), mp.pin_memory(),
                                      ORIGIN, SIZE, T, w, 0.0))
        else:
            m = mp.to(dev())
            tr = torch.empty((n * n, T), device=dev())
            volume_confocal_forward(torch.from_numpy(vol).to(dev()), m, tr, ORIGIN, SIZE, w, 0.0)
            reqs.append(BackprojectRequest(tr, m, ORIGIN, SIZE, (r, r, r), w, 0.0))
    return reqs, albedo


def collect(tickets):
    out = []
    for tk in tickets:
        assert tk.error is None, tk.error
        out.append(tk.result().clone())
        tk.release()
    return out


def test_cfg5_mix_alone_equals_in_mix_and_oracle():
    from totrilib_amd.serve import BackprojectRequest, RenderRequest, ServingEngine
    specs = mix()
    print("specs:", specs, flush=True)
    reqs, albedo = build_requests(specs)
    eng = ServingEngine(device_budget=16 << 30, host_budget=2 << 30, lanes=2)
    alone = []
    for k, r in enumerate(reqs):
        tk = eng.submit(r)
        eng.drain()
        alone.append(collect([tk])[0])
        print(f"alone {k}: {specs[k]} latency {tk.latency * 1e3:.1f} ms", flush=True)
    rng = np.random.default_rng(16)
    arrivals = np.concatenate([[0.0, 0.0], np.cumsum(rng.exponential(0.02, len(reqs) - 2))])
    tickets = eng.run(reqs, list(arrivals))
    mixed = collect(tickets)
    print("jobs per step:", eng.jobs_per_step, flush=True)
    for k, (a, m) in enumerate(zip(alone, mixed)):
        assert torch.equal(a, m), k
        assert torch.count_nonzero(a) > 0, k
    assert eng.dev.arena.used == 0 and eng.host.arena.used == 0

    # the largest render: sampled transient rows against the FP64-pinned oracle
    k = 0
    r = reqs[k]
    assert isinstance(r, RenderRequest) and r.volume.shape == (192, 192, 192)
    mp = r.measurement_points.numpy()
    rows = np.random.default_rng(3).choice(mp.shape[1], 8, replace=False)
    got = alone[k].numpy()
    for i in rows:
        ref = O.r1_forward(albedo[k], mp, r.bin_num, ORIGIN, SIZE, r.bin_width, 0.0,
                           mode=O.PINNED64, i0=int(i), i1=int(i) + 1)
        e = rel_l2(got[i:i + 1], ref)
        print(f"render row {i}: rel-L2 vs FP64 pinned {e:.2e}", flush=True)
        assert e < 1e-5
    # the largest backprojection: sampled columns
    k = 1
    r = reqs[k]
    assert isinstance(r, BackprojectRequest) and tuple(r.volume_resolution) == (192, 192, 192)
    cols = np.stack([np.random.default_rng(4).integers(0, 192, 12),
                     np.random.default_rng(5).integers(0, 192, 12)], 1)
    ref = O.a1_transpose(r.transient.cpu().numpy(), r.measurement_points.cpu().numpy(),
                         192, 192, 192, ORIGIN, SIZE, r.bin_width, 0.0, mode=O.PINNED64,
                         cols=cols)
    got = alone[k].numpy()
    gs, rs = got[cols[:, 0], cols[:, 1]], ref[cols[:, 0], cols[:, 1]]
    e = rel_l2(gs, rs)
    print(f"backprojection columns: rel-L2 vs FP64 pinned {e:.2e}", flush=True)
    assert np.count_nonzero(rs) > 100
    assert e < 1e-5
