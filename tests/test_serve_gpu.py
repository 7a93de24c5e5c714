"""GPU tests of the serving engine: every result of a batched, graph-replayed
step is bit-identical to the single-request entry points on the same inputs,
each request's result is the same alone and in a mix, and admission holds a
fixed device budget."""
import numpy as np
import pytest
import torch

from tests.cases import ORIGIN, SIZE, sparse_albedo, wall_grid

pytestmark = pytest.mark.gpu


def dev():
    return torch.device("cuda:0")


def make_requests(seed=0):
    """Renders and backprojections of mixed sizes that exercise every plan:
    column groups (few scan points), 1/2/4/8/16 bin windows and two bin
    passes, A1 scan-point ranges and the tiny-radius (safe) A1 form, mirror
    and general scans, inputs on the host (pinned and pageable) and resident
    on the device."""
    from totrilib_amd.serve import BackprojectRequest, RenderRequest
    rng = np.random.default_rng(seed)
    # Mirror kernels need a grid whose cell centres are antisymmetric in
    # float: over [-1, 1] that holds for power-of-two sizes (16, 32, 64), not
    # for 20, 24 or 40, which take the general kernels.
    specs = [  # kind, scan side, (X, Y, Z), T, one-way bin width, offset, where
        ("r", 8, (16, 16, 40), 300, 0.009, 0.0, "pinned"),
        ("r", 16, (32, 32, 32), 1100, 0.0025, 0.0, "device"),
        ("b", 8, (16, 16, 40), 300, 0.009, 0.0, "pageable"),
        ("b", 32, (64, 64, 64), 512, 0.005, 0.0, "pinned"),
        ("r", 32, (64, 64, 64), 512, 0.005, 0.0, "pageable"),
        ("b", 12, (16, 16, 20), 300, 0.009, 0.0002, "device"),
        ("b", 12, (16, 24, 20), 300, 0.009, 0.0002, "device"),
        ("r", 6, (16, 16, 24), 64, 0.04, 0.3, "pinned"),
        ("r", 6, (20, 16, 24), 64, 0.04, 0.3, "pinned"),
        ("r", 6, (20, 16, 24), 100, 0.03, 0.1, "device"),
        ("r", 10, (18, 18, 30), 200, 0.02, 0.0, "pinned"),
        ("b", 20, (40, 36, 33), 1500, 0.002, 0.0, "device"),
        # scans without mirror symmetry: the general kernels (A1 row walk;
        # A1 window walk for a negative offset; R1 below-the-box test)
        ("b", 12, (24, 20, 32), 400, 0.006, 0.0, "device", "perturbed"),
        ("b", 10, (20, 20, 24), 300, 0.009, -0.01, "pinned", "perturbed"),
        ("r", 12, (24, 20, 32), 400, 0.006, 0.0, "device", "random"),
        # odd Y: x-pairs only (the x-pair row walk; the x-mirror window walk
        # for a negative offset; the x-pair R1)
        ("b", 10, (16, 15, 24), 300, 0.009, 0.0, "device"),
        ("b", 10, (16, 15, 24), 300, 0.009, -0.01, "pinned"),
        ("r", 10, (16, 15, 24), 300, 0.009, 0.0, "device"),
        ("r", 10, (16, 15, 24), 300, 0.009, -0.01, "pageable"),
    ]
    reqs = []
    for k, (kind, n, res, T, w, off, where, *scan) in enumerate(specs):
        mp_np = wall_grid(n)
        if scan == ["perturbed"]:
            mp_np[0, 5] = np.nextafter(mp_np[0, 5], np.float32(2))
        elif scan == ["random"]:
            mp_np = np.stack([rng.uniform(-1, 1, n * n), rng.uniform(-1, 1, n * n),
                              np.zeros(n * n)]).astype(np.float32)
        mp = torch.from_numpy(mp_np)
        if kind == "r":
            data = torch.from_numpy(sparse_albedo(res, 0.05, 100 + k))
        else:
            data = torch.from_numpy(rng.standard_normal((n * n, T)).astype(np.float32))
        if where == "pinned":
            data = data.pin_memory()
        elif where == "device":
            data, mp = data.to(dev()), mp.to(dev())
        if kind == "r":
            reqs.append(RenderRequest(data, mp, ORIGIN, SIZE, T, w, off))
        else:
            reqs.append(BackprojectRequest(data, mp, ORIGIN, SIZE, res, w, off))
    return reqs


def single_request(req):
    """The same request through the single-request entry points."""
    from totrilib_amd.serve import RenderRequest
    from totrilib_amd.totri_cpp.render import (volume_confocal_forward,
                                               volume_confocal_grad_volume)
    mp = req.measurement_points.to(dev())
    if isinstance(req, RenderRequest):
        out = torch.empty(req.output_shape(), device=dev())
        volume_confocal_forward(req.volume.to(dev()), mp, out, req.volume_bbox_origin,
                                req.volume_bbox_size, req.bin_width, req.bin_offset)
    else:
        out = torch.empty(req.output_shape(), device=dev())
        volume_confocal_grad_volume(out, mp, req.transient.to(dev()), req.volume_bbox_origin,
                                    req.volume_bbox_size, req.bin_width, req.bin_offset)
    torch.cuda.synchronize()
    return out.cpu()


def collect(tickets):
    out = []
    for tk in tickets:
        assert tk.error is None, tk.error
        out.append(tk.result().clone())
        tk.release()
    return out


def paths_allowed(req):
    """Which kernels the request's geometry allows, from the library's host
    queries: R1 (1 xy-mirror, 2 x-pair) or A1 walks (1 x-mirror, 2 xy-mirror
    row walk, 4 general row walk, 8 x-pair row walk)."""
    from totrilib_amd import _lib
    from totrilib_amd.serve import RenderRequest
    lib = _lib.load()
    X, Y, Z, N, T = req.dims()
    o, sz = _lib.f3(req.volume_bbox_origin), _lib.f3(req.volume_bbox_size)
    if isinstance(req, RenderRequest):
        return "r", lib.totri_debug_r1_paths(X, Y, Z, N, T, o, sz)
    return "b", lib.totri_debug_a1_walks(X, Y, Z, N, T, o, sz, req.bin_width, req.bin_offset)


def test_requests_cover_every_path():
    """The mix holds renders for the xy-mirror, x-pair and general R1 and
    backprojections for each A1 walk (the geometry allows; the scans'
    mirror pairs decide on the device)."""
    got = {paths_allowed(r) for r in make_requests()}
    for want in [("r", 3), ("r", 2), ("r", 0), ("b", 15), ("b", 13), ("b", 4), ("b", 1),
                 ("b", 0)]:
        assert want in got, (want, sorted(got))


@pytest.mark.parametrize("graphs,sliced", [(True, None), (False, None), (True, "greedy"),
                                           (True, "one")])
def test_batched_step_matches_single_request_entry_points(graphs, sliced):
    """Every request of the mix through batched steps equals the single-request
    entry point bit for bit.  "greedy" and "one" run the larger requests in
    slices of their work items over several steps (slice cost 20 us of
    estimated work: up to 9 slices per request here); "greedy" lets a step
    that would run alone take all remaining slices, "one" takes one slice per
    step."""
    from totrilib_amd.serve import ServingEngine
    reqs = make_requests()
    eng = ServingEngine(device_budget=1 << 30, host_budget=256 << 20, lanes=2,
                        use_graphs=graphs, slice_cost=2e-5 if sliced else None,
                        greedy_slices=sliced != "one")
    tickets = [eng.submit(r) for r in reqs]
    eng.drain()
    got = collect(tickets)
    assert eng.steps >= 1
    if sliced == "one":
        assert eng.steps > len(eng.lanes) * 2 and max(eng.jobs_per_step) > 1
    for k, (r, g) in enumerate(zip(reqs, got)):
        want = single_request(r)
        assert torch.equal(g, want), (k, float((g - want).abs().max()))
        assert torch.count_nonzero(want) > 0, k


def test_alone_equals_in_mix_and_graph_replays():
    """Each request alone (its own step) and in a mix of overlapping steps
    with Poisson arrivals: identical bits.  Three jobs per step over two
    lanes replays each lane's captured graph several times, with eager
    single-request work and synchronisations between the steps."""
    from totrilib_amd.serve import ServingEngine
    reqs = make_requests(seed=1)
    eng = ServingEngine(device_budget=1 << 30, host_budget=256 << 20, lanes=2, max_jobs=3)
    alone = []
    for r in reqs:
        tk = eng.submit(r)
        eng.drain()
        alone.append(collect([tk])[0])
        single_request(r)  # eager work between replays
    rng = np.random.default_rng(5)
    # a burst of five, then Poisson arrivals
    arrivals = np.concatenate([np.zeros(5), np.cumsum(rng.exponential(0.001, len(reqs) - 5))])
    mixed = collect(eng.run(reqs, list(arrivals)))
    assert max(eng.jobs_per_step) > 1
    assert eng.lanes[0].steps >= 3 and eng.steps > len(reqs)
    for k, (a, m) in enumerate(zip(alone, mixed)):
        assert torch.equal(a, m), k
        assert torch.equal(a, single_request(reqs[k])), k


def test_admission_respects_device_budget():
    from totrilib_amd.serve import BackprojectRequest, ServingEngine, request_pairs
    reqs = make_requests(seed=2)
    # above the largest single request (a backprojection's scan-range
    # partial volumes take up to 128 MiB), below the mix's total
    budget = 128 << 20
    eng = ServingEngine(device_budget=budget, host_budget=64 << 20, lanes=2)
    peak = 0
    tickets = [eng.submit(r) for r in reqs]
    # one request that can never fit
    big = BackprojectRequest(torch.zeros(4096, 2048), torch.from_numpy(wall_grid(64)),
                             ORIGIN, SIZE, (128, 128, 128), 0.0025)
    t_big = eng.submit(big)
    while eng.poll():
        peak = max(peak, eng.dev.arena.used)
    assert peak <= budget
    assert t_big.error is not None and "budget" in t_big.error
    got = collect(tickets)
    for r, g in zip(reqs, got):
        assert torch.equal(g, single_request(r))
    assert eng.dev.arena.used == 0 and eng.host.arena.used == 0
    assert sum(request_pairs(r) for r in reqs) > 0


def test_shortest_job_first_order():
    """With one lane, "sjf" runs a short request submitted after a long one in
    an earlier step (the long one is more than 4x as costly, so they do not
    share a step); "fifo" puts both into one step in arrival order."""
    from totrilib_amd.serve import ServingEngine, request_cost
    reqs = make_requests(seed=3)
    small, big = reqs[0], reqs[3]          # 16x16x40 render, 64^3 backprojection
    assert request_cost(big) > 4 * request_cost(small)
    for policy in ("sjf", "fifo"):
        eng = ServingEngine(device_budget=1 << 30, host_budget=256 << 20, lanes=1,
                            policy=policy)
        tb = eng.submit(big)
        ts = eng.submit(small)
        eng.drain()
        got = collect([tb, ts])
        assert torch.equal(got[0], single_request(big))
        assert torch.equal(got[1], single_request(small))
        if policy == "sjf":
            assert eng.jobs_per_step == [1, 1]
            assert ts.t_done < tb.t_done
        else:
            assert eng.jobs_per_step == [2]


def test_sliced_request_shares_its_steps_with_a_later_short_one():
    """With one lane, a request in slices leaves room for a short request
    submitted after it: the short one joins one of its steps and finishes
    first; both results equal the single-request entry points."""
    from totrilib_amd.serve import ServingEngine, request_cost
    reqs = make_requests(seed=4)
    small, big = reqs[0], reqs[3]          # 16x16x40 render, 64^3 backprojection
    slice_cost = request_cost(big) / 5.5   # six slices
    assert request_cost(small) < slice_cost
    eng = ServingEngine(device_budget=1 << 30, host_budget=256 << 20, lanes=1,
                        slice_cost=slice_cost, group_ratio=None, greedy_slices=False)
    tb = eng.submit(big)
    eng.poll()                             # the first slice starts alone
    ts = eng.submit(small)
    eng.drain()
    got = collect([tb, ts])
    assert torch.equal(got[0], single_request(big))
    assert torch.equal(got[1], single_request(small))
    assert eng.jobs_per_step[0] == 1 and eng.steps == 6 and max(eng.jobs_per_step) == 2
    assert ts.t_done < tb.t_done


def test_resident_inputs_wait_for_the_submitting_stream():
    """A resident input produced by asynchronous work on the caller's current
    stream right before submit() (and a non-contiguous resident view) is read
    only after that work: the lane's stream waits on an event recorded at
    submit.  The long matrix chain keeps the producing kernel queued while
    the step would otherwise start."""
    from totrilib_amd.serve import BackprojectRequest, RenderRequest, ServingEngine
    eng = ServingEngine(device_budget=1 << 30, host_budget=256 << 20, lanes=2)
    rng = np.random.default_rng(8)
    res, n, T, w = (48, 40, 44), 16, 600, 0.004
    vol_np = sparse_albedo(res, 0.05, 81)
    G_np = rng.standard_normal((n * n, T)).astype(np.float32)
    mp_np = wall_grid(n)
    reqs_ref = [RenderRequest(torch.from_numpy(vol_np), torch.from_numpy(mp_np), ORIGIN, SIZE,
                              T, w, 0.0),
                BackprojectRequest(torch.from_numpy(G_np), torch.from_numpy(mp_np), ORIGIN,
                                   SIZE, res, w, 0.0)]
    want = [single_request(r) for r in reqs_ref]
    for trial in range(3):
        a = torch.randn(3072, 3072, device=dev())
        for _ in range(12):
            a = torch.tanh(a @ a * (1.0 / 3072))
        z = a[0, 0] * 0.0
        vol = torch.from_numpy(vol_np).to(dev()) + z      # ready only after the chain
        G = torch.from_numpy(G_np).to(dev()) + z
        mp_t = (torch.from_numpy(np.ascontiguousarray(mp_np.T)).to(dev()) + z).T  # [3, N] view
        assert not mp_t.is_contiguous()
        t1 = eng.submit(RenderRequest(vol, mp_t, ORIGIN, SIZE, T, w, 0.0))
        t2 = eng.submit(BackprojectRequest(G, mp_t, ORIGIN, SIZE, res, w, 0.0))
        del vol, G, mp_t, a
        eng.drain()
        got = collect([t1, t2])
        for k in range(2):
            assert torch.equal(got[k], want[k]), (trial, k)


def test_drain_raises_when_unreleased_results_fill_the_result_arena():
    """Results held by the caller fill the pinned arena: the next request can
    never start, so drain() raises instead of spinning; after release() the
    queue drains."""
    from totrilib_amd.serve import RenderRequest, ServingEngine
    n, T = 16, 1000
    mp = torch.from_numpy(wall_grid(n))
    vol = torch.from_numpy(sparse_albedo((24, 24, 24), 0.05, 82))
    out_bytes = n * n * T * 4
    eng = ServingEngine(device_budget=512 << 20, host_budget=out_bytes + 4096, lanes=2)
    t1 = eng.submit(RenderRequest(vol, mp, ORIGIN, SIZE, T, 0.003, 0.0))
    eng.drain()
    assert t1.done and t1.error is None
    t2 = eng.submit(RenderRequest(vol, mp, ORIGIN, SIZE, T, 0.003, 0.0))
    with pytest.raises(RuntimeError, match="unreleased results"):
        eng.drain()
    assert eng.stalled() and not t2.done
    first = t1.result().clone()
    t1.release()
    eng.drain()
    assert torch.equal(collect([t2])[0], first)
