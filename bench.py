"""Benchmark: confocal render + backprojection (albedo gradient) throughput.

Metric: voxel x scan-point pairs per second for one forward render (R1) plus
one backward pass (A1, the albedo gradient of an L2 loss), at
128^3 voxels / 64x64 scan points / 1024 bins of 0.005 m round trip
(one-way bin width 0.0025 m), seeded 3 % sparse albedo, FP32.

One step = transient = R1(v); G = 2 (transient - target); grad = A1(G).
Inputs live in HBM before the timed region.  The timed region is exactly
`--steps` steps bracketed by device synchronisation; per-step HIP events give
p50/p99.  The JSON line also carries the roofline of the dominant kernel
(FP32 VALU, algorithmic FLOPs per launch / measured launch time) and the CPU
baseline (the oracle's FP32 restatement of the same formulas on host cores).

Usage: python bench.py [--gpus 1] [--steps 20] [--warmup 3] [--no-cpu]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP32 = 157.3e12   # MI355X published FP32 vector peak (packed issue)
PEAK_HBM = 8.0e12      # MI355X HBM3E bytes/s
# Engine clock under these kernels' load: GRBM_GUI_ACTIVE over the dispatch time reads
# 2.26-2.35 GHz on MI355X (peak 2.4 GHz); 1024 SIMDs (256 CUs x 4)
SIMD_CLOCK_HZ = 2.3e9

WORKLOAD = dict(n_side=64, T=1024, res=128, w_round_trip=0.005, occupancy=0.03,
                seed=1, target_seed=2)
ORIGIN = (-1.0, -1.0, 0.5)
SIZE = (2.0, 2.0, 2.0)


def flop_model(mp: np.ndarray, res: int, T: int, w: float, off: float = 0.0):
    """Algorithmic FLOPs per (voxel, scan point) pair, rocprof-compute
    convention (add/mul 1, FMA 2, sqrt/div 1, floor/compare 0):

      R1 per active sample 18; A1 per window bin 17 plus 23 per pair.

    a_fwd = active R1 samples per pair; a_adj = bins the reference's window
    form of A1 visits per pair.  Both are counted in closed form per
    (scan point, column): active bins of a column form one contiguous range
    (radius between the first and last valid voxel-centre depths), and the
    window form telescopes to c(Z) + c(Z-1) - 2 c(0) with
    c(k) = clamp(ceil((|p - q_k| - off) / w), 0, T) at voxel centre k.
    """
    X = Y = Z = res
    s = Z / SIZE[2]
    x = ORIGIN[0] + (np.arange(X) + 0.5) * SIZE[0] / X
    y = ORIGIN[1] + (np.arange(Y) + 0.5) * SIZE[1] / Y
    zc = ORIGIN[2] + (np.arange(Z) + 0.5) * SIZE[2] / Z
    fwd = 0.0
    adj = 0.0
    for i in range(mp.shape[1]):
        wx, wy, wz = (float(v) for v in mp[:, i])
        rho2 = ((wx - x)[:, None] ** 2 + (wy - y)[None, :] ** 2).ravel()
        dlo = max(ORIGIN[2] + 0.5 / s - wz, 0.0)
        dhi = ORIGIN[2] + (Z - 0.5) / s - wz
        if dhi <= 0:
            continue
        r_lo = np.sqrt(rho2 + dlo ** 2)
        r_hi = np.sqrt(rho2 + dhi ** 2)
        t0 = np.clip(np.ceil((r_lo - off) / w), 0, T)
        t1 = np.clip(np.ceil((r_hi - off) / w), 0, T)
        fwd += float(np.sum(t1 - t0))

        def c(k):
            d = np.sqrt(rho2 + (zc[k] - wz) ** 2) if k < Z else \
                np.sqrt(rho2 + (zc[Z - 1] + SIZE[2] / Z - wz) ** 2)
            return np.clip(np.ceil((d - off) / w), 0, T)
        adj += float(np.sum(c(Z) + c(Z - 1) - 2 * c(0)))
    pairs = float(X * Y * Z) * mp.shape[1]
    a_fwd = fwd / pairs
    a_adj = adj / pairs
    return dict(a_fwd=a_fwd, a_adj=a_adj, f_r1=18 * a_fwd, f_a1=17 * a_adj + 23,
                f_pair=18 * a_fwd + 17 * a_adj + 23)


def make_inputs(torch, dev):
    from tests.cases import sparse_albedo, wall_grid
    n, res, T = WORKLOAD["n_side"], WORKLOAD["res"], WORKLOAD["T"]
    mp = wall_grid(n)
    vol = sparse_albedo((res, res, res), WORKLOAD["occupancy"], WORKLOAD["seed"])
    tgt_vol = sparse_albedo((res, res, res), WORKLOAD["occupancy"], WORKLOAD["target_seed"])
    return mp, vol, tgt_vol


def mirror_info(torch, mp):
    """[quads ok, n quads, x-pairs ok, n pairs] of the device-side scan
    mirror detection that selects the symmetric kernels."""
    from totrilib_amd import _lib
    lib = _lib.load()
    N = mp.shape[1]
    ws = torch.empty(lib.totri_debug_mirror_workspace(N), dtype=torch.uint8, device=mp.device)
    info = torch.zeros(4, dtype=torch.int32, device=mp.device)
    _lib.check(lib.totri_debug_mirror_info(mp.data_ptr(), N, ws.data_ptr(), ws.numel(),
                                           info.data_ptr(),
                                           torch.cuda.current_stream().cuda_stream),
               "totri_debug_mirror_info")
    torch.cuda.synchronize()
    return info.cpu().tolist()


def cpu_model() -> str:
    """The host CPU's model name (lscpu's "Model name", from /proc/cpuinfo)."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_threads() -> int:
    """Threads for the CPU leg: the host's cores, capped at the 16-core
    share that one GPU's process gets on a shared host (OMP_NUM_THREADS, when
    set, lowers it)."""
    n = min(os.cpu_count() or 1, 16)
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit() and int(env) > 0:
        n = min(n, int(env))
    return n


def cpu_baseline(mp, vol, T, w):
    """Oracle FP32 restatement (same formulas, brute force as the reference
    kernels) on host cores: R1 on a seeded subset of scan points, A1 on a
    seeded subset of columns; work is linear in both, so the full-size time
    is extrapolated from the subsets (about 10 s of CPU work)."""
    from oracle import oracle as O
    threads = cpu_threads()
    O.set_threads(threads)
    X, Y, Z = vol.shape
    N = mp.shape[1]
    rng = np.random.default_rng(0)
    # R1: rows
    rows = min(N, 1024)
    i0 = int(rng.integers(0, N - rows + 1))
    t0 = time.perf_counter()
    O.r1_forward(vol, mp, T, ORIGIN, SIZE, w, 0.0, mode=O.FP32, i0=i0, i1=i0 + rows)
    t_r1 = (time.perf_counter() - t0) * N / rows
    # A1: columns
    ncols = min(X * Y, 2048)
    cols = np.stack([rng.integers(0, X, ncols), rng.integers(0, Y, ncols)], axis=1)
    G = rng.standard_normal((N, T)).astype(np.float32)
    t0 = time.perf_counter()
    O.a1_transpose(G, mp, X, Y, Z, ORIGIN, SIZE, w, 0.0, mode=O.FP32, cols=cols)
    t_a1 = (time.perf_counter() - t0) * (X * Y) / ncols
    pairs = float(X * Y * Z) * N
    return dict(value=pairs / (t_r1 + t_a1), unit="pairs/s", cores=threads, kind="port",
                threads=threads, host_cores=os.cpu_count(), cpu_model=cpu_model(),
                implementation="C/OpenMP port of the reference formulas (the oracle's "
                               "FP32 mode), not PyTorch",
                sample=(f"oracle FP32 restatement, {threads} OpenMP threads on "
                        f"{os.cpu_count()} host cores ({cpu_model()}; one GPU's 16-core "
                        f"share): R1 on {rows} of {N} scan points, A1 on {ncols} of "
                        f"{X * Y} columns, extrapolated linearly (R1 {t_r1:.1f} s + A1 "
                        f"{t_a1:.1f} s per full step)"))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--copy-probe", action="store_true",
                    help="also run one 256 MiB device-to-device copy (a known-traffic "
                         "kernel that calibrates FETCH_SIZE/WRITE_SIZE counters)")
    ap.add_argument("--workload", choices=["confocal", "serve", "kernels"],
                    default="confocal",
                    help="confocal: the render + gradient step (default); serve: the "
                         "request mix through the serving engine; kernels: each other "
                         "operation of the path at a stated size")
    ap.add_argument("--requests", type=int, default=48, help="serve: requests per run")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this benchmark runs on one GPU; --gpus must be 1 "
                                   "(the multi-GPU sharding is a design note, not built)"}))
        sys.exit(2)
    if args.workload == "serve":
        serve_main(args)
    elif args.workload == "kernels":
        kernels_main(args)
    else:
        confocal_main(args)


def general_scan(mp_np: np.ndarray) -> np.ndarray:
    """The scan with one point moved by one ulp in x: no mirror quads or
    pairs remain, so every kernel takes its general form (the geometry of a
    real capture, an off-centre box or a jittered scan)."""
    mp = mp_np.copy()
    mp[0, 5] = np.nextafter(mp[0, 5], np.float32(2))
    return mp


def kernel_names(info):
    r1 = ("r1_sym_kernel" if info[0] else
          "r1_pair_kernel" if info[2] else "r1_forward_kernel")
    a1 = ("a1_quad_kernel" if info[0] else
          "a1_pair_kernel" if info[2] else "a1_row_kernel")
    return r1, a1


def load_profile(name):
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return {}
    with open(path) as f:
        return json.load(f)


def confocal_main(args):
    import torch
    from totrilib_amd.totri_cpp.render import (volume_confocal_forward,
                                               volume_confocal_grad_volume)

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    T, res = WORKLOAD["T"], WORKLOAD["res"]
    w = WORKLOAD["w_round_trip"] / 2  # one-way radius per bin
    mp_np, vol_np, tgt_np = make_inputs(torch, dev)
    vol = torch.from_numpy(vol_np).to(dev)
    tgt = torch.from_numpy(tgt_np).to(dev)
    N = mp_np.shape[1]
    transient = torch.empty((N, T), device=dev)
    G = torch.empty((N, T), device=dev)
    grad = torch.empty((res, res, res), device=dev)

    def run_leg(mp_leg, steps, warmup):
        """Times `steps` steps (after `warmup`) on the scan mp_leg; returns the
        wall time and per-step event times of the step, R1 and A1."""
        mp = torch.from_numpy(mp_leg).to(dev)
        target = torch.empty((N, T), device=dev)
        volume_confocal_forward(tgt, mp, target, ORIGIN, SIZE, w, 0.0)

        def step(ev=None):
            if ev is not None:
                ev[0].record()
            volume_confocal_forward(vol, mp, transient, ORIGIN, SIZE, w, 0.0)
            if ev is not None:
                ev[1].record()
            torch.sub(transient, target, out=G)
            G.mul_(2.0)
            if ev is not None:
                ev[2].record()
            volume_confocal_grad_volume(grad, mp, G, ORIGIN, SIZE, w, 0.0)
            if ev is not None:
                ev[3].record()

        for _ in range(warmup):
            step()
        torch.cuda.synchronize()
        evs = [[torch.cuda.Event(enable_timing=True) for _ in range(4)] for _ in range(steps)]
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for k in range(steps):
            step(evs[k])
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        step_ms = np.array([e[0].elapsed_time(e[3]) for e in evs])
        r1_ms = np.array([e[0].elapsed_time(e[1]) for e in evs])
        a1_ms = np.array([e[2].elapsed_time(e[3]) for e in evs])
        return wall, step_ms, r1_ms, a1_ms, mirror_info(torch, mp)

    pairs = float(res ** 3) * N
    executed = load_profile("valu_flops.json")
    traffic_p = load_profile("traffic.json")
    issue_p = load_profile("valu.json")

    def roofline(mp_leg, info, r1_avg, a1_avg):
        """Roofline of the leg's dominant kernel: executed FP32 FLOP per
        dispatch (rocprof VALU counters of the same workload) over the
        event-timed call, and the reference formulation's algorithmic FLOP
        over the same time."""
        fm = flop_model(mp_leg, res, T, w)
        r1_name, a1_name = kernel_names(info)
        if r1_avg >= a1_avg:
            dom, alg, dur = r1_name, fm["f_r1"] * pairs, r1_avg
        else:
            dom, alg, dur = a1_name, fm["f_a1"] * pairs, a1_avg
        alg_rate = alg / (dur * 1e-3)
        ex = executed.get(dom)
        ex_rate = ex / (dur * 1e-3) if ex else None
        issue = None
        instr = issue_p.get(dom)
        if instr:
            # transcendental instructions (v_sqrt, v_exp, v_rcp) run in their own
            # pipeline beside the other VALU instructions and are not counted
            trans = issue_p.get("trans", {}).get(dom, 0.0)
            issue_ms = (instr - trans) * 4 / (1024 * SIMD_CLOCK_HZ) * 1e3
            issue = {"wave_instructions": instr, "transcendental": trans,
                     "issue_ms": issue_ms, "call_ms": dur, "frac": issue_ms / dur}
        return fm, {
            "bound": "valu", "kernel": dom,
            "achieved": ex_rate / 1e12 if ex_rate else None,
            "peak": PEAK_FP32 / 1e12, "unit": "TFLOP/s",
            "frac": ex_rate / PEAK_FP32 if ex_rate else None,
            "traffic": traffic_p.get(dom),
            "executed_flop_per_launch": ex,
            "algorithmic_flop_per_launch": alg,
            "algorithmic_achieved": alg_rate / 1e12,
            "algorithmic_frac": alg_rate / PEAK_FP32,
            "valu_issue": issue,
            "note": "achieved/frac: FP32 FLOP the kernel executes per launch (rocprof "
                    "SQ_INSTS_VALU_{ADD,MUL,FMA,TRANS}_F32, 64 lanes x (ADD + MUL + 2 FMA + "
                    "TRANS), profiles/valu_flops.json) over the event-timed call; "
                    "algorithmic_*: the reference formulation's FLOP (18 per active R1 "
                    "sample, 17 per A1 window bin + 23 per pair) over the same time, which "
                    "the mirror kernels exceed by sharing one decision among four samples; "
                    "valu_issue: wave VALU instructions other than transcendental ones x 4 "
                    "cycles at 2.3 GHz (the measured clock) over 1024 SIMDs, over the call"}

    wall, step_ms, r1_ms, a1_ms, info = run_leg(mp_np, args.steps, args.warmup)
    value = pairs * args.steps / wall
    r1_avg, a1_avg = float(r1_ms.mean()), float(a1_ms.mean())
    fm, roof = roofline(mp_np, info, r1_avg, a1_avg)
    r1_name, a1_name = kernel_names(info)

    # the same step on a scan without mirror symmetry (general kernels)
    mp_g = general_scan(mp_np)
    g_steps = max(1, min(args.steps, 10))
    g_wall, g_step, g_r1, g_a1, g_info = run_leg(mp_g, g_steps, 1)
    _, g_roof = roofline(mp_g, g_info, float(g_r1.mean()), float(g_a1.mean()))
    g_r1n, g_a1n = kernel_names(g_info)
    general = {"value": pairs * g_steps / g_wall, "unit": "pairs/s", "steps": g_steps,
               "ms_per_step": g_wall * 1e3 / g_steps,
               "p50_ms": float(np.percentile(g_step, 50)),
               "p99_ms": float(np.percentile(g_step, 99)),
               "kernels_ms": {"r1_forward_avg": float(g_r1.mean()),
                              "a1_backward_avg": float(g_a1.mean()),
                              "r1_kernel": g_r1n, "a1_kernel": g_a1n},
               "roofline": g_roof,
               "scan": "64x64 wall grid with point 5 moved one ulp in x (no mirror pairs)"}

    if args.copy_probe:
        src = torch.ones(64 * 1024 * 1024, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(mp_np, vol_np, T, w)

    line = {
        "metric": "render+backproject voxel·scanpoint pairs/s fwd+bwd at 128³/64²/1024; p50/p99",
        "value": value,
        "unit": "pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "p50_ms": float(np.percentile(step_ms, 50)),
        "p99_ms": float(np.percentile(step_ms, 99)),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded 3% sparse U(0,1) albedo, make_wall_grid scan)",
        "config": {"workload": "confocal R1 render + A1 albedo gradient, 128^3 voxels, "
                               "64x64 scan on [-1,1]^2 m, 1024 bins of 0.005 m round trip, "
                               "box x,y in [-1,1], z in [0.5,2.5] m",
                   "voxels": res ** 3, "scan_points": N, "bins": T,
                   "pairs_per_step": pairs},
        "kernels_ms": {"r1_forward_avg": r1_avg, "a1_backward_avg": a1_avg,
                       "r1_kernel": r1_name, "a1_kernel": a1_name},
        "flop_model": fm,
        "roofline": roof,
        "general_geometry": general,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


# ---------------------------------------------------------------------------
# Serving mix
# ---------------------------------------------------------------------------
SERVE = dict(seed=6, sides=(32, 64, 128), bins=(512, 1024, 2048), res=(64, 128, 192),
             occupancy=(0.01, 0.05), path_length=5.12, device_budget=16 << 30,
             host_budget=6 << 30, loads=(0.5, 1.0, 2.0))


def serve_requests(torch, dev, count):
    """Seeded mix: render or backproject (50/50); scan side, bins and volume
    side uniform over SERVE's sets; occupancy U(1 %, 5 %); one-way bin width
    2.56 m / T (5.12 m of round-trip path over the bins).  Inputs are put in
    pinned host memory, as a service receives them.  A backprojection's input
    is the noisy transient of its own sparse albedo."""
    from tests.cases import sparse_albedo, wall_grid
    from totrilib_amd.serve import BackprojectRequest, RenderRequest
    from totrilib_amd.totri_cpp.render import volume_confocal_forward
    rng = np.random.default_rng(SERVE["seed"])
    reqs = []
    grids = {}
    for k in range(count):
        kind = "render" if rng.random() < 0.5 else "backproject"
        n = int(rng.choice(SERVE["sides"]))
        T = int(rng.choice(SERVE["bins"]))
        r = int(rng.choice(SERVE["res"]))
        p = float(rng.uniform(*SERVE["occupancy"]))
        w = SERVE["path_length"] / 2 / T
        if n not in grids:
            grids[n] = torch.from_numpy(wall_grid(n)).pin_memory()
        mp = grids[n]
        vol = torch.from_numpy(sparse_albedo((r, r, r), p, 1000 + k))
        if kind == "render":
            reqs.append(RenderRequest(vol.pin_memory(), mp, ORIGIN, SIZE, T, w, 0.0))
        else:
            tr = torch.empty((n * n, T), device=dev)
            volume_confocal_forward(vol.to(dev), mp.to(dev), tr, ORIGIN, SIZE, w, 0.0)
            g = torch.Generator(device=dev).manual_seed(2000 + k)
            tr += 1e-3 * tr.max() * torch.randn(tr.shape, device=dev, generator=g)
            reqs.append(BackprojectRequest(tr.cpu().pin_memory(), mp, ORIGIN, SIZE,
                                           (r, r, r), w, 0.0))
    torch.cuda.synchronize()
    return reqs


def per_request_baseline(torch, dev, reqs):
    """Plain per-request launches of the same kernels: allocate, copy in,
    launch, synchronise, copy out, one request at a time."""
    from totrilib_amd.serve import RenderRequest
    from totrilib_amd.totri_cpp.render import (volume_confocal_forward,
                                               volume_confocal_grad_volume)
    lat = []
    t0 = time.perf_counter()
    for r in reqs:
        ts = time.perf_counter()
        mp = r.measurement_points.to(dev)
        if isinstance(r, RenderRequest):
            out = torch.empty(r.output_shape(), device=dev)
            volume_confocal_forward(r.volume.to(dev), mp, out, r.volume_bbox_origin,
                                    r.volume_bbox_size, r.bin_width, r.bin_offset)
        else:
            out = torch.empty(r.output_shape(), device=dev)
            volume_confocal_grad_volume(out, mp, r.transient.to(dev), r.volume_bbox_origin,
                                        r.volume_bbox_size, r.bin_width, r.bin_offset)
        torch.cuda.synchronize()
        out.cpu()
        lat.append(time.perf_counter() - ts)
    return time.perf_counter() - t0, np.array(lat)


def serve_main(args):
    import torch
    from totrilib_amd.serve import ServingEngine, request_pairs

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    reqs = serve_requests(torch, dev, args.requests)
    pairs = float(sum(request_pairs(r) for r in reqs))
    print(f"[serve] {len(reqs)} requests, {pairs:.3e} pairs", file=sys.stderr, flush=True)

    # warm both paths (allocator caches, graph capture) on small requests
    small = [r for r in reqs if request_pairs(r) <= 64 ** 3 * 32 * 32][:4] or reqs[:1]
    per_request_baseline(torch, dev, small)
    eng = ServingEngine(device_budget=SERVE["device_budget"], host_budget=SERVE["host_budget"],
                        lanes=2)
    eng.run(small, [0.0] * len(small), on_done=lambda tk: None)
    # the same engine with arrival-order admission, for comparison
    eng_fifo = ServingEngine(device_budget=SERVE["device_budget"],
                             host_budget=SERVE["host_budget"], lanes=2, policy="fifo")
    eng_fifo.run(small, [0.0] * len(small), on_done=lambda tk: None)

    base_t, base_lat = per_request_baseline(torch, dev, reqs)
    print(f"[serve] per-request baseline {base_t:.2f} s", file=sys.stderr, flush=True)
    # the smaller and larger halves of the mix by pairs, for per-size latency
    cost = np.array([request_pairs(r) for r in reqs], dtype=np.float64)
    large = cost > np.median(cost)
    service = np.asarray(base_lat, dtype=np.float64) * 1e3

    def serve(arrivals, eng=eng):
        steps0 = eng.steps
        jobs0 = len(eng.jobs_per_step)
        t0 = time.perf_counter()
        tks = eng.run(reqs, list(arrivals), on_done=lambda tk: None)
        t_end = max(tk.t_done for tk in tks)
        lat = np.array([tk.latency for tk in tks]) * 1e3
        assert all(tk.error is None for tk in tks), [tk.error for tk in tks if tk.error]
        jps = eng.jobs_per_step[jobs0:]
        slow = lat / service  # latency over the request's own service time
        return dict(span=t_end - t0, p50_ms=float(np.percentile(lat, 50)),
                    p99_ms=float(np.percentile(lat, 99)), steps=eng.steps - steps0,
                    mean_jobs_per_step=float(np.mean(jps)) if jps else 0.0,
                    small_p50_ms=float(np.percentile(lat[~large], 50)),
                    small_p99_ms=float(np.percentile(lat[~large], 99)),
                    large_p50_ms=float(np.percentile(lat[large], 50)),
                    large_p99_ms=float(np.percentile(lat[large], 99)),
                    p99_request_large=bool(large[int(np.argmax(lat))]),
                    slowdown_p50=float(np.percentile(slow, 50)),
                    slowdown_p99=float(np.percentile(slow, 99)))

    sat = serve(np.zeros(len(reqs)))
    print(f"[serve] saturated {sat['span']:.2f} s", file=sys.stderr, flush=True)
    rate_sat = len(reqs) / sat["span"]
    rng = np.random.default_rng(SERVE["seed"] + 1)
    rates = []
    for load in SERVE["loads"]:
        lam = load * rate_sat
        arr = np.cumsum(rng.exponential(1.0 / lam, len(reqs)))
        arr -= arr[0]
        r = serve(arr)
        rf = serve(arr, eng_fifo)
        rates.append(dict(load=load, rate_rps=lam, p50_ms=r["p50_ms"], p99_ms=r["p99_ms"],
                          pairs_per_s=pairs / r["span"], steps=r["steps"],
                          mean_jobs_per_step=r["mean_jobs_per_step"],
                          fifo_p50_ms=rf["p50_ms"], fifo_p99_ms=rf["p99_ms"],
                          fifo_pairs_per_s=pairs / rf["span"],
                          by_size={k: r[k] for k in ("small_p50_ms", "small_p99_ms",
                                                     "large_p50_ms", "large_p99_ms",
                                                     "p99_request_large", "slowdown_p50",
                                                     "slowdown_p99")},
                          fifo_by_size={k: rf[k] for k in ("small_p50_ms", "small_p99_ms",
                                                           "large_p50_ms", "large_p99_ms",
                                                           "slowdown_p50", "slowdown_p99")}))
        print(f"[serve] load {load}: p50 {r['p50_ms']:.1f} ms p99 {r['p99_ms']:.1f} ms "
              f"(fifo p50 {rf['p50_ms']:.1f} p99 {rf['p99_ms']:.1f})",
              file=sys.stderr, flush=True)

    value = pairs / sat["span"]
    base_value = pairs / base_t
    line = {
        "metric": "serving: voxel·scanpoint pairs/s of a seeded render/backproject request "
                  "mix through the continuous-batching engine; p50/p99 latency",
        "value": value,
        "unit": "pairs/s",
        "n_gpus": 1,
        "steps": sat["steps"],
        "warmup": len(small),
        "ms_per_step": sat["span"] * 1e3 / max(sat["steps"], 1),
        "p50_ms": sat["p50_ms"],
        "p99_ms": sat["p99_ms"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded sparse albedo, noisy rendered transients, make_wall_grid scans)",
        "config": {"workload": f"{len(reqs)} requests, 50/50 render (R1) / backproject (A1); "
                               "scan 32^2|64^2|128^2, bins 512|1024|2048 over 5.12 m round "
                               "trip, volume 64^3|128^3|192^3, occupancy U(1%,5%); inputs "
                               "in pinned host memory, results copied back",
                   "requests": len(reqs), "pairs": pairs,
                   "device_budget_bytes": SERVE["device_budget"], "lanes": 2,
                   "policy": "sjf (fifo legs in rates for comparison)"},
        "saturated": sat,
        "per_request_baseline": {"value": base_value, "unit": "pairs/s",
                                 "p50_ms": float(np.percentile(base_lat, 50) * 1e3),
                                 "p99_ms": float(np.percentile(base_lat, 99) * 1e3),
                                 "how": "one request at a time: torch allocations, H2D, "
                                        "single-request entry point, synchronize, D2H"},
        "speedup_vs_per_request": value / base_value,
        "rates": rates,
    }
    print(json.dumps(line))


# ---------------------------------------------------------------------------
# The other operations of the path, one at a time
# ---------------------------------------------------------------------------
def kernels_main(args):
    """Times each remaining operation at a stated size (HIP events around the
    call on its stream, after a warm-up call) and reports its algorithmic
    rate: FLOP per unit in the rocprof-compute convention, counted from the
    reference formulas, and the bytes it must move."""
    import torch
    from tests.cases import sparse_albedo, wall_grid
    from totrilib_amd.reconstruct import (BackprojectionConfocalVelten,
                                          BackprojectionExhaustiveArellano,
                                          BackprojectionExhaustiveVelten)
    from totrilib_amd.render import VolumeRenderConfocal
    from totrilib_amd.render.volume import VolumeRenderExhaustive
    from totrilib_amd.types import BinDef, VolumeDef
    from totrilib_amd.util import Rbf2VolumeGaussian

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    gen = torch.Generator(device=dev).manual_seed(9)
    reps = max(args.steps // 4, 3)

    def timeit(fn):
        fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps

    rows = []

    def row(op, kernel, workload, ms, units, unit, flop_per_unit, nbytes):
        flops = units * flop_per_unit
        rows.append(dict(op=op, kernel=kernel, workload=workload, ms=ms, units=units, unit=unit,
                         units_per_s=units / (ms * 1e-3), flop_per_unit=flop_per_unit,
                         tflops=flops / (ms * 1e-3) / 1e12,
                         frac_fp32=flops / (ms * 1e-3) / PEAK_FP32,
                         gbps=nbytes / (ms * 1e-3) / 1e9))
        print(f"[kernels] {op}: {ms:.3f} ms", file=sys.stderr, flush=True)

    box = VolumeDef((-1.0, -1.0, 0.5), (1.0, 1.0, 2.5), (64, 64, 64))
    # A2: Velten confocal, 64^3 voxels, 32x32 scan, 512 bins of 0.01 m round trip
    n, T, R = 32, 512, 64
    S = n * n
    scan = torch.from_numpy(wall_grid(n)).T.contiguous()[None].to(dev)        # [1, S, 3]
    tr = torch.rand((1, T, S), device=dev, generator=gen)
    ms = timeit(lambda: BackprojectionConfocalVelten.apply(tr, scan, box, BinDef(0.01, 0.0, T)))
    row("A2 Velten confocal backprojection", "velten_confocal_kernel",
        "64^3 voxels, 32x32 scan, 512 bins", ms, R ** 3 * S, "voxel-scan pairs", 14,
        4 * (T * S + R ** 3))
    # The fitting step after it: blob proposals from the positive residual (A2 at half
    # resolution, second-derivative filter, normalise, square, 5 draws)
    from totrilib_amd.fitting import sample_blob_positions, volume_error_map
    model = 0.9 * tr[0]
    scan0 = scan[0]
    gdraw = torch.Generator(device=dev).manual_seed(11)
    ms = timeit(lambda: sample_blob_positions(
        volume_error_map(tr[0], model, scan0, box, BinDef(0.01, 0.0, T)), box, 5, gdraw))
    Rh = R // 2
    row("Blob proposals (fitting step)", "velten_confocal_kernel",
        "32^3 half-resolution map of a 64^3 box, 32x32 scan, 512 bins, 5 draws", ms,
        Rh ** 3 * S, "voxel-scan pairs", 14, 4 * (2 * T * S + 4 * Rh ** 3))
    # A3: Velten exhaustive, 64^3, 16x16 scan points x 16 laser points, 512 bins
    Sx, L = 256, 16
    sp = torch.from_numpy(wall_grid(16)).T.contiguous()[None].to(dev)         # [1, 256, 3]
    lp = torch.from_numpy(wall_grid(4)).T.contiguous()[None].to(dev)          # [1, 16, 3]
    trx = torch.rand((1, T, Sx, L), device=dev, generator=gen)
    ms = timeit(lambda: BackprojectionExhaustiveVelten.apply(trx, sp, lp, box, BinDef(0.01, 0.0, T)))
    row("A3 Velten exhaustive backprojection", "velten_exhaustive_kernel",
        "64^3 voxels, 256 scan x 16 laser points, 512 bins", ms, R ** 3 * Sx * L,
        "voxel-scan-laser triples", 8, 4 * (T * Sx * L + R ** 3))
    # A4: Arellano, same sizes, 5 % of the bins nonzero
    tra = trx * (torch.rand(trx.shape, device=dev, generator=gen) < 0.05)
    nz = int(torch.count_nonzero(tra))
    sp3, lp3 = sp.transpose(1, 2).contiguous(), lp.transpose(1, 2).contiguous()
    ms = timeit(lambda: BackprojectionExhaustiveArellano.apply(tra, sp3, lp3, box, BinDef(0.01, 0.0, T)))
    row("A4 Arellano backprojection", "arellano_kernel",
        "64^3 voxels, 256 scan x 16 laser points, 512 bins, 5% nonzero", ms, 64 * 64 * nz,
        "column-bin pairs (nonzero bins)", 30, 4 * (T * Sx * L + R ** 3))
    # R2 + A5: exhaustive render and its adjoint, 64^3 (5 % occupied), 256 x 16, 512 bins
    vol = torch.from_numpy(sparse_albedo((R, R, R), 0.05, 3)).permute(2, 1, 0).contiguous()[None].to(dev)
    nnz = int(torch.count_nonzero(vol))
    v = vol.clone().requires_grad_(True)
    ms = timeit(lambda: VolumeRenderExhaustive.apply(vol, sp3, lp3, box, BinDef(0.01, 0.0, T)))
    row("R2 exhaustive render", "r2_forward_kernel",
        "64^3 voxels (5% occupied), 256 scan x 16 laser points, 512 bins", ms, nnz * Sx * L,
        "occupied-voxel-scan-laser triples", 40, 4 * (T * Sx * L + R ** 3))
    out = VolumeRenderExhaustive.apply(v, sp3, lp3, box, BinDef(0.01, 0.0, T))
    G = torch.randn(out.shape, device=dev, generator=gen)
    ms = timeit(lambda: torch.autograd.grad(out, v, G, retain_graph=True))
    row("A5 exhaustive render adjoint", "r2_adjoint_kernel",
        "64^3 voxels, 256 scan x 16 laser points, 512 bins", ms, R ** 3 * Sx * L,
        "voxel-scan-laser triples", 40, 4 * (T * Sx * L + R ** 3))
    # S1 + S2: 20,000 RBFs -> 128^3, truncated at 3 sigma
    M, RR = 20000, 128
    rbf = torch.cat([torch.rand((1, M, 2), device=dev, generator=gen) * 2 - 1,
                     torch.rand((1, M, 1), device=dev, generator=gen) * 2 + 0.5,
                     torch.rand((1, M, 1), device=dev, generator=gen) * 0.04 + 0.01,
                     torch.rand((1, M, 1), device=dev, generator=gen)], 2)
    dv = 2.0 / RR
    vd = VolumeDef((-1 + dv / 2, -1 + dv / 2, 0.5 + dv / 2), (1 - dv / 2, 1 - dv / 2, 2.5 - dv / 2),
                   (RR, RR, RR))
    sig = rbf[0, :, 3].double()
    in_support = float(torch.sum(4.0 / 3.0 * np.pi * (3 * sig) ** 3) / dv ** 3)
    ms = timeit(lambda: Rbf2VolumeGaussian.apply(rbf, vd, 3.0))
    row("S1 RBF splat", "splat_bin_kernel, splat_forward_kernel",
        "20,000 RBFs -> 128^3, 3 sigma", ms, in_support,
        "voxel-RBF pairs within 3 sigma", 14, 4 * (2 * RR ** 3 + 5 * M))
    r = rbf.clone().requires_grad_(True)
    V = Rbf2VolumeGaussian.apply(r, vd, 3.0)
    GV = torch.randn(V.shape, device=dev, generator=gen)
    ms = timeit(lambda: torch.autograd.grad(V, r, GV, retain_graph=True))
    row("S2 RBF splat gradient", "splat_grad_kernel", "20,000 RBFs -> 128^3, 3 sigma", ms,
        in_support, "voxel-RBF pairs within 3 sigma", 25, 4 * (4 * RR ** 3 + 10 * M))
    # the reference's untruncated splat (truncation 0): every RBF at every voxel
    ms = timeit(lambda: Rbf2VolumeGaussian.apply(rbf, vd, 0.0))
    row("S1 RBF splat, untruncated", "splat_forward_kernel", "20,000 RBFs -> 128^3, no truncation",
        ms, float(M) * RR ** 3, "voxel-RBF pairs", 14, 4 * (2 * RR ** 3 + 5 * M))
    r0 = rbf.clone().requires_grad_(True)
    V0 = Rbf2VolumeGaussian.apply(r0, vd, 0.0)
    ms = timeit(lambda: torch.autograd.grad(V0, r0, GV, retain_graph=True))
    row("S2 RBF splat gradient, untruncated", "splat_grad_all_kernel",
        "20,000 RBFs -> 128^3, no truncation", ms, float(M) * RR ** 3, "voxel-RBF pairs", 25,
        4 * (4 * RR ** 3 + 10 * M))
    # A6: scan-point gradient, 128^3 / 64x64 / 1024 bins
    mp = torch.from_numpy(wall_grid(64)).to(dev).requires_grad_(True)
    alb = torch.from_numpy(sparse_albedo((128, 128, 128), 0.03, 1)).to(dev)
    w = WORKLOAD["w_round_trip"] / 2
    trc = VolumeRenderConfocal.apply(alb, mp, ORIGIN, SIZE, 1024, w, 0.0)
    Gc = torch.randn(trc.shape, device=dev, generator=gen)
    ms = timeit(lambda: torch.autograd.grad(trc, mp, Gc, retain_graph=True))
    fm = flop_model(mp.detach().cpu().numpy(), 128, 1024, w)
    row("A6 scan-point gradient", "a6_sym_kernel", "128^3 voxels, 64x64 scan, 1024 bins", ms,
        fm["a_fwd"] * 128 ** 3 * 4096, "active samples", 30, 4 * (2 * 4096 * 1024 + 128 ** 3))
    # The other A1 / R1 / A6 paths on the same workload: scans with x-pairs
    # only (the grid moved 0.01 m in y, so no y-mirrors) and without mirror
    # pairs (one point moved by an ulp)
    from totrilib_amd.totri_cpp.render import (volume_confocal_forward,
                                               volume_confocal_grad_measurement_points,
                                               volume_confocal_grad_volume)
    g1 = torch.empty(128, 128, 128, device=dev)
    tr1 = torch.empty(4096, 1024, device=dev)
    gm1 = torch.empty(3, 4096, device=dev)
    Gd = Gc.detach().contiguous()
    mp_x = wall_grid(64)
    mp_x[1] += np.float32(0.01)
    mp_g = wall_grid(64)
    mp_g[0, 5] = np.nextafter(mp_g[0, 5], np.float32(2))
    pairs = 128.0 ** 3 * 4096
    for name, kern, mpn in (("A1 albedo gradient, x-pair scan", "a1_pair_kernel", mp_x),
                            ("A1 albedo gradient, general scan", "a1_row_kernel", mp_g)):
        m1 = torch.from_numpy(mpn).to(dev)
        ms = timeit(lambda: volume_confocal_grad_volume(g1, m1, Gd, ORIGIN, SIZE, w, 0.0))
        fm1 = flop_model(mpn, 128, 1024, w)
        row(name, kern, "128^3 voxels, 64x64 scan, 1024 bins", ms, pairs,
            "voxel-scan pairs", fm1["f_a1"], 4 * (4096 * 1024 + 128 ** 3))
    mx = torch.from_numpy(mp_x).to(dev)
    fmx = flop_model(mp_x, 128, 1024, w)
    ms = timeit(lambda: volume_confocal_forward(alb, mx, tr1, ORIGIN, SIZE, w, 0.0))
    row("R1 render, x-pair scan", "r1_pair_kernel", "128^3 voxels, 64x64 scan, 1024 bins",
        ms, pairs, "voxel-scan pairs", fmx["f_r1"], 4 * (4096 * 1024 + 128 ** 3))
    m1 = torch.from_numpy(mp_g).to(dev)
    fm1 = flop_model(mp_g, 128, 1024, w)
    ms = timeit(lambda: volume_confocal_forward(alb, m1, tr1, ORIGIN, SIZE, w, 0.0))
    row("R1 render, general scan", "r1_forward_kernel", "128^3 voxels, 64x64 scan, 1024 bins",
        ms, pairs, "voxel-scan pairs", fm1["f_r1"], 4 * (4096 * 1024 + 128 ** 3))
    ms = timeit(lambda: volume_confocal_grad_measurement_points(gm1, alb, m1, Gd, ORIGIN, SIZE,
                                                                 w, 0.0))
    row("A6 scan-point gradient, general scan", "a6_kernel", "128^3 voxels, 64x64 scan, 1024 bins",
        ms, fm1["a_fwd"] * 128 ** 3 * 4096, "active samples", 30,
        4 * (2 * 4096 * 1024 + 128 ** 3))
    ms = timeit(lambda: volume_confocal_grad_measurement_points(gm1, alb, mx, Gd, ORIGIN, SIZE,
                                                                 w, 0.0))
    row("A6 scan-point gradient, x-pair scan", "a6_pair_kernel",
        "128^3 voxels, 64x64 scan, 1024 bins", ms, fmx["a_fwd"] * 128 ** 3 * 4096,
        "active samples", 30, 4 * (2 * 4096 * 1024 + 128 ** 3))
    # f-k migration: 1024 bins on a 64 x 64 wall; the spectrum is [2048, 64, 64] complex
    from totrilib_amd import _lib
    from totrilib_amd.reconstruct import FkMigration
    from totrilib_amd.reconstruct.fk_migration import _stolt
    Df, Hf, Wf = 1024, 64, 64
    npts = 2 * Df * Hf * Wf
    trf = torch.rand((1, Df, Hf, Wf), device=dev, generator=gen)
    fkm = FkMigration(Df * 0.0025, 2.0, 2.0)
    ms = timeit(lambda: fkm(trf))
    # FLOP: the two 3-D FFTs at 5 N log2 N each, about 80 per Stolt element.  Bytes per
    # (bin, wall point) the five kernels must move (W/2+1 = 33 of 64 columns one-sided):
    # slices 4 + 4.1, depth 4.1 + 8.3, Stolt 8.3 + 8, inverse depth 8 + 8, slices 8 + 4
    row("f-k migration", "fk_depth_fwd_kernel, fk_depth_inv_kernel, fk_stolt_kernel",
        "1024 bins, 64x64 wall, spectrum 2048x64x64", ms, npts, "spectrum points",
        10 * np.log2(npts) + 80, 65 * Df * Hf * Wf)
    spec = torch.randn((1, 2 * Df, Hf, Wf // 2 + 1), dtype=torch.complex64, device=dev,
                       generator=gen)
    gout = torch.zeros((2 * Df, Hf, Wf), dtype=torch.complex64, device=dev)
    ms = timeit(lambda: _stolt(spec, gout, Df * 0.0025, 2.0, 2.0,
                               _lib.TOTRI_FK_UNSHIFTED_ONESIDED, True))
    row("f-k Stolt interpolation", "fk_stolt_kernel",
        "one-sided spectrum 2048x64x33 -> 2048x64x64, z > 0 half written", ms, npts // 2,
        "output points", 80, 16 * Df * Hf * Wf)
    # The large solve at one GPU's share: 256^3 volume, 2048 bins of 0.0025 m round trip,
    # 8192 of the 256x256 scan's points (rows 0-15 and 240-255, so the share keeps its
    # mirror quads, as a scan-point sharding would), FP32 / BF16 / FP16 transient storage
    from totrilib_amd.totri_cpp.render import (volume_confocal_forward,
                                               volume_confocal_grad_volume)
    full = wall_grid(256)
    rows_kept = list(range(16)) + list(range(240, 256))
    idx = np.concatenate([np.arange(r * 256, (r + 1) * 256) for r in rows_kept])
    mp4_np = np.ascontiguousarray(full[:, idx])
    mp4 = torch.from_numpy(mp4_np).to(dev)
    vol4 = torch.from_numpy(sparse_albedo((256, 256, 256), 0.01, 5)).to(dev)
    T4, w4, N4 = 2048, 0.00125, mp4_np.shape[1]
    pairs4 = 256 ** 3 * N4
    fm4 = flop_model(mp4_np[:, ::16], 256, T4, w4)    # per-pair averages from 512 points
    gvol = torch.empty((256, 256, 256), device=dev)
    for dt, name in ((torch.float32, "FP32"), (torch.bfloat16, "BF16"), (torch.float16, "FP16")):
        tr4 = torch.empty((N4, T4), dtype=dt, device=dev)
        G4 = torch.randn((N4, T4), device=dev, generator=gen).to(dt)
        ms_r = timeit(lambda: volume_confocal_forward(vol4, mp4, tr4, ORIGIN, SIZE, w4, 0.0))
        ms_a = timeit(lambda: volume_confocal_grad_volume(gvol, mp4, G4, ORIGIN, SIZE, w4, 0.0))
        esz = 4 if dt == torch.float32 else 2
        row(f"Large solve share, R1 + A1, {name} transient", "r1_sym_kernel, a1_quad_kernel",
            "256^3 voxels, 8192 of 256x256 scan points, 2048 bins", ms_r + ms_a, pairs4,
            "voxel-scan pairs (render + backprojection)", fm4["f_pair"],
            2 * (4 * 256 ** 3 + esz * N4 * T4))
    line = {"metric": "per-operation time and algorithmic rate", "value": None, "unit": "ms",
            "n_gpus": 1, "steps": reps, "warmup": 1, "higher_is_better": False,
            "dtype": "f32", "data": "synthetic (seeded)", "rows": rows,
            "peak_fp32_tflops": PEAK_FP32 / 1e12, "peak_hbm_gbps": PEAK_HBM / 1e9}
    print(json.dumps(line))


if __name__ == "__main__":
    main()
