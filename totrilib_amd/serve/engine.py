"""Continuous-batching request loop for confocal renders and backprojections.

Requests are the two operations of VolumeRenderConfocal (reference
totri/render/volume.py:12-76): a render R1 (`forward`: albedo volume ->
transient) and a backprojection A1 (`apply_transposed`: transient -> volume),
each with its own scan, bin count and volume size.

How a request moves through the engine:

* submit() puts it in a queue.
* Admission (every poll()) takes queued requests in priority order while
  their device memory -- input copies, output, scratch -- fits in the
  fixed-size HBM arena, and their host result buffer fits in the pinned
  arena.  Inputs that already live on the device are used in place
  (resident data).  The default policy, "sjf", orders by estimated service
  time minus time waited (shortest job first with aging, so a large request
  overtakes once it has waited longer than its extra cost); "fifo" keeps
  arrival order.
* A step starts whenever a lane is idle and admitted requests wait: up to
  TOTRI_BATCH_MAX_JOBS of them, in the same order (under "sjf" only those at
  most `group_ratio` times as costly as the first, so a short job does not
  wait for a long one in the same step), are copied in (on the lane's
  stream), run by
  one batched step of persistent kernels (totri_confocal_batch_run), and
  copied out into pinned host memory.  A lane's step is a HIP graph captured
  once and replayed for every step, since its launches do not depend on the
  step's contents (the kernels read a device-side job table that the replay
  uploads first).
* With two or more lanes, a step is queued on the device while the previous
  one runs, so requests that arrive meanwhile start as soon as the previous
  step's work drains, and copies overlap compute.
* poll() retires steps whose completion event has fired: their requests get
  their result and completion time, and their memory returns to the arenas
  (the result's pinned block when the caller releases the ticket).

Every job's result is bit-identical to the single-request entry points on the
same inputs, whatever shares its step (same plan, same per-item code, fixed
summation order).
"""
from __future__ import annotations

import ctypes
import dataclasses
import math
import time
from collections import deque
from typing import Deque, List, Optional, Sequence, Tuple

import torch

from totrilib_amd import _lib
from totrilib_amd.serve.arena import TensorArena

_ALIGN = 256


@dataclasses.dataclass
class RenderRequest:
    """R1: volume [X, Y, Z] albedo -> transient [N, num_bins] (FP32)."""
    volume: torch.Tensor
    measurement_points: torch.Tensor
    volume_bbox_origin: Tuple[float, float, float]
    volume_bbox_size: Tuple[float, float, float]
    bin_num: int
    bin_width: float
    bin_offset: float = 0.0

    kind = _lib.TOTRI_JOB_RENDER

    def dims(self) -> Tuple[int, int, int, int, int]:
        X, Y, Z = self.volume.shape
        return X, Y, Z, self.measurement_points.shape[1], int(self.bin_num)

    def output_shape(self) -> Tuple[int, ...]:
        X, Y, Z, N, T = self.dims()
        return (N, T)


@dataclasses.dataclass
class BackprojectRequest:
    """A1: transient [N, T] -> volume [X, Y, Z] = R1^T transient (FP32)."""
    transient: torch.Tensor
    measurement_points: torch.Tensor
    volume_bbox_origin: Tuple[float, float, float]
    volume_bbox_size: Tuple[float, float, float]
    volume_resolution: Tuple[int, int, int]
    bin_width: float
    bin_offset: float = 0.0

    kind = _lib.TOTRI_JOB_BACKPROJECT

    def dims(self) -> Tuple[int, int, int, int, int]:
        X, Y, Z = (int(v) for v in self.volume_resolution)
        N, T = self.transient.shape
        return X, Y, Z, N, T

    def output_shape(self) -> Tuple[int, ...]:
        return tuple(int(v) for v in self.volume_resolution)


def request_pairs(req) -> int:
    """Voxel x scan-point pairs of a request (the throughput unit)."""
    X, Y, Z, N, _ = req.dims()
    return X * Y * Z * N


# Nominal rate for the shortest-job-first key, in voxel x scan-point x bin
# triples per second: the active samples per pair grow with the bin count
# for a fixed path length.  The key compares this estimate with time waited.
_TRIPLES_PER_S = 8e14


def request_cost(req) -> float:
    """Estimated service time of a request in seconds (ordering only)."""
    X, Y, Z, N, T = req.dims()
    return X * Y * Z * N * T / _TRIPLES_PER_S


@dataclasses.dataclass
class Ticket:
    """Handle of a submitted request."""
    request: object
    id: int
    t_submit: float
    t_admit: Optional[float] = None
    t_done: Optional[float] = None
    error: Optional[str] = None
    cost: float = 0.0
    _host_out: Optional[torch.Tensor] = None
    _release: Optional[object] = None
    # device-resident inputs made contiguous at submit, and an event recorded
    # on the submitting stream after them: the step's stream waits on it
    _resident: Optional[dict] = None
    _ready: Optional[torch.cuda.Event] = None

    @property
    def done(self) -> bool:
        return self.t_done is not None

    @property
    def latency(self) -> Optional[float]:
        return None if self.t_done is None else self.t_done - self.t_submit

    def result(self) -> torch.Tensor:
        """The output in pinned host memory (valid until release())."""
        if self.error is not None:
            raise RuntimeError(self.error)
        if self._host_out is None:
            raise RuntimeError("request has not completed")
        return self._host_out

    def release(self) -> None:
        """Return the result's pinned block to the engine."""
        if self._release is not None:
            self._release()
            self._release = None
            self._host_out = None


class _Job:
    """An admitted request: its arena blocks and the C job description.  A
    job of `slices` > 1 runs over one or more steps, all on one lane (`lane`,
    set at its first step); `slice` is the next slice to run and `run` the
    number of slices its current step runs."""

    def __init__(self, ticket: Ticket, dev_blocks: List[int], host_block: int):
        self.ticket = ticket
        self.dev_blocks = dev_blocks
        self.host_block = host_block
        self.copies_in: List[Tuple[torch.Tensor, torch.Tensor]] = []
        self.inputs: List[torch.Tensor] = []
        self.out_dev: Optional[torch.Tensor] = None
        self.out_host: Optional[torch.Tensor] = None
        self.cjob = _lib.ConfocalJob()
        self.slices = 1
        self.slice = 0
        self.run = 1
        self.lane: Optional[int] = None

    def step_cost(self) -> float:
        """Estimated service time of one slice."""
        return self.ticket.cost / self.slices

    def remaining_cost(self) -> float:
        return self.ticket.cost * (self.slices - self.slice) / self.slices


class _Lane:
    """A stream with its own step table (pinned host + device) and graph."""

    def __init__(self, device: torch.device, table_bytes: int):
        self.stream = torch.cuda.Stream(device=device)
        self.host_table = torch.zeros(table_bytes, dtype=torch.uint8, pin_memory=True)
        self.dev_table = torch.zeros(table_bytes, dtype=torch.uint8, device=device)
        self.graph: Optional[torch.cuda.CUDAGraph] = None
        self.jobs: List[_Job] = []
        self.event: Optional[torch.cuda.Event] = None
        self.steps = 0


class ServingEngine:
    """Continuous batching of confocal render / backprojection requests.

    device_budget  bytes of HBM for request data and scratch (one allocation)
    host_budget    bytes of pinned host memory for results (one allocation)
    lanes          steps that may be queued on the device at once
    use_graphs     replay each lane's step as a captured HIP graph
    policy         "sjf" (shortest job first with aging) or "fifo"
    group_ratio    under "sjf", a step takes only jobs at most this many times
                   as costly as its first (None: no limit)
    slice_cost     under "sjf", a request estimated to take longer than this
                   (seconds) runs in slices of about this cost over several
                   steps, each a consecutive share of its work items, so that
                   requests arriving meanwhile join the next step instead of
                   waiting for the whole request; results are bit-identical
                   to the unsliced request (None: no slicing)
    greedy_slices  a sliced request whose step would run alone (nothing else
                   admitted or queued, every other lane idle) runs all its
                   remaining slices in that step: an idle lane serves
                   arrivals meanwhile
    """

    def __init__(self, device_budget: int = 16 << 30, host_budget: int = 4 << 30,
                 lanes: int = 2, max_jobs: int = _lib.TOTRI_BATCH_MAX_JOBS,
                 use_graphs: bool = True, device: Optional[torch.device] = None,
                 policy: str = "sjf", group_ratio: Optional[float] = 4.0,
                 slice_cost: Optional[float] = 0.03, greedy_slices: bool = True):
        if policy not in ("sjf", "fifo"):
            raise ValueError(f"policy must be 'sjf' or 'fifo', not {policy!r}")
        if slice_cost is not None and not slice_cost > 0:
            raise ValueError("slice_cost must be positive or None")
        self.policy = policy
        self.group_ratio = group_ratio
        self.slice_cost = slice_cost if policy == "sjf" else None
        self.greedy_slices = greedy_slices
        if not torch.cuda.is_available():
            raise RuntimeError("ServingEngine needs a GPU")
        self.lib = _lib.load()
        self.device = device or torch.device("cuda", torch.cuda.current_device())
        if not 1 <= max_jobs <= _lib.TOTRI_BATCH_MAX_JOBS:
            raise ValueError(f"max_jobs must be in [1, {_lib.TOTRI_BATCH_MAX_JOBS}]")
        self.max_jobs = max_jobs
        self.use_graphs = use_graphs
        self.table_bytes = int(self.lib.totri_confocal_batch_table_bytes())
        self.dev = TensorArena(device_budget, self.device, align=_ALIGN)
        self.host = TensorArena(host_budget, self.device, pinned=True, align=_ALIGN)
        self.lanes = [_Lane(self.device, self.table_bytes) for _ in range(max(1, lanes))]
        self.queue: Deque[Ticket] = deque()
        self.admitted: Deque[_Job] = deque()
        self._next_id = 0
        self.steps = 0
        self.jobs_per_step: List[int] = []

    # -- submission -------------------------------------------------------
    def submit(self, request, t_submit: Optional[float] = None) -> Ticket:
        """Queue a request.  Inputs already on the device are used in place;
        work that produces them on the caller's current stream (of the
        engine's device) before submit() is complete before the request's
        step reads them."""
        self._validate(request)
        tk = Ticket(request, self._next_id, time.perf_counter() if t_submit is None else t_submit)
        tk.cost = request_cost(request)
        data = request.volume if isinstance(request, RenderRequest) else request.transient
        resident = {}
        for name, t in (("measurement_points", request.measurement_points), ("data", data)):
            if t.is_cuda:
                resident[name] = t.contiguous()
        if resident:
            tk._resident = resident
            tk._ready = torch.cuda.Event()
            tk._ready.record(torch.cuda.current_stream(self.device))
        self._next_id += 1
        self.queue.append(tk)
        return tk

    def _validate(self, req) -> None:
        if not isinstance(req, (RenderRequest, BackprojectRequest)):
            raise TypeError(f"unsupported request {type(req)}")
        X, Y, Z, N, T = req.dims()
        mp = req.measurement_points
        if mp.ndim != 2 or mp.shape[0] != 3 or mp.dtype != torch.float32:
            raise ValueError("measurement_points must be a float32 [3, N] tensor")
        data = req.volume if isinstance(req, RenderRequest) else req.transient
        if data.dtype != torch.float32:
            raise TypeError("request data must be float32")
        if isinstance(req, RenderRequest) and data.ndim != 3:
            raise ValueError("volume must be [X, Y, Z]")
        if isinstance(req, BackprojectRequest) and (data.ndim != 2 or data.shape[0] != N):
            raise ValueError("transient must be [N, T] with N = measurement points")
        if min(X, Y, N, T) < 1 or Z < 2:
            raise ValueError("requests need X, Y, N, T >= 1 and Z >= 2")
        for t in (mp, data):
            if t.is_cuda and t.device != self.device:
                raise ValueError(f"tensor on {t.device}, engine on {self.device}")

    # -- admission --------------------------------------------------------
    def _footprint(self, req) -> Tuple[List[int], int, int]:
        """Device block sizes [inputs copied in..., output, scratch], the
        host result size and the scratch size."""
        X, Y, Z, N, T = req.dims()
        cj = _lib.ConfocalJob(kind=req.kind, X=X, Y=Y, Z=Z, N=N, T=T,
                              volume_origin=_lib.f3(req.volume_bbox_origin),
                              volume_size=_lib.f3(req.volume_bbox_size),
                              bin_width=float(req.bin_width), bin_offset=float(req.bin_offset))
        ws = int(self.lib.totri_confocal_job_workspace(ctypes.byref(cj)))
        data = req.volume if isinstance(req, RenderRequest) else req.transient
        sizes = []
        for t in (req.measurement_points, data):
            if not t.is_cuda:
                sizes.append(t.numel() * 4)
        out = math.prod(req.output_shape()) * 4
        sizes.append(out)
        sizes.append(ws)
        return sizes, out, ws

    def _key(self, tk: Ticket, now: float, cost: Optional[float] = None):
        """Priority of a ticket (smaller first): under "sjf" its remaining
        estimated cost less the time it has waited."""
        if self.policy == "fifo":
            return tk.id
        return (tk.cost if cost is None else cost) - (now - tk.t_submit)

    def _admit(self, now: float) -> None:
        if not self.queue:
            return
        taken = set()
        for tk in sorted(self.queue, key=lambda t: self._key(t, now)):
            req = tk.request
            sizes, out_bytes, ws = self._footprint(req)
            if not self.dev.arena.fits(sizes) or out_bytes > self.host.arena.capacity:
                taken.add(tk.id)
                tk.error = "request exceeds the engine's memory budget"
                tk.t_done = now
                continue
            blocks = self.dev.arena.alloc_many(sizes)
            if blocks is None:
                break
            hb = self.host.arena.alloc(out_bytes)
            if hb is None:
                for b in blocks:
                    self.dev.arena.free(b)
                break
            taken.add(tk.id)
            tk.t_admit = now
            self.admitted.append(self._make_job(tk, blocks, hb, ws))
        self.queue = deque(tk for tk in self.queue if tk.id not in taken)

    def _make_job(self, tk: Ticket, blocks: List[int], hb: int, ws: int) -> _Job:
        req = tk.request
        job = _Job(tk, blocks, hb)
        X, Y, Z, N, T = req.dims()
        data = req.volume if isinstance(req, RenderRequest) else req.transient
        k = 0
        placed = []
        for name, t in (("measurement_points", req.measurement_points), ("data", data)):
            if t.is_cuda:
                placed.append(tk._resident[name])
            else:
                dst = self.dev.view(blocks[k], t.shape)
                job.copies_in.append((dst, t))
                placed.append(dst)
                k += 1
        mp_dev, data_dev = placed
        out_shape = req.output_shape()
        job.out_dev = self.dev.view(blocks[k], out_shape)
        job.out_host = self.host.view(hb, out_shape)
        ws_ptr = self.dev.ptr(blocks[k + 1]) if ws > 0 else None
        if self.slice_cost is not None and tk.cost > self.slice_cost:
            job.slices = min(int(math.ceil(tk.cost / self.slice_cost)), 256)
        cj = job.cjob
        cj.kind = req.kind
        cj.slice = 0
        cj.slices = job.slices
        cj.slice_run = 1
        cj.measurement_points = mp_dev.data_ptr()
        if isinstance(req, RenderRequest):
            cj.volume = data_dev.data_ptr()
            cj.transient = job.out_dev.data_ptr()
        else:
            cj.volume = job.out_dev.data_ptr()
            cj.transient = data_dev.data_ptr()
        cj.X, cj.Y, cj.Z, cj.N, cj.T = X, Y, Z, N, T
        cj.volume_origin = _lib.f3(req.volume_bbox_origin)
        cj.volume_size = _lib.f3(req.volume_bbox_size)
        cj.bin_width = float(req.bin_width)
        cj.bin_offset = float(req.bin_offset)
        cj.workspace = ws_ptr
        cj.workspace_bytes = ws
        # keep device-resident inputs alive while the job runs
        job.inputs = placed
        return job

    # -- steps --------------------------------------------------------------
    def _launch(self, lane: _Lane, now: float) -> None:
        li = self.lanes.index(lane)
        jobs = []
        # a sliced job continues on the lane of its first slice
        cands = [j for j in self.admitted if j.lane is None or j.lane == li]
        for j in sorted(cands, key=lambda j: self._key(j.ticket, now, j.remaining_cost())):
            if len(jobs) == self.max_jobs:
                break
            if (self.policy == "sjf" and self.group_ratio is not None and jobs
                    and j.step_cost() > self.group_ratio * jobs[0].step_cost()):
                continue
            jobs.append(j)
        if not jobs:
            return
        chosen = set(id(j) for j in jobs)
        self.admitted = deque(j for j in self.admitted if id(j) not in chosen)
        # A sliced job runs one slice when other work waits or another lane is
        # busy (requests arriving meanwhile then join its next step), and all
        # its remaining slices otherwise (an idle lane serves new arrivals).
        contended = (not self.greedy_slices or bool(self.queue) or bool(self.admitted)
                     or any(l.event is not None for l in self.lanes if l is not lane))
        for j in jobs:
            j.lane = li
            j.run = 1 if contended else j.slices - j.slice
            j.cjob.slice = j.slice
            j.cjob.slice_run = j.run
        arr = (_lib.ConfocalJob * len(jobs))(*[j.cjob for j in jobs])
        rc = self.lib.totri_confocal_batch_prepare(arr, len(jobs), lane.host_table.data_ptr(),
                                                   self.table_bytes)
        _lib.check(rc, "totri_confocal_batch_prepare")
        st = lane.stream
        # the lane's host table may be rewritten only after its last step's
        # upload; the lane is idle here, so that upload has completed.
        with torch.cuda.stream(st):
            for j in jobs:
                if j.ticket._ready is not None:
                    st.wait_event(j.ticket._ready)
            for j in jobs:
                if j.slice == 0:
                    for dst, src in j.copies_in:
                        dst.copy_(src, non_blocking=True)
            if self.use_graphs:
                if lane.graph is None:
                    lane.graph = torch.cuda.CUDAGraph()
                    with torch.cuda.graph(lane.graph, stream=st):
                        self._run_step(lane)
                lane.graph.replay()
            else:
                self._run_step(lane)
            for j in jobs:
                if j.slice + j.run == j.slices:
                    j.out_host.copy_(j.out_dev, non_blocking=True)
            ev = torch.cuda.Event()
            ev.record(st)
        lane.jobs = jobs
        lane.event = ev
        lane.steps += 1
        self.steps += 1
        self.jobs_per_step.append(len(jobs))

    def _run_step(self, lane: _Lane) -> None:
        rc = self.lib.totri_confocal_batch_run(lane.host_table.data_ptr(),
                                               lane.dev_table.data_ptr(), self.table_bytes,
                                               lane.stream.cuda_stream)
        _lib.check(rc, "totri_confocal_batch_run")

    def _retire(self, lane: _Lane, now: float) -> None:
        for j in lane.jobs:
            if j.slice + j.run < j.slices:  # more slices to run: back to the admitted jobs
                j.slice += j.run
                self.admitted.append(j)
                continue
            for b in j.dev_blocks:
                self.dev.arena.free(b)
            tk = j.ticket
            tk._host_out = j.out_host
            hb = j.host_block
            tk._release = (lambda hb=hb: self.host.arena.free(hb))
            tk.t_done = now
            tk._resident = None
            tk._ready = None
        lane.jobs = []
        lane.event = None

    def poll(self) -> int:
        """Retire finished steps, admit queued requests, start steps on idle
        lanes.  Returns the number of requests still queued or running."""
        now = time.perf_counter()
        for lane in self.lanes:
            if lane.event is not None and lane.event.query():
                self._retire(lane, now)
        self._admit(now)
        for lane in self.lanes:
            if not self.admitted:
                break
            if lane.event is None:
                self._launch(lane, now)
        return self.pending()

    def pending(self) -> int:
        running = sum(len(l.jobs) for l in self.lanes)
        return len(self.queue) + len(self.admitted) + running

    def stalled(self) -> bool:
        """True when queued requests can never start without the caller:
        nothing runs and nothing is admitted, so the device arena is empty,
        yet admission took nothing -- the pinned result arena is held by
        results the caller has not released."""
        return (bool(self.queue) and not self.admitted
                and all(l.event is None for l in self.lanes))

    def _raise_stalled(self) -> None:
        raise RuntimeError(
            f"{len(self.queue)} queued request(s) cannot start: the pinned result arena "
            f"({self.host.arena.used} of {self.host.arena.capacity} bytes in use) is held by "
            "unreleased results; release() finished tickets or raise host_budget")

    def drain(self, sleep: float = 1e-4) -> None:
        """Serve until nothing is queued or running.  Raises RuntimeError when
        the queue is stalled on unreleased results (see stalled())."""
        while self.poll():
            if self.stalled():
                self._raise_stalled()
            time.sleep(sleep)

    def run(self, requests: Sequence, arrivals: Sequence[float], sleep: float = 1e-4,
            on_done=None) -> List[Ticket]:
        """Submit each request at its arrival time (seconds from now), serve
        until all are done, and return the tickets in request order.

        `on_done(ticket)` is called once per finished request, after which
        its result is released; without it, the caller releases results
        (the pinned arena must hold every unreleased result)."""
        order = sorted(range(len(requests)), key=lambda k: arrivals[k])
        tickets: List[Optional[Ticket]] = [None] * len(requests)
        t0 = time.perf_counter()
        nxt = 0
        while nxt < len(order) or self.pending():
            now = time.perf_counter()
            while nxt < len(order) and t0 + arrivals[order[nxt]] <= now:
                k = order[nxt]
                tickets[k] = self.submit(requests[k], t_submit=t0 + arrivals[k])
                nxt += 1
            self.poll()
            if on_done is not None:
                for tk in tickets:
                    if tk is not None and tk.done and tk._release is not None:
                        on_done(tk)
                        tk.release()
            if self.stalled():
                self._raise_stalled()
            time.sleep(sleep)
        return tickets  # type: ignore[return-value]
