"""Continuous-batching solve service on one GPU.

Requests ask for (D^dag D)^-1 on k right-hand sides on one of several
device-resident gauge configurations, with their own mass.  A lane is one
``wl_batch`` (a fixed set of solver columns on one configuration, its own
HIP stream); every right-hand side occupies one column from admission to
retirement, so requests that arrive while others are on the GPU join the
running lanes as columns free up, and lanes on different configurations
overlap on the device.  Admission is bounded by an HBM budget covering lane
workspaces (resident gauge fields are counted too); idle lanes are released
when a waiting request needs the memory.

This has no counterpart in the reference: it generalizes the 12 sequential
``solve_normal`` calls of the point-source propagator loop
(core/measurements.py:277-371) into a shared request stream.
"""

from __future__ import annotations

import collections
import time
from ctypes import byref, c_int, c_void_p
from dataclasses import dataclass, field
from typing import Any, Callable, Dict, List, Optional, Sequence, Tuple

import numpy as np
import torch

from . import _native as nat


@dataclass
class Request:
    rid: int
    key: Tuple[Tuple[int, ...], int]  # (lattice shape, gauge field id)
    mass: float
    rhs: Optional[torch.Tensor]        # (k, *L, 4, Nc) complex, on the device (None: made at admission)
    arrival: float = 0.0               # seconds since the start of the run
    make_rhs: Optional[Callable[[], torch.Tensor]] = None  # builds rhs at admission when rhs is None
    k: int = 0                         # right-hand sides of a request whose rhs is made at admission
    # filled by the server
    x: Optional[torch.Tensor] = None
    admitted: Optional[float] = None
    finished: Optional[float] = None
    stats: List[Dict[str, Any]] = field(default_factory=list)
    _left: int = 0

    @property
    def nrhs(self) -> int:
        return int(self.rhs.shape[0]) if self.rhs is not None else int(self.k)

    @property
    def latency(self) -> float:
        return float("nan") if self.finished is None else self.finished - self.arrival


class Lane:
    """A wl_batch of `capacity` columns on one resident configuration."""

    def __init__(self, ctx: nat.Context, gauge: nat.Gauge, capacity: int, sp: nat.WlSolveParams):
        self.ctx = ctx
        self.gauge = gauge
        self.capacity = int(capacity)
        self.stream = torch.cuda.Stream()
        h = c_void_p()
        nat.check(ctx._lib.wl_batch_create(ctx.h, gauge.h, self.capacity, byref(sp), c_void_p(self.stream.cuda_stream),
                                           byref(h)))
        self.h = h
        self.owner: List[Optional[Tuple[Request, int]]] = [None] * self.capacity
        self._stats = (nat.WlSolveStats * self.capacity)()
        self._state = (c_int * self.capacity)()
        self.stepped = False

    @property
    def free(self) -> int:
        return sum(1 for o in self.owner if o is None)

    @property
    def running(self) -> int:
        return self.capacity - self.free

    def admit(self, req: Request) -> None:
        lib = self.ctx._lib
        if req.rhs is None:
            req.rhs = req.make_rhs()
        dt = nat.WL_C128 if req.rhs.dtype == torch.complex128 else nat.WL_C64
        slots = [i for i, o in enumerate(self.owner) if o is None][: req.nrhs]
        assert len(slots) == req.nrhs
        caller = torch.cuda.current_stream()  # the stream that produced req.rhs
        with torch.cuda.stream(self.stream):
            self.stream.wait_stream(caller)
            for j, slot in enumerate(slots):
                nat.check(lib.wl_batch_admit(self.h, slot, c_void_p(req.rhs[j].data_ptr()), dt, float(req.mass)))
                self.owner[slot] = (req, j)
        req._left = req.nrhs
        if req.x is None:
            req.x = torch.empty_like(req.rhs)
            req.stats = [dict() for _ in range(req.nrhs)]

    def step(self) -> None:
        if self.running:
            nat.check(self.ctx._lib.wl_batch_step(self.h))
            self.stepped = True

    def ready(self) -> bool:
        """Whether the work queued on this lane has finished (no waiting)."""
        return self.stream.query()

    def harvest(self, now_fn) -> List[Request]:
        """Wait for this lane's queued work, retire finished columns and return
        the requests whose every column has finished."""
        if not self.stepped:
            return []
        self.stepped = False
        lib = self.ctx._lib
        nat.check(lib.wl_batch_query(self.h, self._stats, self._state))
        done: List[Request] = []
        for slot in range(self.capacity):
            if self._state[slot] != 2:
                continue
            req, j = self.owner[slot]
            dt = nat.WL_C128 if req.x.dtype == torch.complex128 else nat.WL_C64
            nat.check(lib.wl_batch_retire(self.h, slot, c_void_p(req.x[j].data_ptr()), dt))
            req.stats[j] = nat.stats_list(self._stats, self.capacity)[slot]
            self.owner[slot] = None
            req._left -= 1
            if req._left == 0:
                done.append(req)
        if done:
            self.stream.synchronize()  # solutions written
            t = now_fn()
            for r in done:
                r.finished = t
        return done

    def close(self) -> None:
        if getattr(self, "h", None) is not None and self.h.value:
            self.ctx._lib.wl_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Scheduler: FIFO admission per lane with an HBM budget, continuous
    batching, lanes on separate streams.  Each lane always has one block of
    iterations queued while it has work; a lane is harvested (finished
    requests retired, freed columns reused) as soon as its own block has
    finished, so a lane on a large lattice does not hold back the lanes on
    small ones."""

    IDLE_POLL_S = 2e-4  # host sleep while every lane's block is still running

    def __init__(self, gauge_fields: Dict[Tuple[Tuple[int, ...], int], Any], nc: int = 3, bc=(1, 1, 1, -1),
                 hbm_budget_bytes: float = 64e9, capacity: int = 16, inner: str = "f16", recon: int = 12,
                 tol: float = 1e-10, maxiter: int = 20000, delta: float = 0.01, poll_every: int = 8,
                 use_graph: bool = True, backend: Optional["LaneBackend"] = None):
        self.nc = nc
        self.capacity = int(capacity)
        self.budget = float(hbm_budget_bytes)
        self.backend = backend or NativeBackend(nc, bc, capacity, inner, recon, tol, maxiter, delta, poll_every,
                                                use_graph)
        self.resident_bytes = 0
        for key, U in gauge_fields.items():
            self.resident_bytes += self.backend.add_field((tuple(key[0]), key[1]), U)
        self.lanes: Dict[Tuple[Tuple[int, ...], int], Any] = {}
        self.lane_bytes: Dict[Tuple[Tuple[int, ...], int], int] = {}
        self.admitted_bytes = 0

    def lane_cost(self, L) -> int:
        return self.backend.lane_cost(tuple(L))

    def _get_lane(self, key) -> Optional[Lane]:
        if key in self.lanes:
            return self.lanes[key]
        cost = self.lane_cost(key[0])
        if self.resident_bytes + self.admitted_bytes + cost > self.budget:
            # release idle lanes until the new one fits
            for k2 in list(self.lanes):
                if self.lanes[k2].running == 0:
                    self.lanes[k2].close()
                    self.admitted_bytes -= self.lane_bytes.pop(k2)
                    del self.lanes[k2]
                    if self.resident_bytes + self.admitted_bytes + cost <= self.budget:
                        break
            if self.resident_bytes + self.admitted_bytes + cost > self.budget:
                return None
        lane = self.backend.new_lane(key)
        self.lanes[key] = lane
        self.lane_bytes[key] = cost
        self.admitted_bytes += cost
        return lane

    def run(self, requests: Sequence[Request], realtime: bool = True, release: bool = False) -> Dict[str, Any]:
        """Serve `requests` (sorted by arrival).  With realtime=True arrivals
        follow the wall clock; otherwise all are available at t=0.  With
        release=True a finished request drops its right-hand sides and
        solution (its stats are kept), so that long request streams do not
        hold every field on the device."""
        for r in requests:
            if r.nrhs > self.capacity:
                raise ValueError(f"request {r.rid} has {r.nrhs} right-hand sides > lane capacity {self.capacity}")
        pending = collections.deque()
        future = collections.deque(sorted(requests, key=lambda r: r.arrival))
        t0 = time.perf_counter()
        now = lambda: time.perf_counter() - t0
        completed: List[Request] = []
        while future or pending or any(l.running for l in self.lanes.values()):
            t = now()
            while future and (not realtime or future[0].arrival <= t):
                pending.append(future.popleft())
            # admission: FIFO, but a request blocked on its lane does not block other lanes
            blocked = set()
            for _ in range(len(pending)):
                r = pending.popleft()
                if r.key in blocked:
                    pending.append(r)
                    continue
                lane = self._get_lane(r.key)
                if lane is None or lane.free < r.nrhs:
                    blocked.add(r.key)
                    pending.append(r)
                    continue
                lane.admit(r)
                r.admitted = now()
            active = [l for l in self.lanes.values() if l.running]
            if not active:
                if pending and not future:
                    raise RuntimeError(f"request {pending[0].rid} ({pending[0].nrhs} right-hand sides on "
                                       f"{pending[0].key[0]}) does not fit in the HBM budget")
                if future and realtime:
                    time.sleep(max(0.0, min(0.002, future[0].arrival - now())))
                elif pending and future and not realtime:
                    raise RuntimeError("admission stalled")
                continue
            for lane in active:
                if not lane.stepped:
                    lane.step()
            ready = [lane for lane in active if lane.stepped and lane.ready()]
            if not ready:
                time.sleep(self.IDLE_POLL_S)
                continue
            for lane in ready:
                fin = lane.harvest(now)
                if release:
                    for r in fin:
                        r.rhs = None
                        r.x = None
                completed.extend(fin)
        makespan = now()
        for l in list(self.lanes.values()):
            l.close()
        self.lanes.clear()
        self.admitted_bytes = 0
        return summarize(completed, makespan)


class NativeBackend:
    """Lanes on libwilson_amd: resident gauge fields per (lattice, field id)."""

    def __init__(self, nc, bc, capacity, inner, recon, tol, maxiter, delta, poll_every, use_graph):
        self.nc = nc
        self.bc = tuple(bc)
        self.capacity = int(capacity)
        self.inner = nat.PREC[inner]
        self.sp, _ = nat.solve_params(tol, maxiter, nat.WL_STOP_TRUE_RESIDUAL, self.inner, recon, delta, poll_every,
                                      use_graph)
        self.ctxs: Dict[Tuple[int, ...], nat.Context] = {}
        self.gauges: Dict[Tuple[Tuple[int, ...], int], nat.Gauge] = {}

    def ctx(self, L) -> nat.Context:
        L = tuple(L)
        if L not in self.ctxs:
            self.ctxs[L] = nat.Context(L, self.nc, 1, 0.0, 1.0, self.bc)
        return self.ctxs[L]

    def add_field(self, key, U) -> int:
        ctx = self.ctx(key[0])
        Ut = U if isinstance(U, torch.Tensor) else torch.from_numpy(np.ascontiguousarray(U))
        self.gauges[key] = nat.Gauge(ctx, Ut.cuda(), nat.WL_LAYOUT_BMXYIJ)
        V = int(np.prod(key[0]))
        return V * 4 * (self.nc * self.nc) * 16 * 2  # FP64 master + lower-precision encodings (bound)

    def lane_cost(self, L) -> int:
        ctx = self.ctx(L)
        return int(ctx._lib.wl_batch_bytes(ctx.h, self.capacity, self.inner))

    def new_lane(self, key) -> Lane:
        return Lane(self.ctxs[key[0]], self.gauges[key], self.capacity, self.sp)


def summarize(done: Sequence[Request], makespan: float) -> Dict[str, Any]:
    lat = np.array([r.latency for r in done])
    nrhs = int(sum(r.nrhs for r in done))
    return {
        "requests": len(done),
        "rhs": nrhs,
        "makespan_s": round(makespan, 4),
        "rhs_per_s": round(nrhs / makespan, 3) if makespan > 0 else float("nan"),
        "requests_per_s": round(len(done) / makespan, 3) if makespan > 0 else float("nan"),
        "latency_p50_ms": round(float(np.percentile(lat, 50)) * 1e3, 2) if len(lat) else None,
        "latency_p99_ms": round(float(np.percentile(lat, 99)) * 1e3, 2) if len(lat) else None,
        "true_resid_max": max((s.get("true_resid", 0.0) for r in done for s in r.stats), default=None),
        "all_converged": all(s.get("status") == "converged" for r in done for s in r.stats),
    }


def fp64_queue(requests: Sequence[Request], gauges: Dict, ctxs: Dict, tol: float = 1e-10, maxiter: int = 20000,
               realtime: bool = True) -> Dict[str, Any]:
    """Unbatched FP64 baseline: one FIFO queue, each right-hand side solved
    alone with plain FP64 CG on D^dag D."""
    stream = torch.cuda.Stream()
    sp, _ = nat.solve_params(tol, maxiter, nat.WL_STOP_TRUE_RESIDUAL, nat.WL_F64, 18, 0.0, 8, True)
    fields_cache: Dict[Tuple[int, ...], Tuple[nat.Field, nat.Field]] = {}
    t0 = time.perf_counter()
    now = lambda: time.perf_counter() - t0
    done = []
    for r in sorted(requests, key=lambda q: q.arrival):
        if realtime and now() < r.arrival:
            time.sleep(r.arrival - now())
        L = r.key[0]
        ctx = ctxs[L]
        if L not in fields_cache:
            fields_cache[L] = (nat.Field(ctx, nat.WL_F64, 1), nat.Field(ctx, nat.WL_F64, 1))
        b, x = fields_cache[L]
        ctx.set_mass(r.mass)
        r.x = torch.empty_like(r.rhs)
        r.stats = []
        caller = torch.cuda.current_stream()
        with torch.cuda.stream(stream):
            stream.wait_stream(caller)
            for j in range(r.nrhs):
                b.import_(r.rhs[j:j + 1])
                st = (nat.WlSolveStats * 1)()
                nat.check(ctx._lib.wl_cg_normal_field(ctx.h, gauges[r.key].h, b.h, x.h, 0, 1, byref(sp), st,
                                                      c_void_p(stream.cuda_stream)))
                x.export(r.x[j:j + 1])
                r.stats.append(nat.stats_list(st, 1)[0])
        stream.synchronize()
        r.finished = now()
        done.append(r)
    return summarize(done, now())


def _gaussian_rhs(shape, seed: int, dtype) -> torch.Tensor:
    gen = torch.Generator(device="cuda")
    gen.manual_seed(seed)
    re = torch.randn(shape, generator=gen, device="cuda", dtype=torch.float64)
    im = torch.randn(shape, generator=gen, device="cuda", dtype=torch.float64)
    return torch.complex(re, im).to(dtype) / np.sqrt(2.0)


def poisson_requests(rate_hz: float, n: int, lattices: Sequence[Tuple[int, ...]], fields_per_lattice: int,
                     mass_range=(0.02, 0.2), rhs_range=(1, 12), seed: int = 0, nc: int = 3,
                     dtype=torch.complex128, lazy: bool = False) -> List[Request]:
    """n requests with exponential inter-arrival times (mean 1/rate_hz), the
    lattice, field, mass and RHS count drawn uniformly; complex Gaussian
    right-hand sides generated on the device (seeded per request), now or,
    with lazy=True, when the request is admitted."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for i in range(n):
        t += float(rng.exponential(1.0 / rate_hz))
        L = tuple(lattices[int(rng.integers(len(lattices)))])
        fid = int(rng.integers(fields_per_lattice))
        mass = float(rng.uniform(*mass_range))
        k = int(rng.integers(rhs_range[0], rhs_range[1] + 1))
        shape = (k, *L, 4, nc)
        rseed = seed * 1000003 + i
        if lazy:
            out.append(Request(rid=i, key=(L, fid), mass=mass, rhs=None, arrival=t, k=k,
                               make_rhs=lambda shape=shape, rseed=rseed: _gaussian_rhs(shape, rseed, dtype)))
        else:
            out.append(Request(rid=i, key=(L, fid), mass=mass, rhs=_gaussian_rhs(shape, rseed, dtype), arrival=t))
    return out
