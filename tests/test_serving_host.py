"""Scheduling logic of the continuous-batching server on a CPU fake backend:
FIFO admission per lane, HBM budget with idle-lane release, no head-of-line
blocking across lanes, continuous admission into running lanes, and a clear
error for a request that can never fit."""

import numpy as np
import pytest

from jaxqft_amd.serving import Request, Server


class FakeLane:
    """Columns finish after a number of steps that depends on the mass; the
    'solution' of column j is rhs[j] / (m + 4)^2.  A step's block stays in
    flight for `polls` readiness checks (a lane on a large lattice)."""

    def __init__(self, key, capacity, log, polls=0):
        self.key = key
        self.capacity = capacity
        self.owner = [None] * capacity
        self.left = [0] * capacity
        self.stepped = False
        self.log = log
        self.closed = False
        self.polls = polls
        self.wait = 0

    @property
    def free(self):
        return sum(1 for o in self.owner if o is None)

    @property
    def running(self):
        return self.capacity - self.free

    def admit(self, req):
        slots = [i for i, o in enumerate(self.owner) if o is None][: req.nrhs]
        for j, s in enumerate(slots):
            self.owner[s] = (req, j)
            self.left[s] = 1 + int(round(req.mass * 20))
        req._left = req.nrhs
        if req.x is None:
            req.x = np.zeros_like(req.rhs)
            req.stats = [dict() for _ in range(req.nrhs)]
        self.log.append(("admit", req.rid, self.key))

    def step(self):
        if self.running:
            for s in range(self.capacity):
                if self.owner[s] is not None and self.left[s] > 0:
                    self.left[s] -= 1
            self.stepped = True
            self.wait = self.polls
            self.log.append(("step", None, self.key))

    def ready(self):
        if self.wait > 0:
            self.wait -= 1
            return False
        return True

    def harvest(self, now_fn):
        if not self.stepped:
            return []
        self.stepped = False
        done = []
        for s in range(self.capacity):
            if self.owner[s] is None or self.left[s] > 0:
                continue
            req, j = self.owner[s]
            req.x[j] = req.rhs[j] / (req.mass + 4.0) ** 2
            req.stats[j] = {"status": "converged", "true_resid": 0.0, "iters": 1}
            self.owner[s] = None
            req._left -= 1
            if req._left == 0:
                done.append(req)
                self.log.append(("done", req.rid, self.key))
        t = now_fn()
        for r in done:
            r.finished = t
        return done

    def close(self):
        self.closed = True


class FakeBackend:
    def __init__(self, capacity, lane_bytes=100, polls=None):
        self.capacity = capacity
        self.lane_bytes = lane_bytes
        self.log = []
        self.live = 0
        self.max_live = 0
        self.lanes = []
        self.polls = polls or {}

    def add_field(self, key, U):
        return 10

    def lane_cost(self, L):
        return self.lane_bytes

    def new_lane(self, key):
        lane = FakeLane(key, self.capacity, self.log, self.polls.get(key, 0))
        self.lanes.append(lane)
        live = sum(1 for l in self.lanes if not l.closed)
        self.max_live = max(self.max_live, live)
        return lane


def make_requests(n, keys, rng, kmax=4):
    out = []
    for i in range(n):
        key = keys[int(rng.integers(len(keys)))]
        k = int(rng.integers(1, kmax + 1))
        out.append(Request(rid=i, key=key, mass=float(rng.uniform(0.02, 0.2)),
                           rhs=rng.standard_normal((k, 2, 2)) + 0j, arrival=0.0))
    return out


def fields_for(keys):
    return {k: None for k in keys}


def test_all_requests_complete_with_correct_results():
    rng = np.random.default_rng(0)
    keys = [((4, 4, 4, 4), 0), ((4, 4, 4, 4), 1), ((8, 8, 8, 8), 0)]
    be = FakeBackend(capacity=6)
    srv = Server(fields_for(keys), capacity=6, hbm_budget_bytes=1e9, backend=be)
    reqs = make_requests(25, keys, rng)
    res = srv.run(reqs, realtime=False)
    assert res["requests"] == 25 and res["rhs"] == sum(r.nrhs for r in reqs)
    for r in reqs:
        np.testing.assert_allclose(r.x, r.rhs / (r.mass + 4) ** 2)
        assert r.finished >= r.admitted


def test_budget_limits_concurrent_lanes_and_releases_idle_ones():
    rng = np.random.default_rng(1)
    keys = [((4, 4, 4, 4), f) for f in range(4)]
    be = FakeBackend(capacity=4, lane_bytes=100)
    # resident fields 4 x 10 bytes; room for exactly two lanes
    srv = Server(fields_for(keys), capacity=4, hbm_budget_bytes=40 + 200, backend=be)
    res = srv.run(make_requests(30, keys, rng), realtime=False)
    assert res["requests"] == 30
    assert be.max_live <= 2
    assert any(l.closed for l in be.lanes)  # idle lanes were released for others


def test_continuous_admission_into_running_lane():
    """A later request joins a lane whose earlier columns are still running."""
    key = ((4, 4, 4, 4), 0)
    be = FakeBackend(capacity=8)
    srv = Server(fields_for([key]), capacity=8, hbm_budget_bytes=1e9, backend=be)
    slow = Request(rid=0, key=key, mass=0.2, rhs=np.ones((3, 2, 2), complex))   # 5 steps
    fast = [Request(rid=i, key=key, mass=0.02, rhs=np.ones((2, 2, 2), complex)) for i in range(1, 5)]  # 1 step
    srv.run([slow] + fast, realtime=False)
    order = [(e[0], e[1]) for e in be.log]
    # request 4 cannot be admitted at t=0 (3 + 2 + 2 = 7 of 8 columns busy) and
    # must enter while the slow request is still running
    i_admit4 = order.index(("admit", 4))
    i_done0 = order.index(("done", 0))
    assert i_admit4 < i_done0
    assert order.index(("done", 1)) < i_admit4


def test_no_head_of_line_blocking_across_lanes():
    a, b = ((4, 4, 4, 4), 0), ((4, 4, 4, 4), 1)
    be = FakeBackend(capacity=4)
    srv = Server(fields_for([a, b]), capacity=4, hbm_budget_bytes=1e9, backend=be)
    reqs = [Request(rid=0, key=a, mass=0.2, rhs=np.ones((4, 2, 2), complex)),
            Request(rid=1, key=a, mass=0.2, rhs=np.ones((4, 2, 2), complex)),   # blocked behind 0
            Request(rid=2, key=b, mass=0.02, rhs=np.ones((1, 2, 2), complex))]  # must not wait for 1
    srv.run(reqs, realtime=False)
    order = [(e[0], e[1]) for e in be.log]
    assert order.index(("admit", 2)) < order.index(("admit", 1))


def test_request_larger_than_budget_raises():
    key = ((4, 4, 4, 4), 0)
    be = FakeBackend(capacity=4, lane_bytes=1000)
    srv = Server(fields_for([key]), capacity=4, hbm_budget_bytes=500, backend=be)
    with pytest.raises(RuntimeError, match="HBM budget"):
        srv.run([Request(rid=0, key=key, mass=0.1, rhs=np.ones((1, 2, 2), complex))], realtime=False)
    with pytest.raises(ValueError, match="capacity"):
        Server(fields_for([key]), capacity=4, hbm_budget_bytes=1e9, backend=FakeBackend(4)).run(
            [Request(rid=0, key=key, mass=0.1, rhs=np.ones((5, 2, 2), complex))], realtime=False)


def test_slow_lane_does_not_hold_back_fast_lanes():
    """A lane whose block of iterations is still in flight (a large lattice)
    is not waited for: the other lane is harvested and stepped again meanwhile,
    so its requests finish before the slow lane's first block does."""
    big, small = ((32, 32, 32, 32), 0), ((8, 8, 8, 8), 0)
    be = FakeBackend(capacity=4, polls={big: 50})
    srv = Server(fields_for([big, small]), capacity=4, hbm_budget_bytes=1e9, backend=be)
    srv.IDLE_POLL_S = 0.0
    reqs = [Request(rid=0, key=big, mass=0.02, rhs=np.ones((2, 2, 2), complex))]
    reqs += [Request(rid=i, key=small, mass=0.1, rhs=np.ones((2, 2, 2), complex)) for i in range(1, 6)]
    res = srv.run(reqs, realtime=False)
    assert res["requests"] == 6
    order = [(e[0], e[1], e[2]) for e in be.log]
    first_big_done = order.index(("done", 0, big))
    small_done = [i for i, e in enumerate(order) if e[0] == "done" and e[2] == small]
    assert len(small_done) == 5 and max(small_done) < first_big_done
    # the small lane was stepped many times while the big lane's only block was in flight
    steps_small = sum(1 for e in order[:first_big_done] if e[0] == "step" and e[2] == small)
    steps_big = sum(1 for e in order[:first_big_done] if e[0] == "step" and e[2] == big)
    assert steps_big == 1 and steps_small >= 5
