"""Serving: the scheduler and dispatcher logic on the host (no GPU)."""
import threading
import time

import numpy as np

from spec_spars_amd.serve import Request, Scheduler, Server, poisson_requests, summarize


def _q(*specs):
    return [Request(i, g, "sparsify", t, k=k) for i, (g, t, k) in enumerate(specs)]


def test_scheduler_fifo_busy_graph_and_slots():
    s = Scheduler(slots=2)
    q = _q((0, 0.0, 8), (0, 0.1, 8), (1, 0.2, 8), (2, 0.3, 8))
    need = lambda r: 0
    i = s.pick(q, 1 << 40, need)
    assert i == 0
    s.started(q.pop(i), 0)
    i = s.pick(q, 1 << 40, need)            # request 1 is on the busy graph 0
    assert q[i].rid == 2
    s.started(q.pop(i), 0)
    assert s.pick(q, 1 << 40, need) is None  # both slots taken
    s.finished(Request(0, 0, "sparsify", 0.0), 0)
    i = s.pick(q, 1 << 40, need)
    assert q[i].rid == 1


def test_scheduler_memory_admission():
    s = Scheduler(slots=4, reserve_bytes=100)
    q = _q((0, 0.0, 8), (1, 0.1, 8))
    need = lambda r: 500 if r.graph == 0 else 200
    assert s.pick(q, 550, need) == 1          # 550 - 100 reserve: only the 200-byte request fits
    s.started(q.pop(1), 200)
    assert s.pick(q, 550, need) is None       # 550 - 100 - 200 left: 250 < 500
    s.finished(Request(1, 1, "sparsify", 0.1), 200)
    assert s.pick(q, 550, need) is None       # still 450 < 500 ...
    assert s.pick(q, 650, need) == 0          # ... until memory frees up


class _FakeRegistry:
    """Executor stand-in: each request sleeps its graph's time; one graph at
    a time is enforced by asserting exclusivity."""

    def __init__(self, times, mem=1 << 40):
        self.times = times
        self.mem = mem
        self.active = set()
        self.lock = threading.Lock()
        self.max_conc = 0

    def free_bytes(self):
        return self.mem

    def need(self, r, precision):
        return 0

    def execute_group(self, group, precision, keep_arrays=True):
        return [self.execute(r, precision, keep_arrays) for r in group]

    def execute(self, r, precision, keep_arrays=True):
        with self.lock:
            assert r.graph not in self.active
            self.active.add(r.graph)
            self.max_conc = max(self.max_conc, len(self.active))
        time.sleep(self.times[r.graph])
        with self.lock:
            self.active.discard(r.graph)
        if r.op == "sparsify":
            return {"edges": 100, "kept": 50}
        return {"rhs": r.k}


def test_server_runs_every_request_with_overlap():
    reg = _FakeRegistry([0.02, 0.03, 0.01, 0.02])
    reqs = poisson_requests(rate=200, duration=0.3, graphs=4, seed=1)
    srv = Server(reg, slots=3, reserve_bytes=0)
    out = srv.run(reqs)
    assert out["completed"] == len(reqs) and out["dropped"] == 0 and out["failed"] == 0
    assert 1 < reg.max_conc <= 3                       # requests overlap, at most `slots` at once
    assert all(r.end >= r.start >= r.arrival for r in reqs)
    assert out["p99_ms"] >= out["p50_ms"] > 0


def test_server_deadline_drops_and_failures_are_contained():
    class Failing(_FakeRegistry):
        def execute(self, r, precision, keep_arrays=True):
            if r.rid == 0:
                raise RuntimeError("boom")
            return super().execute(r, precision, keep_arrays)

    reg = Failing([0.05])
    # requests of 64 columns: four go in one merged pass (256 columns); the
    # failing request fails its whole pass, the server goes on
    reqs = [Request(i, 0, "sparsify", 0.0, k=64) for i in range(40)]
    out = Server(reg, slots=2, reserve_bytes=0).run(reqs, deadline=0.2)
    assert out["failed"] == 4
    assert [r.rid for r in reqs if r.error] == [0, 1, 2, 3]
    assert out["completed"] > 0 and out["dropped"] > 0
    assert out["completed"] + out["failed"] + out["dropped"] == 40


def test_server_merges_sparsifies_on_one_graph():
    class Rec(_FakeRegistry):
        groups = []

        def execute_group(self, group, precision, keep_arrays=True):
            Rec.groups.append([(r.rid, r.op) for r in group])
            return super().execute_group(group, precision, keep_arrays)

    reg = Rec([0.05])
    reqs = [Request(0, 0, "solve", 0.0, k=16)] + [Request(i, 0, "sparsify", 0.001 * i, k=64) for i in range(1, 7)]
    out = Server(reg, slots=1, reserve_bytes=0, max_cols=256).run(reqs)
    assert out["completed"] == 7
    # while the solve runs the sparsifies queue, then go four (256 columns) and two at a time
    assert [len(g) for g in Rec.groups] == [1, 4, 2]
    assert all(op == "sparsify" for g in Rec.groups[1:] for _, op in g)


def test_poisson_requests_shape():
    reqs = poisson_requests(rate=200, duration=5.0, graphs=7, seed=3)
    assert abs(len(reqs) - 1000) < 150
    arr = np.array([r.arrival for r in reqs])
    assert np.all(np.diff(arr) > 0) and arr[-1] < 5.0
    frac = np.mean([r.op == "sparsify" for r in reqs])
    assert 0.4 < frac < 0.6
    assert {r.eps for r in reqs if r.op == "sparsify"} <= {0.2, 0.3, 0.5}


def test_summarize_percentiles():
    rs = [Request(i, 0, "sparsify", 0.0) for i in range(100)]
    for i, r in enumerate(rs):
        r.start, r.end, r.result = 0.0, (i + 1) / 1000.0, {"edges": 10}
    s = summarize(rs, [], 1.0)
    assert abs(s["p50_ms"] - 50.5) < 1 and abs(s["p99_ms"] - 99.01) < 0.5
    assert s["sparsified_edges_per_s"] == 1000


def test_server_groups_solves_on_one_graph():
    class Rec(_FakeRegistry):
        groups = []

        def execute_group(self, group, precision, keep_arrays=True):
            Rec.groups.append([r.rid for r in group])
            return super().execute_group(group, precision, keep_arrays)

    reg = Rec([0.05, 0.01])
    reqs = [Request(0, 0, "sparsify", 0.0, k=32)] + [Request(i, 0, "solve", 0.001 * i, k=16) for i in range(1, 6)] + \
        [Request(6, 1, "solve", 0.002, k=16)]
    out = Server(reg, slots=2, reserve_bytes=0, max_rhs=48).run(reqs)
    assert out["completed"] == 7
    # while graph 0 runs the sparsify, its solves queue up and then go three at a time (48 RHS)
    g0 = [g for g in Rec.groups if g[0] in (1, 2, 3, 4, 5)]
    assert [len(g) for g in g0] == [3, 2]
    assert sorted(x for g in Rec.groups for x in g) == list(range(7))


def test_rhs_block_is_each_requests_columns_side_by_side():
    """A merged solve's right-hand sides (one gather over the graph's
    Gaussian block) are each request's seeded columns with its seeded signs,
    in request order: the block a request gets alone, bit for bit.  Requests
    of more columns than the block holds are refused."""
    import pytest
    from spec_spars_amd.serve import Registry

    class HostRegistry(Registry):
        def __init__(self, n, width=64):
            self.rhs_width = width
            self.rhs_base = {0: np.random.default_rng(3).standard_normal((n, width))}

    reg = HostRegistry(1000)
    reqs = [(16, 5), (8, 6), (16, 7)]
    b = reg.rhs_block(0, reqs)
    c = 0
    for k, seed in reqs:
        rng = np.random.default_rng(seed)
        cols, signs = rng.permutation(64)[:k], rng.choice([-1.0, 1.0], size=k)
        assert np.array_equal(b[:, c:c + k], reg.rhs_base[0][:, cols] * signs)
        assert np.array_equal(b[:, c:c + k], reg.rhs(0, k, seed))
        c += k
    assert b.shape == (1000, 40) and b.flags.c_contiguous
    with pytest.raises(ValueError):
        reg.rhs(0, 65, 1)
