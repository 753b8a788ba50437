"""Host logic of the serving loop (no GPU): admission by memory footprint,
slot reuse and return, every request answered once, latency bookkeeping.
The device servers are replaced by a fake with the same interface whose
requests finish after a number of steps fixed by their coefficients."""
import threading
import time

import numpy as np

from edlgt_amd.serving import SolveService, latency_summary, poisson_stream, serve_plain


class FakeServer:
    def __init__(self, slot_bytes, shared_bytes, max_slots, log):
        self.slot_bytes, self.shared_bytes, self.max_slots = slot_bytes, shared_bytes, max_slots
        self.slots = []  # [active, req, steps_left, coef]
        self.log = log
        self.max_active = 0
        self.lock = threading.Lock()

    def info(self):
        a = sum(1 for s in self.slots if s[0])
        return dict(active=a, allocated=len(self.slots), max_slots=self.max_slots,
                    slot_bytes=self.slot_bytes, shared_bytes=self.shared_bytes)

    def submit(self, rid, coef, seed):
        for s in self.slots:
            if not s[0]:
                s[:] = [True, rid, 1 + int(coef[1] * 3), coef]
                break
        else:
            if len(self.slots) >= self.max_slots:
                return False
            self.slots.append([True, rid, 1 + int(coef[1] * 3), coef])
        self.max_active = max(self.max_active, self.info()["active"])
        return True

    def step(self):
        time.sleep(0.001)
        out = []
        for s in self.slots:
            if s[0]:
                s[2] -= 1
                if s[2] == 0:
                    s[0] = False
                    out.append((s[1], np.array([-s[3][1]]), np.zeros(1), 4, True))
        return out

    def trim(self):
        self.slots = [s for s in self.slots if s[0]]


def test_service_admission_and_completion():
    sizes = {"small": 10, "big": 100}
    fakes = {}

    def factory(p, _):
        fakes[p] = FakeServer(sizes[p], 5, 8, None)
        return fakes[p]

    budget = 250  # shared 10 + at most 2 big slots (200) + a few small ones
    svc = SolveService({"small": None, "big": None}, hbm_budget=budget, server_factory=factory)
    reqs = poisson_stream(60, 2000.0, ["small", "big"], lambda rng: [1.0, rng.uniform(0.3, 3.0)], seed=3)
    res = svc.run(reqs)
    assert sorted(res) == list(range(60))
    for r in reqs:
        out = res[r["id"]]
        assert out["problem"] == r["problem"]
        assert out["evals"][0] == -r["coef"][1]
        assert out["done"] >= out["admitted"] >= r["arrival"] - 1e-9
        assert abs(out["latency"] - (out["done"] - r["arrival"])) < 1e-12
    assert svc.peak_bytes <= budget
    assert fakes["big"].max_active <= 2
    # idle slots are returned at the end
    assert all(f.info()["allocated"] == 0 for f in fakes.values())
    s = latency_summary(res, svc.wall)
    assert s["requests"] == 60 and s["p99_ms"] >= s["p50_ms"] > 0


def test_service_rejects_problem_larger_than_budget():
    def factory(p, _):
        return FakeServer(1000, 0, 4, None)

    try:
        SolveService({"huge": None}, hbm_budget=500, server_factory=factory)
    except MemoryError:
        return
    raise AssertionError("expected MemoryError")


def test_plain_baseline_is_serial():
    reqs = poisson_stream(10, 1000.0, ["a"], lambda rng: [1.0, rng.uniform(0.3, 3.0)], seed=5)
    calls = []

    def solve(r):
        calls.append(r["id"])
        time.sleep(0.002)
        return np.array([r["coef"][1]])

    res = serve_plain(solve, reqs)
    assert calls == sorted(calls)
    done = [res[i]["done"] for i in calls]
    starts = [res[i]["admitted"] for i in calls]
    assert all(starts[i + 1] >= done[i] for i in range(len(calls) - 1))


def test_cold_start_limit_until_first_result():
    """With warm starts, a burst runs at most cold_limit requests of a problem
    until one has finished (its cache is empty before that); afterwards the
    slots fill up to max_slots.  Without warm starts there is no limit."""
    for warm, limit, expect_first in ((64, 2, 2), (0, 2, 8)):
        fakes = {}

        def factory(p, _):
            fakes[p] = FakeServer(1, 0, 8, None)
            return fakes[p]

        svc = SolveService({"a": None}, hbm_budget=10_000, server_factory=factory, warm_cache=warm,
                           cold_limit=limit)
        reqs = [dict(id=i, problem="a", coef=[1.0, 1.0 + (i % 3)], seed=i, arrival=0.0) for i in range(20)]
        seen = []
        orig = fakes["a"].submit

        def submit(rid, coef, seed, orig=orig, f=fakes["a"]):
            ok = orig(rid, coef, seed)
            if ok:
                seen.append((rid, f.info()["active"], svc._finished["a"]))
            return ok

        fakes["a"].submit = submit
        res = svc.run(reqs)
        assert sorted(res) == list(range(20))
        before = [a for _, a, fin in seen if fin == 0]
        assert max(before) == expect_first
        assert fakes["a"].max_active == 8
