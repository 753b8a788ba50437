"""Serving loop for lowest-eigenvalue requests with continuous batching.

A request asks for the lowest n_eigs eigenvalues of H(c) = sum_t c[t] H_t on
one of several resident problems (a sector basis plus its multi-term CSR
matrix, assembled once).  The reference answers each such request with a
fresh build_Hamiltonian + diagonalize_Hamiltonian
(edlgt/models/quantum_model.py:606-611, edlgt/modeling/qmb_hamiltonian.py:
173-332); here every problem has one continuous-batching server
(edlgt_server_*) driven by its own host thread on its own HIP stream, so the
problems' device work overlaps, and inside a server all active requests share
one fused SpMM per step.

Admission is by device-memory footprint: a request enters its server when a
free allocated slot exists, or when the bytes of one more slot fit in the
budget shared by all problems; a server that goes idle returns its slots.
With warm starts, a problem whose cache is still empty runs at most
`cold_limit` requests: the others wait until one has finished and then start
from its Ritz vectors instead of cold (a cold Z3 4x4 solve takes about three
times the block steps of a warm one, and per-slot orthogonalisation, not the
shared SpMM, sets the cost of a step there; on the 64-request Z3 sample,
all present at once, cold_limit 1 / 2 / none served 14.8 / 10.0 / 10.3
solves/s).
Latency is arrival -> result on the host clock.
"""
import threading
import time

import numpy as np

from .device import Server


def poisson_stream(n_requests, rate, problems, coef_fn, seed):
    """Synthetic request stream: exponential inter-arrival times at `rate`
    per second, problem drawn uniformly from `problems`, coefficients from
    coef_fn(rng) -> sequence.  Returns a list of dicts sorted by arrival."""
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate, n_requests)
    t = np.cumsum(gaps) - gaps[0]
    out = []
    for i in range(n_requests):
        p = problems[int(rng.integers(len(problems)))]
        out.append(dict(id=i, problem=p, coef=list(coef_fn(rng)), seed=10_000 + i, arrival=float(t[i])))
    return out


class SolveService:
    """Continuous-batching solves over several resident multi-term problems."""

    def __init__(self, problems, n_eigs=4, tol=1e-8, max_basis=32, max_slots=16, warm_cache=0,
                 hbm_budget=None, max_matvecs=200000, server_factory=None, cold_limit=1):
        """problems: {name: CSRMatrix (multi-term)}.  hbm_budget: bytes for
        solver slots, warm caches and workspaces over all problems (default:
        60% of the free device memory).  server_factory(name, problem) may
        replace the device Server (same submit/step/info/trim interface)."""
        self.problems = dict(problems)
        self.n_eigs = n_eigs
        if server_factory is None:
            import torch
            self._stream_ctx = torch.cuda.stream
            if hbm_budget is None:
                free, _ = torch.cuda.mem_get_info()
                hbm_budget = int(0.6 * free)
            self.streams = {p: torch.cuda.Stream() for p in self.problems}

            def server_factory(p, csr):
                return Server(csr, n_eigs=n_eigs, tol=tol, max_basis=max_basis, max_matvecs=max_matvecs,
                              max_slots=max_slots, warm_cache=warm_cache, stream=self.streams[p])
        else:
            import contextlib
            self._stream_ctx = lambda s: contextlib.nullcontext()
            self.streams = {p: None for p in self.problems}
        if hbm_budget is None:
            raise ValueError("hbm_budget is required with a custom server_factory")
        self.budget = int(hbm_budget)
        self.servers = {p: server_factory(p, csr) for p, csr in self.problems.items()}
        self._info = {p: s.info() for p, s in self.servers.items()}
        for p, inf in self._info.items():
            if inf["slot_bytes"] + inf["shared_bytes"] > self.budget:
                raise MemoryError(f"problem {p}: one slot ({inf['slot_bytes']} B) does not fit the budget")
        self._lock = threading.Lock()
        self._starved = False
        self.peak_bytes = 0
        # requests finished per problem (its warm cache is empty before the first)
        self.cold_limit = int(cold_limit) if warm_cache > 0 else 0
        self._finished = {p: 0 for p in self.problems}

    def _bytes(self):
        # shared bytes (warm cache, SpMM workspace) are reserved from the start
        return sum(i["allocated"] * i["slot_bytes"] + i["shared_bytes"] for i in self._info.values())

    def _try_admit(self, p, req):
        """Admission under the lock: a free allocated slot, or one more slot
        within the budget."""
        srv = self.servers[p]
        with self._lock:
            inf = self._info[p]
            if self.cold_limit and self._finished[p] == 0 and inf["active"] >= self.cold_limit:
                return False
            if inf["allocated"] <= inf["active"]:
                if inf["allocated"] >= inf["max_slots"]:
                    return False
                if self._bytes() + inf["slot_bytes"] > self.budget:
                    self._starved = True
                    return False
            ok = srv.submit(req["id"], req["coef"], req["seed"])
            self._info[p] = srv.info()
            self.peak_bytes = max(self.peak_bytes, self._bytes())
            return ok

    def run(self, requests):
        """Serves `requests` (dicts with id, problem, coef, seed, arrival in
        seconds from the start) and returns {id: result dict} with evals,
        resid_est, matvecs, converged, arrival, admitted, done, latency."""
        t0 = time.perf_counter()
        queues = {p: [r for r in requests if r["problem"] == p] for p in self.problems}
        results = {}
        errors = []

        def worker(p):
            srv, q, pos = self.servers[p], queues[p], 0
            admitted = {}
            try:
                with self._stream_ctx(self.streams[p]):
                    while pos < len(q) or self._info[p]["active"]:
                        now = time.perf_counter() - t0
                        while pos < len(q) and q[pos]["arrival"] <= now and self._try_admit(p, q[pos]):
                            admitted[q[pos]["id"]] = time.perf_counter() - t0
                            pos += 1
                        if self._info[p]["active"]:
                            done = srv.step()
                            tdone = time.perf_counter() - t0
                            with self._lock:
                                self._info[p] = srv.info()
                                self._finished[p] += len(done)
                            for rid, ev, es, mv, cv in done:
                                req = next(r for r in q if r["id"] == rid)
                                results[rid] = dict(evals=ev, resid_est=es, matvecs=mv, converged=cv,
                                                    problem=p, coef=req["coef"], arrival=req["arrival"],
                                                    admitted=admitted.pop(rid), done=tdone,
                                                    latency=tdone - req["arrival"])
                        elif pos < len(q):
                            if self._starved and self._info[p]["allocated"]:
                                with self._lock:  # give idle slots back to a starved problem
                                    srv.trim()
                                    self._info[p] = srv.info()
                                    self._starved = False
                            wait = q[pos]["arrival"] - (time.perf_counter() - t0)
                            time.sleep(min(max(wait, 0.0), 0.002) if wait > 0 else 0.0005)
                    if self._info[p]["allocated"]:
                        with self._lock:  # done: return the slots to the shared budget
                            srv.trim()
                            self._info[p] = srv.info()
            except Exception as e:  # surfaced after join
                errors.append(e)

        threads = [threading.Thread(target=worker, args=(p,)) for p in self.problems]
        for th in threads:
            th.start()
        for th in threads:
            th.join()
        if errors:
            raise errors[0]
        self.wall = time.perf_counter() - t0
        return results


def latency_summary(results, wall):
    lat = np.array([r["latency"] for r in results.values()])
    return dict(requests=len(results), wall_s=wall, solves_per_s=len(results) / wall,
                p50_ms=float(np.percentile(lat, 50) * 1e3), p99_ms=float(np.percentile(lat, 99) * 1e3),
                mean_ms=float(lat.mean() * 1e3))


def serve_plain(solve_one, requests):
    """The one-at-a-time baseline: requests in arrival order, each solved by
    solve_one(req) -> evals after the previous one finished (waiting for its
    arrival when idle).  Returns {id: result dict} like SolveService.run."""
    t0 = time.perf_counter()
    results = {}
    for req in sorted(requests, key=lambda r: r["arrival"]):
        now = time.perf_counter() - t0
        if req["arrival"] > now:
            time.sleep(req["arrival"] - now)
        start = time.perf_counter() - t0
        ev = solve_one(req)
        tdone = time.perf_counter() - t0
        results[req["id"]] = dict(evals=ev, problem=req["problem"], coef=req["coef"], arrival=req["arrival"],
                                  admitted=start, done=tdone, latency=tdone - req["arrival"])
    serve_plain.wall = time.perf_counter() - t0
    return results
