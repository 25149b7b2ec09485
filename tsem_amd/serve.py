"""Served Poisson solve stream on one GPU.

Requests (a jittered Kuhn mesh of n^3 hexes, order p, seeded problem) arrive
over time and are solved by mixed-precision PCG to a relative residual.
Continuous batching: `slots` workers, each with its own context (HIP
stream), take the next arrived request as soon as they are free and the
request's HBM footprint fits the budget, so requests that arrive while
others are on the GPU start at once and their kernels overlap on the device.
Each worker keeps its spaces cached per (n, p): a new request of a cached
topology only replaces the vertex coordinates (tsem_space_set_geometry) and
rebuilds its load vector and boundary data; cached spaces are dropped (least
recently used first) when the budget needs their memory.

Latency of a request = completion time - arrival time (queueing included);
`Server.run` returns per-request records and p50/p99 summaries.
"""
import heapq
import threading
import time
from dataclasses import dataclass
from math import comb

import numpy as np

from . import api, mesh, problems


@dataclass
class Request:
    id: int
    n: int
    p: int
    seed: int
    arrival: float = 0.0          # seconds after the stream starts
    # filled in by the server
    start: float = None
    done: float = None
    dofs: int = 0
    iterations: int = 0
    relres: float = None
    worker: int = -1
    solution: object = None
    error: str = None


def footprint(n, p):
    """Estimate of the device bytes of one space of n^3 Kuhn cells at order p
    (maps, element-local buffers, factors, solver vectors)."""
    E = 6 * n ** 3
    K = comb(p + 3, 3)
    N = (p * n + 1) ** 3
    terms = 3.2 * N
    return int(terms * 16 + E * K * 24 + E * 200 + N * 150) + (64 << 20)


def make_requests(count, seed=0, n_range=(8, 32), p_range=(3, 8), rate=None):
    """count requests with n, p uniform in the ranges, seed = request id, and
    Poisson-process arrivals of the given rate (all at 0 when rate is None).
    The sizes depend on `seed` only, not on the rate."""
    sizes = np.random.default_rng(seed)
    gaps = np.random.default_rng(seed + 1)
    t = 0.0
    out = []
    for i in range(count):
        if rate:
            t += gaps.exponential(1.0 / rate)
        out.append(Request(i, int(sizes.integers(n_range[0], n_range[1] + 1)),
                           int(sizes.integers(p_range[0], p_range[1] + 1)), i, t))
    return out


class Server:
    def __init__(self, device=0, slots=4, hbm_budget=None, tol=1e-10, keep_solutions=False):
        import torch
        self.device = device
        self.slots = slots
        if hbm_budget is None:
            free, _ = torch.cuda.mem_get_info(device)
            hbm_budget = int(0.7 * free)
        self.budget = hbm_budget
        self.tol = tol
        self.keep = keep_solutions
        self.ctx = [api.Context(device) for _ in range(slots)]
        self.cache = [dict() for _ in range(slots)]   # worker -> {(n, p): (space, last use)}
        self.busy = [None] * slots                    # (n, p) a worker is solving on
        self.in_use = 0                               # estimated bytes of all spaces held
        self.lock = threading.Condition()
        self.pending = []
        self.t0 = None

    # ------------------------------------------------------------- memory --
    def _reserve(self, w, key):
        """With the lock held: True when worker w may take a request of
        topology key, evicting idle cached spaces (least recently used
        first) to fit the budget; an idle device takes any request."""
        need = 0 if key in self.cache[w] else footprint(*key)
        while self.in_use + need > self.budget:
            victims = [(t, ww, k) for ww, c in enumerate(self.cache) for k, (sp, t) in c.items()
                       if not (ww == w and k == key) and self.busy[ww] != k]
            if not victims:
                if any(b is not None for b in self.busy):
                    return False
                break
            _, ww, k = min(victims)
            sp, _ = self.cache[ww].pop(k)
            if sp is not None:
                sp.close()
            self.in_use -= footprint(*k)
        self.in_use += need
        if need:
            self.cache[w][key] = (None, time.perf_counter())  # placeholder until built
        self.busy[w] = key
        return True

    # ------------------------------------------------------------ workers --
    def _next(self, w):
        with self.lock:
            while True:
                if not self.pending:
                    return None
                req = self.pending[0][2]
                now = time.perf_counter() - self.t0
                if req.arrival > now:
                    self.lock.wait(timeout=req.arrival - now)
                    continue
                if self._reserve(w, (req.n, req.p)):
                    heapq.heappop(self.pending)
                    req.start = time.perf_counter() - self.t0
                    return req
                self.lock.wait(timeout=0.01)

    def _space(self, w, n, p):
        key = (n, p)
        with self.lock:
            sp, _ = self.cache[w].get(key, (None, 0))
        if sp is None:
            xyz, tet = mesh.kuhn_cube(n)
            sp = api.Space(self.ctx[w], p, xyz, tet, qorder=p + 3)
        with self.lock:
            self.cache[w][key] = (sp, time.perf_counter())
        return sp

    def solve_one(self, w, req):
        """Setup and solve of one request on worker w's context."""
        sp = self._space(w, req.n, req.p)
        xyz, _ = mesh.kuhn_cube(req.n, jitter=0.1, seed=req.seed)
        sp.set_geometry(xyz)
        b = api.poisson_setup(sp, problems.Manufactured(req.seed))
        x, st = sp.pcg(b, tol=self.tol, criterion=1, precond=1, mixed=True, inner_tol=1e-4, check_every=10,
                       max_iter=100000)
        req.dofs, req.iterations, req.relres = sp.ndof, st["iterations"], st["relres"]
        if self.keep:
            req.solution = x.cpu().numpy()
        return req

    def _worker(self, w):
        import torch
        torch.cuda.set_device(self.device)
        while True:
            req = self._next(w)
            if req is None:
                return
            try:
                self.solve_one(w, req)
            except Exception as e:  # reported by run()
                req.error = repr(e)
            finally:
                req.done = time.perf_counter() - self.t0
                req.worker = w
                with self.lock:
                    self.busy[w] = None
                    self.lock.notify_all()

    def run(self, requests):
        """Serve the requests (arrival times relative to now); returns them
        with timings filled in, plus a summary."""
        self.pending = [(r.arrival, r.id, r) for r in requests]
        heapq.heapify(self.pending)
        self.t0 = time.perf_counter()
        threads = [threading.Thread(target=self._worker, args=(w,), daemon=True) for w in range(self.slots)]
        for t in threads:
            t.start()
        for t in threads:
            t.join()
        wall = time.perf_counter() - self.t0
        failed = [r for r in requests if r.error or r.done is None]
        if failed:
            raise RuntimeError("%d requests failed, first: %s" % (len(failed), failed[0].error))
        lat = np.array([r.done - r.arrival for r in requests])
        dofs = sum(r.dofs for r in requests)
        return requests, dict(requests=len(requests), wall_s=wall, dofs_per_s=dofs / wall,
                              p50_ms=float(np.percentile(lat, 50)) * 1e3, p99_ms=float(np.percentile(lat, 99)) * 1e3,
                              max_relres=max(r.relres for r in requests))

    def close(self):
        for c in self.cache:
            for sp, _ in c.values():
                if sp is not None:
                    sp.close()
            c.clear()
        self.in_use = 0
        for ctx in self.ctx:
            ctx.close()
