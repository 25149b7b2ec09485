"""Served Poisson solve stream on one GPU.

Requests (a jittered Kuhn mesh of n^3 hexes, order p, seeded problem) arrive
over time and are solved by mixed-precision PCG to a relative residual.
Continuous batching: `slots` workers, each with its own context (HIP
stream), take the next arrived request as soon as they are free and the
request's HBM footprint fits the budget, so requests that arrive while
others are on the GPU start at once and their kernels overlap on the device.
Spaces are cached per (n, p) in a pool shared by the workers, each with its
own context (HIP stream): a new request of a cached topology only replaces
the vertex coordinates (tsem_space_set_geometry) and rebuilds its load
vector and boundary data; idle cached spaces are dropped (least recently
used first) when the budget needs their memory.  Workers are split into a
small-request lane and a general lane, so small requests never queue behind
large ones; and (classes=True) the contexts of small topologies are
foreground, those of large ones background, so a large solve pauses between
its chunks of iterations while a small one runs (tsem_ctx_set_class).

Latency of a request = completion time - arrival time (queueing included);
`Server.run` returns per-request records and p50/p99 summaries.
"""
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
    built: bool = False           # the request's space had to be built
    phases: dict = None           # seconds per phase: space, geometry, setup, solve
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


def cost(n, p):
    """Work estimate of a request: DOFs x order x cells per edge (PCG
    iterations grow with both)."""
    return (p * n + 1) ** 3 * p * n


def median_cost(n_range=(8, 32), p_range=(3, 8)):
    return float(np.median([cost(n, p) for n in range(n_range[0], n_range[1] + 1)
                            for p in range(p_range[0], p_range[1] + 1)]))


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
    def __init__(self, device=0, slots=8, small_slots=None, small_cost=None, hbm_budget=None, tol=1e-10,
                 keep_solutions=False, classes=True):
        """slots workers; the first small_slots of them (default half) take only
        requests whose cost is at most small_cost (default: the median over the
        request distribution), so small requests never queue behind large
        ones; the other workers take the earliest large request, else the
        earliest small one."""
        import torch
        self.device = device
        self.slots = slots
        self.small_slots = slots // 2 if small_slots is None else small_slots
        self.small_cost = median_cost() if small_cost is None else small_cost
        if hbm_budget is None:
            free, _ = torch.cuda.mem_get_info(device)
            hbm_budget = int(0.7 * free)
        self.budget = hbm_budget
        self.tol = tol
        self.keep = keep_solutions
        self.classes = classes
        # cached spaces per (n, p), shared by all workers; each space has its
        # own context (stream) and serves one request at a time
        self.pool = {}                                # (n, p) -> [[space, ctx, busy, last use], ...]
        self.in_use = 0                               # estimated bytes of all spaces held
        self.lock = threading.Condition()
        self.pending = []
        self.t0 = None

    # ------------------------------------------------------------- memory --
    def _reserve(self, key):
        """With the lock held: an entry of the pool for topology key, marked
        busy (an idle cached space, or a placeholder to build when the budget
        allows after evicting idle spaces, least recently used first); None
        when the request must wait.  An idle device takes any request."""
        for ent in self.pool.get(key, []):
            if not ent[2]:
                ent[2] = True
                return ent
        need = footprint(*key)
        while self.in_use + need > self.budget:
            victims = [(ent[3], k, i) for k, lst in self.pool.items() for i, ent in enumerate(lst) if not ent[2]]
            if not victims:
                if any(ent[2] for lst in self.pool.values() for ent in lst):
                    return None
                break
            _, k, i = min(victims)
            sp, ctx, _, _ = self.pool[k].pop(i)
            sp.close()
            ctx.close()
            self.in_use -= footprint(*k)
        self.in_use += need
        ent = [None, None, True, time.perf_counter()]
        self.pool.setdefault(key, []).append(ent)
        return ent

    # ------------------------------------------------------------ workers --
    def _next(self, w):
        """The next request for worker w (arrival order within its lane) and
        its pool entry, or None when no request is left."""
        small_lane = w < self.small_slots
        with self.lock:
            while True:
                if not self.pending:
                    return None
                now = time.perf_counter() - self.t0
                arrived = [r for r in self.pending if r.arrival <= now]
                small = [r for r in arrived if cost(r.n, r.p) <= self.small_cost]
                large = [r for r in arrived if cost(r.n, r.p) > self.small_cost]
                pick = (small[:1] if small_lane else (large[:1] or small[:1]))
                if pick:
                    ent = self._reserve((pick[0].n, pick[0].p))
                    if ent is not None:
                        req = pick[0]
                        self.pending.remove(req)
                        req.start = now
                        return req, ent
                # wait for an arrival or a finished request
                upcoming = [r.arrival - now for r in self.pending if r.arrival > now]
                self.lock.wait(timeout=min(upcoming + [0.05]))

    def _space(self, ent, n, p, req=None):
        if ent[0] is None:
            if req is not None:
                req.built = True
            ctx = api.Context(self.device)
            if self.classes:
                ctx.set_class(1 if cost(n, p) <= self.small_cost else 2)
            xyz, tet = mesh.kuhn_cube(n)
            sp = api.Space(ctx, p, xyz, tet, qorder=p + 3)
            with self.lock:
                ent[0], ent[1] = sp, ctx
        return ent[0]

    def solve_one(self, ent, req):
        """Setup and solve of one request on the space of a pool entry.  A
        foreground request holds a foreground span from its first device
        work to its solution; a background one waits before each phase while
        foreground work is active (up to 50 ms), and its solve pauses between
        chunks for at most the previous chunk's duration."""
        t0 = time.perf_counter()
        sp = self._space(ent, req.n, req.p, req)
        ctx = ent[1]
        fg = self.classes and cost(req.n, req.p) <= self.small_cost
        if fg:
            ctx.foreground_begin()
        try:
            t1 = time.perf_counter()
            ctx.background_wait(0.05)
            xyz, _ = mesh.kuhn_cube(req.n, jitter=0.1, seed=req.seed)
            sp.set_geometry(xyz)
            t2 = time.perf_counter()
            ctx.background_wait(0.05)
            # the load vector in pieces of about 2M points: torch work of all
            # workers shares one stream, so no request queues much ahead of
            # another; background requests pause between the pieces
            chunk = max(1, int(2e6 // sp.nq3))
            b = api.poisson_setup(sp, problems.Manufactured(req.seed), chunk=chunk,
                                  pause=None if fg else (lambda: ctx.background_wait(0.01)))
            t3 = time.perf_counter()
            x, st = sp.pcg(b, tol=self.tol, criterion=1, precond=1, mixed=True, inner_tol=1e-2, check_every=10,
                           max_iter=100000)
        finally:
            if fg:
                ctx.foreground_end()
        req.phases = dict(space=t1 - t0, geometry=t2 - t1, setup=t3 - t2, solve=time.perf_counter() - t3)
        req.dofs, req.iterations, req.relres = sp.ndof, st["iterations"], st["relres"]
        if self.keep:
            req.solution = x.cpu().numpy()
        return req

    def _worker(self, w):
        import torch
        torch.cuda.set_device(self.device)
        while True:
            nxt = self._next(w)
            if nxt is None:
                return
            req, ent = nxt
            try:
                self.solve_one(ent, req)
            except Exception as e:  # reported by run()
                req.error = repr(e)
            finally:
                req.done = time.perf_counter() - self.t0
                req.worker = w
                with self.lock:
                    if ent[0] is None:  # build failed: give the memory back
                        self.pool[(req.n, req.p)].remove(ent)
                        self.in_use -= footprint(req.n, req.p)
                    ent[2] = False
                    ent[3] = time.perf_counter()
                    self.lock.notify_all()

    def run(self, requests):
        """Serve the requests (arrival times relative to now); returns them
        with timings filled in, plus a summary."""
        self.pending = sorted(requests, key=lambda r: (r.arrival, r.id))
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
        with self.lock:
            for lst in self.pool.values():
                for sp, ctx, _, _ in lst:
                    if sp is not None:
                        sp.close()
                        ctx.close()
            self.pool.clear()
            self.in_use = 0
