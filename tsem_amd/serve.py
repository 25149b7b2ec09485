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
    def __init__(self, device=0, slots=4, small_slots=None, small_cost=None, hbm_budget=None, tol=1e-10,
                 keep_solutions=False, classes=True):
        """slots workers; the first small_slots of them (default half) take only
        requests whose cost is at most small_cost (default: the median over the
        request distribution), so small requests never queue behind large
        ones; the other workers take the smallest large request, else the
        smallest small one (shortest job first: a queue of requests of mixed
        sizes waits least in total that way)."""
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
        """The next request for worker w (the smallest arrived one of its
        lane) and its pool entry, or None when no request is left."""
        small_lane = w < self.small_slots
        with self.lock:
            while True:
                if not self.pending:
                    return None
                now = time.perf_counter() - self.t0
                arrived = [r for r in self.pending if r.arrival <= now]
                small = [r for r in arrived if cost(r.n, r.p) <= self.small_cost]
                large = [r for r in arrived if cost(r.n, r.p) > self.small_cost]
                # shortest first within a lane (estimated work), arrival order among equals
                small.sort(key=lambda r: (cost(r.n, r.p), r.arrival))
                large.sort(key=lambda r: (cost(r.n, r.p), r.arrival))
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


# ---------------------------------------------------------------- C ABI ----
import ctypes as C  # noqa: E402

from . import lib as L  # noqa: E402

_LOAD = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p)
_VALUE = C.CFUNCTYPE(C.c_int, C.c_void_p, C.POINTER(C.c_double), C.c_int64, C.POINTER(C.c_double))
_FLUX = C.CFUNCTYPE(C.c_int, C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_int64,
                    C.POINTER(C.c_double))


class ServeOpts(C.Structure):
    _fields_ = [("slots", C.c_int), ("small_slots", C.c_int), ("small_cost", C.c_double),
                ("hbm_budget", C.c_int64), ("classes", C.c_int)]


REQ_POISSON, REQ_EIGEN, REQ_SYSTEM = 0, 1, 2


class ServeRequest(C.Structure):
    _fields_ = [("topology", C.c_int64), ("M", C.c_int), ("qorder", C.c_int), ("nv", C.c_int), ("nt", C.c_int),
                ("xyz", C.c_void_p), ("tet", C.c_void_p), ("labels", C.c_void_p),
                ("dirichlet", C.c_int32), ("neumann", C.c_int32),
                ("f", _LOAD), ("u", _VALUE), ("g", _FLUX), ("user", C.c_void_p),
                ("tol", C.c_double), ("mixed", C.c_int), ("precond", C.c_int),
                ("kind", C.c_int), ("manufactured", C.c_void_p), ("nev", C.c_int), ("seed", C.c_uint64),
                ("eigenvalues", C.c_void_p), ("space", C.c_void_p), ("b_dev", C.c_void_p), ("x_dev", C.c_void_p)]


class ServeResult(C.Structure):
    _fields_ = [("status", C.c_int), ("built", C.c_int), ("n_dof", C.c_int64), ("iterations", C.c_int),
                ("outer", C.c_int), ("relres", C.c_double), ("queued_s", C.c_double), ("space_s", C.c_double),
                ("setup_s", C.c_double), ("solve_s", C.c_double), ("total_s", C.c_double)]


class _DeviceArray:
    """A device buffer of doubles seen by torch (__cuda_array_interface__)."""

    def __init__(self, ptr, shape):
        self.__cuda_array_interface__ = {"shape": shape, "typestr": "<f8", "data": (ptr, False), "version": 3,
                                         "strides": None}


class NativeServer:
    """The served stream of the C ABI (tsem_serve_open / submit / poll /
    close): worker threads, lanes, space pool and solve classes in the
    library.  Poisson requests take a problems.Manufactured (evaluated by
    the library on the device, no Python in the request's path) or any
    `problem` with f, u and du/dn as torch functions of points, called back
    from the workers (f on the library's stream for the request, u and du/dn
    on the device with the values back to the host).  Eigen requests
    (submit_eigen) and system requests on a caller's space (submit_system,
    e.g. a Hartree solve) share the workers and the lanes."""

    def __init__(self, device=0, slots=4, small_slots=-1, small_cost=0.0, hbm_budget=0, classes=True):
        self.device = device
        o = ServeOpts(slots, small_slots, small_cost, hbm_budget, 1 if classes else 0)
        h = C.c_void_p()
        L.check(L.load().tsem_serve_open(device, C.byref(o), C.byref(h)))
        self.h = h
        self.keep = {}  # request id -> the callbacks and arrays it uses

    def _submit(self, r, keep):
        rid = C.c_int64()
        L.check(L.load().tsem_serve_submit(self.h, C.byref(r), C.byref(rid)))
        self.keep[rid.value] = keep
        return rid.value

    def submit(self, xyz, tet, M, problem, labels=None, dirichlet=0b000001, neumann=0b111110, topology=0,
               qorder=0, tol=1e-10, mixed=1, precond=1, callbacks=False):
        """A Poisson request; a problem with a device definition
        (problems.Manufactured) is evaluated by the library unless
        callbacks=True."""
        xyz = np.ascontiguousarray(xyz, dtype=np.float64)
        tet = np.ascontiguousarray(tet, dtype=np.int32)
        lab = None if labels is None else np.ascontiguousarray(labels, dtype=np.int32)
        if hasattr(problem, "definition") and not callbacks:
            d = problem.definition()
            r = ServeRequest(topology, M, qorder, len(xyz), len(tet), xyz.ctypes.data, tet.ctypes.data,
                             None if lab is None else lab.ctypes.data, dirichlet, neumann, _LOAD(), _VALUE(),
                             _FLUX(), None, tol, int(mixed), precond, REQ_POISSON, C.addressof(d), 0, 0, None,
                             None, None, None)
            return self._submit(r, dict(cols=1, keep=(d,)))
        import torch
        dev = "cuda:%d" % self.device

        def load(user, pts, n, out, stream):
            try:
                with torch.cuda.stream(torch.cuda.ExternalStream(stream, device=dev)):
                    p = torch.as_tensor(_DeviceArray(pts, (n, 3)), device=dev)
                    torch.as_tensor(_DeviceArray(out, (n,)), device=dev).copy_(problem.f(p))
                return 0
            except Exception:
                return 1

        def value(user, pts, n, out):
            try:
                p = torch.from_numpy(np.ctypeslib.as_array(pts, (n, 3)).copy()).to(dev)
                np.ctypeslib.as_array(out, (n,))[:] = problem.u(p).cpu().numpy()
                return 0
            except Exception:
                return 1

        def flux(user, pts, nrm, n, out):
            try:
                p = torch.from_numpy(np.ctypeslib.as_array(pts, (n, 3)).copy()).to(dev)
                q = torch.from_numpy(np.ctypeslib.as_array(nrm, (n, 3)).copy()).to(dev)
                np.ctypeslib.as_array(out, (n,))[:] = problem.dudn(p, [q[:, d] for d in range(3)]).cpu().numpy()
                return 0
            except Exception:
                return 1

        cb = (_LOAD(load), _VALUE(value), _FLUX(flux))
        r = ServeRequest(topology, M, qorder, len(xyz), len(tet), xyz.ctypes.data, tet.ctypes.data,
                         None if lab is None else lab.ctypes.data, dirichlet, neumann, cb[0], cb[1], cb[2], None,
                         tol, int(mixed), precond, REQ_POISSON, None, 0, 0, None, None, None, None)
        return self._submit(r, dict(cols=1, keep=cb))

    def submit_eigen(self, xyz, tet, M, nev, labels=None, dirichlet=0b111111, topology=0, qorder=0, tol=1e-8,
                     seed=7):
        """An eigen request: the nev lowest pairs of the Laplacian with zero
        Dirichlet data on the labelled parts in `dirichlet`; poll returns the
        eigenvectors (n_dof x nev, column-major flattened) and the result's
        "eigenvalues"."""
        xyz = np.ascontiguousarray(xyz, dtype=np.float64)
        tet = np.ascontiguousarray(tet, dtype=np.int32)
        lab = None if labels is None else np.ascontiguousarray(labels, dtype=np.int32)
        lam = np.zeros(nev)
        r = ServeRequest(topology, M, qorder, len(xyz), len(tet), xyz.ctypes.data, tet.ctypes.data,
                         None if lab is None else lab.ctypes.data, dirichlet, 0, _LOAD(), _VALUE(), _FLUX(), None,
                         tol, 1, 4, REQ_EIGEN, None, nev, seed, lam.ctypes.data, None, None, None)
        return self._submit(r, dict(cols=nev, eig=lam))

    def submit_system(self, space, b, x, tol=1e-10, mixed=1, precond=1):
        """A system request: PCG on `space` (an api.Space, its current
        elimination) for the device vector b into x (initial guess in,
        solution out); both stay untouched by the caller until poll says
        done.  Device work queued before the call on any stream is finished
        first."""
        import torch
        torch.cuda.synchronize(space.device)
        r = ServeRequest(0, 0, 0, 0, 0, None, None, None, 0, 0, _LOAD(), _VALUE(), _FLUX(), None, tol, int(mixed),
                         precond, REQ_SYSTEM, None, 0, 0, None, space.h.value, b.data_ptr(), x.data_ptr())
        return self._submit(r, dict(cols=0, keep=(b, x)))

    def poll(self, rid, wait=True):
        """(done, x, result dict): x the solution (host) once done -- n_dof
        values, n_dof x nev (row j = eigenvector j) for an eigen request,
        None for a system request (its solution is in the caller's x) -- and
        the request released; (False, None, None) while it runs."""
        res = ServeResult()
        done = C.c_int()
        rc = L.load().tsem_serve_poll(self.h, rid, 1 if wait else 0, None, C.byref(res), C.byref(done))
        if not done.value:
            L.check(rc)
            return False, None, None
        info = self.keep.get(rid, dict(cols=1))
        count = res.n_dof * info["cols"]
        x = np.zeros(max(1, count))
        rc = L.load().tsem_serve_poll(self.h, rid, 1, x.ctypes.data, C.byref(res), C.byref(done))
        self.keep.pop(rid, None)
        L.check(rc)
        out = {k: getattr(res, k) for k, _ in ServeResult._fields_}
        if info["cols"] == 0:
            return True, None, out
        if "eig" in info:
            out["eigenvalues"] = info["eig"].copy()
            return True, x[:count].reshape(info["cols"], res.n_dof), out
        return True, x[:count], out

    def run(self, requests):
        """Serve Requests (n, p, seed, arrival; the Python Server's stream) on
        the C ABI server: meshes are made first, each request is submitted at
        its arrival time (relative to now) with a topology key per (n, p),
        then all are collected.  Fills start/done/dofs/iterations/relres/
        built/phases and returns (requests, summary) as Server.run does."""
        from . import mesh, problems
        data = {r.id: mesh.kuhn_cube(r.n, jitter=0.1, seed=r.seed) for r in requests}
        order = sorted(requests, key=lambda r: (r.arrival, r.id))
        ids = []
        t0 = time.perf_counter()
        for r in order:
            dt = r.arrival - (time.perf_counter() - t0)
            if dt > 0:
                time.sleep(dt)
            xyz, tet = data[r.id]
            sub = time.perf_counter() - t0
            ids.append((r, sub, self.submit(xyz, tet, r.p, problems.Manufactured(r.seed),
                                            labels=mesh.cube_plane_marks(xyz), topology=1000 * r.n + r.p)))
        for r, sub, rid in ids:
            _, _, res = self.poll(rid)
            r.start = sub + res["queued_s"]
            r.done = sub + res["total_s"]
            r.dofs, r.iterations, r.relres, r.built = res["n_dof"], res["iterations"], res["relres"], bool(res["built"])
            r.phases = dict(space=res["space_s"], setup=res["setup_s"], solve=res["solve_s"])
        wall = max(r.done for r in requests)
        lat = np.array([r.done - r.arrival for r in requests])
        return requests, dict(requests=len(requests), wall_s=wall, dofs_per_s=sum(r.dofs for r in requests) / wall,
                              p50_ms=float(np.percentile(lat, 50)) * 1e3,
                              p99_ms=float(np.percentile(lat, 99)) * 1e3,
                              max_relres=max(r.relres for r in requests))

    def close(self):
        if self.h:
            L.check(L.load().tsem_serve_close(self.h))
            self.h = None
            self.keep.clear()


def stream(count=2000, load=0.8, slots=4, seed=2026, sat_count=300, log=print, rate=None, small_slots=-1):
    """The served stream of the C ABI server at one offered load, measured
    end to end: count Poisson requests (n uniform in 8..32, p in 3..8, seed =
    request id, jittered Kuhn meshes, manufactured problems evaluated in the
    library) with Poisson arrivals at `load` x the saturated rate.

    Steps: (1) meshes made up front; (2) one request of every (n, p) bucket
    served, so the pool holds each topology's space (server warm-up); (3)
    saturated rate: the first sat_count requests submitted at once; (4) the
    stream; (5) isolated time per bucket: one request of the bucket alone on
    the idle server, its space warm (the second of two runs).  Latency =
    completion - arrival; each request's latency is set against its bucket's
    isolated time.  Returns a summary dict."""
    import json
    from . import mesh as M, problems as P
    reqs = make_requests(count, seed=seed)
    t = time.perf_counter()
    data = {}
    for i, r in enumerate(reqs):
        xyz, tet = M.kuhn_cube(r.n, jitter=0.1, seed=r.seed)
        data[r.id] = (xyz, tet, M.cube_plane_marks(xyz))
        if (i + 1) % 500 == 0:
            log("stream: %d meshes (%.0f s)" % (i + 1, time.perf_counter() - t))
    srv = NativeServer(slots=slots, small_slots=small_slots,
                       small_cost=float(np.median([6 * n ** 3 * comb(p + 3, 3) for n in range(8, 33)
                                                   for p in range(3, 9)])))

    def submit(r):
        xyz, tet, lab = data[r.id]
        return srv.submit(xyz, tet, r.p, P.Manufactured(r.seed), labels=lab, topology=1000 * r.n + r.p)

    first = {}
    for r in reqs:
        first.setdefault((r.n, r.p), r)
    t = time.perf_counter()
    ids = [submit(r) for r in first.values()]
    for k, rid in enumerate(ids):
        srv.poll(rid)
        if (k + 1) % 25 == 0:
            log("stream: warm-up %d of %d buckets (%.0f s)" % (k + 1, len(ids), time.perf_counter() - t))
    # saturated rate
    t = time.perf_counter()
    ids = [submit(r) for r in reqs[:sat_count]]
    for rid in ids:
        srv.poll(rid)
    sat_wall = time.perf_counter() - t
    sat_rate = len(ids) / sat_wall
    rate = load * sat_rate if not rate else rate
    log("stream: saturated %.2f requests/s over %d requests; offered %.2f requests/s" % (sat_rate, len(ids), rate))
    arrivals = np.cumsum(np.random.default_rng(seed + 1).exponential(1.0 / rate, count))
    sub = [None] * count
    lock = threading.Lock()

    def feeder():
        t0 = feeder.t0
        for i, r in enumerate(reqs):
            dt = arrivals[i] - (time.perf_counter() - t0)
            if dt > 0:
                time.sleep(dt)
            rid = submit(r)
            with lock:
                sub[i] = (time.perf_counter() - t0, rid)

    feeder.t0 = time.perf_counter()
    th = threading.Thread(target=feeder)
    th.start()
    res = [None] * count
    for i in range(count):
        while True:
            with lock:
                s = sub[i]
            if s is not None:
                break
            time.sleep(0.002)
        _, _, out = srv.poll(s[1])
        res[i] = out
        if (i + 1) % 100 == 0:
            log("stream: %d of %d done (%.0f s)" % (i + 1, count, time.perf_counter() - feeder.t0))
    th.join()
    # isolated time per bucket, on the idle server with the space warm
    iso = {}
    for key, r in first.items():
        for _ in range(2):
            _, _, out = srv.poll(submit(r))
        iso[key] = out["total_s"] - out["queued_s"]
    srv.close()
    lat = np.array([sub[i][0] + res[i]["total_s"] - arrivals[i] for i in range(count)])
    ratio = np.array([lat[i] / iso[(r.n, r.p)] for i, r in enumerate(reqs)])
    dofs = sum(res[i]["n_dof"] for i in range(count))
    wall = max(sub[i][0] + res[i]["total_s"] for i in range(count))
    worst = [dict(n=reqs[i].n, p=reqs[i].p, latency_s=float(lat[i]), isolated_s=iso[(reqs[i].n, reqs[i].p)],
                  queued_s=res[i]["queued_s"]) for i in np.argsort(ratio)[::-1][:5]]
    small = np.array([iso[(r.n, r.p)] <= np.median(list(iso.values())) for r in reqs])
    out = dict(workload="served_poisson_n8-32_p3-8", requests=count, slots=slots, small_slots=small_slots, load=load,
               saturated_requests_per_s=sat_rate, offered_requests_per_s=rate, wall_s=wall, dofs_per_s=dofs / wall,
               p50_ms=float(np.percentile(lat, 50)) * 1e3, p99_ms=float(np.percentile(lat, 99)) * 1e3,
               latency_over_isolated={"p50": float(np.percentile(ratio, 50)), "p90": float(np.percentile(ratio, 90)),
                                      "p99": float(np.percentile(ratio, 99)), "max": float(ratio.max()),
                                      "p99_small_half": float(np.percentile(ratio[small], 99)),
                                      "p99_large_half": float(np.percentile(ratio[~small], 99))},
               isolated_ms={"p50": float(np.percentile(list(iso.values()), 50)) * 1e3,
                            "max": float(max(iso.values())) * 1e3, "buckets": len(iso)},
               max_relres=float(max(res[i]["relres"] for i in range(count))), worst=worst)
    log("stream: " + json.dumps(out))
    return out


def main():
    import argparse
    import json
    import sys
    ap = argparse.ArgumentParser(description="served Poisson stream on the C ABI server at one offered load")
    ap.add_argument("--requests", type=int, default=2000)
    ap.add_argument("--load", type=float, default=0.8)
    ap.add_argument("--slots", type=int, default=4)
    ap.add_argument("--sat", type=int, default=300, help="requests of the saturated-rate measurement")
    ap.add_argument("--rate", type=float, default=0.0, help="offered requests/s (default: load x saturated rate)")
    ap.add_argument("--small-slots", type=int, default=-1, help="workers that take only small requests (-1: half)")
    a = ap.parse_args()
    out = stream(a.requests, a.load, a.slots, sat_count=a.sat, rate=a.rate or None, small_slots=a.small_slots,
                 log=lambda s: print("[serve]", s, file=sys.stderr, flush=True))
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
