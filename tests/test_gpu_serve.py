"""Served solve stream: results are the same alone and in the mix, and the
HBM budget only delays requests."""
import numpy as np
import pytest

from tsem_amd import serve

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[case]", *a, flush=True)


def _reqs():
    return serve.make_requests(6, seed=3, n_range=(3, 5), p_range=(3, 5))


def test_results_same_alone_and_in_the_mix(ctx):
    _say("serve mix")
    mixed = serve.Server(slots=3, keep_solutions=True)
    reqs, summary = mixed.run(_reqs())
    mixed.close()
    assert summary["max_relres"] <= 1e-10
    assert len({r.worker for r in reqs}) > 1  # the work was spread over several streams
    alone = serve.Server(slots=1, keep_solutions=True)
    for r in _reqs():
        got, _ = alone.run([r])
        ref = reqs[r.id].solution
        # same space, geometry and deterministic kernels: bit-identical
        assert np.array_equal(got[0].solution, ref), r
    alone.close()


def test_budget_serializes_and_evicts(ctx):
    _say("serve budget")
    reqs = _reqs()
    small = max(serve.footprint(r.n, r.p) for r in reqs)
    s = serve.Server(slots=3, hbm_budget=small)
    reqs, summary = s.run(reqs)
    assert summary["max_relres"] <= 1e-10
    # with room for one space at a time, no two requests overlap
    spans = sorted((r.start, r.done) for r in reqs)
    assert all(spans[i + 1][0] >= spans[i][1] - 1e-6 for i in range(len(spans) - 1))
    assert s.in_use <= small
    s.close()


def test_foreground_and_background_classes(ctx):
    """With small topologies foreground and large ones background (large
    solves pause between chunks while a small one runs), every request still
    gets the bit-identical result it gets alone."""
    _say("serve classes")
    reqs = serve.make_requests(8, seed=5, n_range=(3, 6), p_range=(3, 6))
    costs = sorted(serve.cost(r.n, r.p) for r in reqs)
    split = costs[len(costs) // 2]
    assert costs[0] <= split < costs[-1]
    mixed = serve.Server(slots=4, small_slots=2, small_cost=split, keep_solutions=True)
    got, summary = mixed.run(reqs)
    mixed.close()
    assert summary["max_relres"] <= 1e-10
    alone = serve.Server(slots=1, keep_solutions=True, classes=False)
    for r in serve.make_requests(8, seed=5, n_range=(3, 6), p_range=(3, 6)):
        (one,), _ = alone.run([r])
        assert np.array_equal(one.solution, got[r.id].solution), (r, one.relres, got[r.id].relres)
    alone.close()


# ------------------------------------------------ C ABI served stream ----
def _python_solution(ctx, n, p, seed):
    """The request solved on its own through the Python path (poisson_setup
    and the mixed PCG the Python server runs)."""
    from tsem_amd import api, mesh, problems
    xyz, tet = mesh.kuhn_cube(n, jitter=0.1, seed=seed)
    sp = api.Space(ctx, p, xyz, tet, qorder=p + 3)
    b = api.poisson_setup(sp, problems.Manufactured(seed))
    x, st = sp.pcg(b, tol=1e-10, criterion=1, precond=1, mixed=True, inner_tol=1e-2, check_every=10,
                   max_iter=100000)
    out = x.cpu().numpy()
    sp.close()
    return out, st


def _submit(srv, n, p, seed, topology=0):
    from tsem_amd import mesh, problems
    xyz, tet = mesh.kuhn_cube(n, jitter=0.1, seed=seed)
    return srv.submit(xyz, tet, p, problems.Manufactured(seed), labels=mesh.cube_plane_marks(xyz),
                      topology=topology)


def test_native_server_matches_python_path(ctx):
    """tsem_serve_*: requests of several sizes in flight together on four
    workers give the Python path's solutions.  Load and boundary data are
    assembled by other code paths (C++ normals, chunked load), so the two
    solves agree to their tolerance (relative residual 1e-10; the solutions
    to 1e-9), not bit for bit."""
    _say("native serve mix")
    reqs = [(4, 3, 1), (6, 4, 2), (3, 5, 3), (5, 6, 4), (7, 3, 5), (4, 4, 6)]
    srv = serve.NativeServer(slots=4, small_cost=3e4)
    ids = [_submit(srv, n, p, seed) for n, p, seed in reqs]
    got = [srv.poll(i) for i in ids]
    srv.close()
    for (n, p, seed), (done, x, r) in zip(reqs, got):
        ref, st = _python_solution(ctx, n, p, seed)
        assert done and r["status"] == 0 and r["built"] == 1
        assert r["n_dof"] == len(ref) and r["relres"] <= 1e-10, r
        assert np.abs(x - ref).max() <= 1e-9 * np.abs(ref).max(), (n, p, seed)


def test_native_server_reuses_cached_space(ctx):
    """A request with the topology key of a finished one reuses its space
    (only the coordinates change) and still matches the Python path."""
    _say("native serve reuse")
    srv = serve.NativeServer(slots=2)
    a = _submit(srv, 4, 4, 11, topology=7)
    _, xa, ra = srv.poll(a)
    b = _submit(srv, 4, 4, 12, topology=7)
    _, xb, rb = srv.poll(b)
    srv.close()
    assert ra["built"] == 1 and rb["built"] == 0
    for x, seed in ((xa, 11), (xb, 12)):
        ref, _ = _python_solution(ctx, 4, 4, seed)
        assert np.abs(x - ref).max() <= 1e-9 * np.abs(ref).max()


def test_native_server_reapplies_marks_on_reuse(ctx):
    """Two requests with one topology key but different boundary parts
    (Dirichlet on x = 0 then on every face): the second reuses the cached
    space and is solved with its own marks, as a fresh space would be."""
    from tsem_amd import api, mesh, problems
    _say("native serve marks")
    srv = serve.NativeServer(slots=1)
    xyz, tet = mesh.kuhn_cube(4, jitter=0.1, seed=21)
    prob = problems.Manufactured(21)
    lab = mesh.cube_plane_marks(xyz)
    a = srv.submit(xyz, tet, 4, prob, labels=lab, topology=9)
    _, xa, ra = srv.poll(a)
    b = srv.submit(xyz, tet, 4, prob, labels=lab, dirichlet=0b111111, neumann=0, topology=9)
    _, xb, rb = srv.poll(b)
    srv.close()
    assert ra["built"] == 1 and rb["built"] == 0 and rb["relres"] <= 1e-10
    for x, dp, nb in ((xa, 0b000001, 0b111110), (xb, 0b111111, 0)):
        sp = api.Space(ctx, 4, xyz, tet, qorder=7)
        bb = api.poisson_setup(sp, prob, dirichlet_planes=dp, neumann_planes=nb)
        ref, _ = sp.pcg(bb, tol=1e-10, criterion=1, mixed=True, check_every=10)
        ref = ref.cpu().numpy()
        sp.close()
        assert np.abs(x - ref).max() <= 1e-9 * np.abs(ref).max(), dp


def test_native_server_reports_callback_failure(ctx):
    """A load callback that fails ends that request with TSEM_EINVAL and its
    message; the server goes on with the next request."""
    from tsem_amd import lib, mesh, problems

    class Broken(problems.Manufactured):
        def f(self, p):
            raise RuntimeError("no load")

    _say("native serve error")
    srv = serve.NativeServer(slots=1)
    xyz, tet = mesh.kuhn_cube(3)
    bad = srv.submit(xyz, tet, 3, Broken(1), labels=mesh.cube_plane_marks(xyz), callbacks=True)
    good = _submit(srv, 3, 3, 2)
    with pytest.raises(lib.TsemError) as e:
        srv.poll(bad)
    assert e.value.code == 2 and "load callback" in str(e.value)
    done, x, r = srv.poll(good)
    srv.close()
    assert done and r["status"] == 0 and r["relres"] <= 1e-10


def test_native_manufactured_matches_callbacks(ctx):
    """A Poisson request whose problem the library evaluates on the device
    (no callbacks) and the same request through the Python callbacks, which
    call the same device evaluator: the same solution bit for bit."""
    from tsem_amd import mesh, problems
    _say("native serve manufactured vs callbacks")
    srv = serve.NativeServer(slots=2)
    a = _submit(srv, 4, 4, 31)
    xyz, tet = mesh.kuhn_cube(4, jitter=0.1, seed=31)
    b = srv.submit(xyz, tet, 4, problems.Manufactured(31), labels=mesh.cube_plane_marks(xyz), callbacks=True)
    _, xa, ra = srv.poll(a)
    _, xb, rb = srv.poll(b)
    srv.close()
    assert ra["relres"] <= 1e-10 and np.array_equal(xa, xb)


def test_native_eigen_requests_in_the_mix(ctx):
    """Eigen requests served among Poisson requests: the 6 lowest
    eigenvalues of the Dirichlet Laplacian (n = 3, p = 4) within 1e-8 of the
    dense generalized eigensolve of the oracle's pencil, and the same
    request alone on a one-worker server gives the same bits."""
    import scipy.linalg as la
    import harness
    from tsem_amd import mesh
    _say("native serve eigen")
    A, B, free, xyz, tet = harness.dirichlet_pencil(3, 4)
    w = la.eigh(A, B, eigvals_only=True, subset_by_index=[0, 5])
    lab = mesh.cube_plane_marks(xyz)
    srv = serve.NativeServer(slots=4, small_cost=3e4)
    e1 = srv.submit_eigen(xyz, tet, 4, 6, labels=lab, seed=3, qorder=7, tol=1e-9)
    pois = [_submit(srv, 5, 4, 7), _submit(srv, 4, 5, 8), _submit(srv, 6, 3, 9)]
    _, X1, r1 = srv.poll(e1)
    for i in pois:
        _, _, r = srv.poll(i)
        assert r["relres"] <= 1e-10
    srv.close()
    lam = r1["eigenvalues"]
    print("[result] served eigenvalues max rel err %.2e" % (np.abs(lam - w) / w).max(), flush=True)
    assert r1["status"] == 0 and np.abs(lam - w).max() <= 1e-8 * w[-1]
    assert X1.shape == (6, r1["n_dof"]) and np.abs(X1[:, ~free]).max() == 0
    alone = serve.NativeServer(slots=1)
    e2 = alone.submit_eigen(xyz, tet, 4, 6, labels=lab, seed=3, qorder=7, tol=1e-9)
    _, X2, r2 = alone.poll(e2)
    alone.close()
    assert np.array_equal(r2["eigenvalues"], lam) and np.array_equal(X2, X1)


def test_native_system_request_matches_direct_solve(ctx):
    """A system request (PCG on the caller's space, as the Kohn-Sham Hartree
    solve uses) served beside other requests equals the same PCG called
    directly, bit for bit."""
    import torch
    from tsem_amd import api, mesh, problems
    _say("native serve system")
    xyz, tet = mesh.kuhn_cube(5, jitter=0.1, seed=41)
    sp = api.Space(ctx, 4, xyz, tet, qorder=7)
    b = api.poisson_setup(sp, problems.Manufactured(41))
    x1, st1 = sp.pcg(b, tol=1e-10, criterion=1, precond=1, mixed=True, inner_tol=1e-2, check_every=10)
    srv = serve.NativeServer(slots=3, small_cost=3e4)
    other = [_submit(srv, 6, 4, 42), _submit(srv, 4, 6, 43)]
    x2 = sp.zeros()
    rid = srv.submit_system(sp, b, x2, tol=1e-10, mixed=1, precond=1)
    done, none, r = srv.poll(rid)
    for i in other:
        srv.poll(i)
    srv.close()
    torch.cuda.synchronize()
    assert done and none is None and r["status"] == 0 and r["iterations"] == st1["iterations"]
    assert torch.equal(x1, x2)
    sp.close()


def test_kohn_sham_hartree_through_the_server(ctx):
    """The helium SCF with every step's Hartree solve served as a system
    request gives the energies of the direct solves (the same PCG, so the
    same bits)."""
    from tsem_amd import ks
    _say("kohn-sham served hartree")
    srv = serve.NativeServer(slots=2)
    a = ks.KohnShamAtom(ctx, n=8, p=3, server=srv)
    ra = a.run(steps=3)
    srv.close()
    b = ks.KohnShamAtom(ctx, n=8, p=3)
    rb = b.run(steps=3)
    ea = [h["energy"] for h in a.history]
    eb = [h["energy"] for h in b.history]
    print("[result] served %s direct %s" % (ea, eb), flush=True)
    assert ea == eb
    a.sp.close()
    b.sp.close()
