"""GPU parity of the HIP path against the CPU oracle (same seeded meshes).

Bars: gather and scatter bit-exact (integer index work and fixed-order FP64
sums); operator apply within 1e-13 (FP64) / 1e-6 (FP32, on the FP32-rounded input)
of the largest output entry; solves within 1e-10 relative of the
oracle's solution when both are driven to a relative residual of 1e-13.
"""
import numpy as np
import pytest

import harness
import oracle as O
from tsem_amd import api, mesh, problems

pytestmark = pytest.mark.gpu

# n^3 Kuhn cells, order M, jitter, seed; M = 10 and 12 (K = 286, 455) run the
# LDS element kernel (the MFMA kernels take K <= 192)
CASES = [(2, 4, 0.1, 3), (3, 6, 0.15, 5), (2, 8, 0.1, 11), (2, 3, 0.1, 2), (1, 10, 0.1, 4), (1, 12, 0.1, 9)]


def _say(*a):
    print("[case]", *a, flush=True)


@pytest.fixture(scope="module", params=CASES, ids=lambda c: "n%d_p%d" % (c[0], c[1]))
def pair(request, ctx):
    n, M, jit, seed = request.param
    _say("setup", request.param)
    xyz, tet = mesh.kuhn_cube(n, jitter=jit, seed=seed)
    osp = O.OracleSpace(M, xyz, tet, qorder=M + 3)
    psp = api.Space(ctx, M, xyz, tet, qorder=M + 3)
    return osp, psp


def test_gather_bit_exact(pair):
    import torch
    osp, psp = pair
    _say("gather", psp.M, psp.ndof)
    x = np.random.default_rng(1).standard_normal(psp.ndof)
    xl_o = osp.gather(x)
    xl_p = psp.gather(torch.from_numpy(x).cuda()).cpu().numpy()
    assert np.array_equal(xl_o, xl_p)


def test_scatter_bit_exact(pair):
    import torch
    osp, psp = pair
    _say("scatter", psp.M, psp.ndof)
    off, ind, val = osp.maps()
    yl = np.random.default_rng(2).standard_normal((osp.nt, osp.K))
    # oracle order: element by element, mode by mode, term by term
    y = np.zeros(osp.ndof)
    flat = yl.ravel()
    for s in range(len(flat)):
        for t in range(off[s], off[s + 1]):
            y[ind[t]] += val[t] * flat[s]
    yp = psp.scatter(torch.from_numpy(yl).cuda()).cpu().numpy()
    assert np.array_equal(y, yp)


# FP32 bound: 1e-6 of max |y| for every FP32 element kernel (scalar, FP32
# MFMA, and the fp16-split MFMA kernel, whose two-term operands carry about 22
# bits: a CPU model of its arithmetic gives 1-2e-7 on these meshes)
FP32_BOUND = {0: 1e-6, 1: 1e-6, 2: 1e-6}


@pytest.mark.parametrize("path", [0, 1, 2], ids=["mfma_fp16split", "scalar", "mfma_fp32"])
@pytest.mark.parametrize("coef", [(1.0, 0.0), (0.0, 1.0), (0.5, -2.0)])
def test_apply_matches_assembled(pair, coef, path):
    import torch
    from tsem_amd import lib
    osp, psp = pair
    _say("apply", psp.M, coef, "path", path)
    lib.check(lib.load().tsem_set_apply_path(path))
    try:
        osp.assemble(*coef)
        x = np.random.default_rng(3).standard_normal(osp.ndof)
        y_ref = osp.vmult(x)
        y64 = psp.apply(torch.from_numpy(x).cuda(), stiff=coef[0], mass=coef[1]).cpu().numpy()
        scale = np.abs(y_ref).max()
        assert np.abs(y64 - y_ref).max() <= 1e-13 * scale
        # FP32: against the exact operator on the FP32-rounded input, so the
        # bound measures the kernel's arithmetic, not the input rounding
        x32 = x.astype(np.float32)
        y_ref32 = osp.vmult(x32.astype(np.float64))
        y32 = psp.apply(torch.from_numpy(x32).cuda(), stiff=coef[0], mass=coef[1]).double().cpu().numpy()
        err32 = np.abs(y32 - y_ref32).max() / np.abs(y_ref32).max()
        print("[result] FP32 apply M %d coef %s path %d: max error %.2e of max |y|" % (psp.M, coef, path, err32),
              flush=True)
        assert err32 <= FP32_BOUND[path]
    finally:
        lib.check(lib.load().tsem_set_apply_path(0))


def test_apply32_slab_matches_csr_scatter(pair):
    """The FP32 operator with entity-slab element results (the default) and
    with element-major results through the CSR scatter: same operator to
    FP32 rounding (the sums run in another order), each deterministic."""
    import torch
    from tsem_amd import lib
    osp, psp = pair
    _say("slab", psp.M, psp.ndof)
    L = lib.load()
    active, total, _, jt = psp.slab_state()
    # slab tables whenever the transposed face rows have at most 8 terms
    # (orders up to 10); written by the matrix-core element kernel (K <= 192)
    assert total == (psp.nt * psp.K if psp.M <= 10 else 0), (total, jt)
    assert active == (psp.K <= 192)
    x = torch.from_numpy(np.random.default_rng(5).standard_normal(psp.ndof)).float().cuda()
    try:
        out = {}
        for slab in (1, 0):
            lib.check(L.tsem_set_slab(slab))
            a = psp.apply(x).double().cpu().numpy()
            b = psp.apply(x).double().cpu().numpy()
            assert np.array_equal(a, b)
            out[slab] = a
        scale = np.abs(out[0]).max()
        assert np.abs(out[1] - out[0]).max() <= 2e-6 * scale
    finally:
        lib.check(L.tsem_set_slab(1))


def test_apply_deterministic(pair):
    import torch
    osp, psp = pair
    x = torch.from_numpy(np.random.default_rng(4).standard_normal(psp.ndof)).cuda()
    a = psp.apply(x).cpu().numpy()
    b = psp.apply(x).cpu().numpy()
    assert np.array_equal(a, b)


@pytest.mark.parametrize("n,M,jit,seed", [(2, 4, 0.1, 3), (3, 5, 0.1, 8)])
def test_poisson_solution_matches_oracle(ctx, n, M, jit, seed):
    import torch
    _say("poisson", n, M)
    prob = problems.Manufactured(1234)
    xyz, tet = mesh.kuhn_cube(n, jitter=jit, seed=seed)
    q = M + 4
    S, b_o = harness.setup_poisson(M, xyz, tet, prob, qorder=q)
    x_o, st_o = S.pcg(b_o, tol=1e-13, criterion=1)
    assert st_o["converged"]
    psp = api.Space(ctx, M, xyz, tet, qorder=q)
    b = api.poisson_setup(psp, prob)
    # load vector and eliminated right-hand side agree to round-off
    assert np.abs(b.cpu().numpy() - b_o).max() <= 1e-12 * np.abs(b_o).max()
    x, st = psp.pcg(b, tol=1e-13, criterion=1)
    assert st["converged"] and st["relres"] <= 1e-12
    xp = x.cpu().numpy()
    assert np.linalg.norm(xp - x_o) <= 1e-10 * np.linalg.norm(x_o)
    # reference stopping rule (absolute l1 of the recursive residual) also works
    x2, st2 = psp.pcg(b, tol=1e-10, criterion=0)
    assert st2["converged"] and st2["l1"] < 1e-10


@pytest.mark.parametrize("mode,inner_tol", [(1, 1e-2), (1, 1e-4), (2, 1e-4)])
@pytest.mark.parametrize("precond", [1, 4])
def test_mixed_precision_reaches_fp64_accuracy(ctx, precond, mode, inner_tol):
    """Both mixed forms reach the FP64 solve's accuracy; reliable updates keep
    the FP32 recursion, so they need about the FP64 iteration count (with the
    FP32 V-cycle only when the updates come at least every 10^-2 of residual
    reduction: it drifts from conjugacy over longer stretches)."""
    _say("mixed", precond, mode, inner_tol)
    prob = problems.Manufactured(1234)
    M = 5
    xyz, tet = mesh.kuhn_cube(3, jitter=0.1, seed=9)
    psp = api.Space(ctx, M, xyz, tet, qorder=M + 4)
    b = api.poisson_setup(psp, prob)
    x64, s64 = psp.pcg(b, tol=1e-13, criterion=1, precond=precond)
    xm, sm = psp.pcg(b, tol=1e-13, criterion=1, precond=precond, mixed=mode, inner_tol=inner_tol, check_every=5)
    assert s64["converged"] and sm["converged"] and sm["relres"] <= 1e-13, sm
    a, c = x64.cpu().numpy(), xm.cpu().numpy()
    print("[result] mixed %d precond %d: %.2e" % (mode, precond, np.linalg.norm(a - c) / np.linalg.norm(a)),
          flush=True)
    assert np.linalg.norm(a - c) <= 1e-10 * np.linalg.norm(a)
    if mode == 1 and (precond == 1 or inner_tol >= 1e-2):
        assert sm["outer"] >= 1 and sm["iterations"] <= 1.15 * s64["iterations"] + 10, (sm, s64)


def test_pmg_matches_jacobi_solution(ctx):
    import torch
    _say("pmg")
    prob = problems.Manufactured(1234)
    M = 6
    xyz, tet = mesh.kuhn_cube(3, jitter=0.1, seed=10)
    psp = api.Space(ctx, M, xyz, tet, qorder=M + 3)
    b = api.poisson_setup(psp, prob)
    xj, sj = psp.pcg(b, tol=1e-13, criterion=1, precond=1)
    xm, sm = psp.pcg(b, tol=1e-13, criterion=1, precond=4, smooth_steps=2)
    assert sj["converged"] and sm["converged"]
    assert sm["iterations"] < sj["iterations"] / 3
    a, c = xj.cpu().numpy(), xm.cpu().numpy()
    assert np.linalg.norm(a - c) <= 1e-10 * np.linalg.norm(a)


def test_dirichlet_reimposed_after_zero_elimination(ctx):
    """Data imposed on DOFs that an earlier call already eliminated (the
    Hartree solve of every SCF step) give the same system as on a fresh space."""
    prob = problems.Manufactured(1234)
    xyz, tet = mesh.kuhn_cube(2, jitter=0.1, seed=12)
    a = api.Space(ctx, 4, xyz, tet, qorder=8)
    b = api.Space(ctx, 4, xyz, tet, qorder=8)
    ba = api.poisson_setup(a, prob)
    b.set_cube_marks()
    b.zero_bc(1)
    bb = api.poisson_setup(b, prob)
    assert np.abs(ba.cpu().numpy() - bb.cpu().numpy()).max() <= 1e-13 * np.abs(ba.cpu().numpy()).max()
    xa, _ = a.pcg(ba, tol=1e-12)
    xb, _ = b.pcg(bb, tol=1e-12)
    assert np.abs(xa.cpu().numpy() - xb.cpu().numpy()).max() <= 1e-11 * np.abs(xa.cpu().numpy()).max()


@pytest.mark.parametrize("precond", [1, 4])
def test_graph_replays_match_eager(ctx, precond):
    """PCG chunks replayed as captured graphs (many replays per solve, eager
    copies and synchronizations between them, the cached graph reused by the
    next solve) give the iterates of eager launches, bit for bit."""
    from tsem_amd import lib
    _say("graphs", precond)
    prob = problems.Manufactured(1234)
    xyz, tet = mesh.kuhn_cube(3, jitter=0.1, seed=13)
    psp = api.Space(ctx, 5, xyz, tet, qorder=8)
    b = api.poisson_setup(psp, prob)
    res = {}
    try:
        for g in (0, 1, 1):
            lib.check(lib.load().tsem_set_graphs(g))
            for mixed in (False, True):
                x, st = psp.pcg(b, tol=1e-11, precond=precond, mixed=mixed, check_every=7)
                res.setdefault((g, mixed), []).append((x.cpu().numpy(), st["iterations"]))
    finally:
        lib.check(lib.load().tsem_set_graphs(1))
    for mixed in (False, True):
        x0, i0 = res[(0, mixed)][0]
        for x1, i1 in res[(1, mixed)]:
            assert i1 == i0 and np.array_equal(x1, x0)


@pytest.mark.parametrize("mixed", [False, True])
def test_graphs_follow_pmg_rebuilds(ctx, mixed):
    """A captured PCG chunk is never replayed after the p-multigrid levels it
    captured were rebuilt: p-MG solve, then a new geometry and new boundary
    data on the same space (both drop the levels), then a p-MG solve again.
    With graphs on, both solves give the iterates of eager launches, bit
    for bit, and the second one the solution of a fresh space."""
    from tsem_amd import lib
    _say("graphs after rebuild", mixed)
    prob = problems.Manufactured(1234)
    xyz, tet = mesh.kuhn_cube(3, jitter=0.1, seed=13)
    xyz2, _ = mesh.kuhn_cube(3, jitter=0.1, seed=14)
    out = {}
    try:
        for g in (0, 1):
            lib.check(lib.load().tsem_set_graphs(g))
            psp = api.Space(ctx, 5, xyz, tet, qorder=8)
            b = api.poisson_setup(psp, prob)
            x1, s1 = psp.pcg(b, tol=1e-11, precond=4, mixed=mixed, check_every=7)
            psp.set_geometry(xyz2)
            b2 = api.poisson_setup(psp, prob)
            x2, s2 = psp.pcg(b2, tol=1e-11, precond=4, mixed=mixed, check_every=7)
            out[g] = (x1.cpu().numpy(), x2.cpu().numpy(), s1["iterations"], s2["iterations"])
            psp.close()
    finally:
        lib.check(lib.load().tsem_set_graphs(1))
    assert out[0][2:] == out[1][2:]
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])
    fresh = api.Space(ctx, 5, xyz2, tet, qorder=8)
    bf = api.poisson_setup(fresh, prob)
    xf, _ = fresh.pcg(bf, tol=1e-11, precond=4, mixed=mixed, check_every=7)
    assert np.linalg.norm(xf.cpu().numpy() - out[1][1]) <= 1e-9 * np.linalg.norm(xf.cpu().numpy())
    fresh.close()


@pytest.mark.parametrize("mixed", [False, True])
def test_pmg_levels_kept_for_new_boundary_data(ctx, mixed):
    """New Dirichlet data on the parts eliminated before leave the
    elimination mask as it was, so the space keeps its p-multigrid levels
    and captured graphs: a second problem solved on the same space after a
    first one gives, with graphs on and off alike (bit for bit), the
    solution of a fresh space."""
    from tsem_amd import lib
    _say("p-MG kept for new boundary data", mixed)
    prob1, prob2 = problems.Manufactured(1234), problems.Manufactured(99)
    xyz, tet = mesh.kuhn_cube(3, jitter=0.1, seed=13)
    out = {}
    try:
        for g in (0, 1):
            lib.check(lib.load().tsem_set_graphs(g))
            psp = api.Space(ctx, 5, xyz, tet, qorder=8)
            b1 = api.poisson_setup(psp, prob1)
            psp.pcg(b1, tol=1e-11, precond=4, mixed=mixed, check_every=7)
            b2 = api.poisson_setup(psp, prob2)
            x2, s2 = psp.pcg(b2, tol=1e-11, precond=4, mixed=mixed, check_every=7)
            out[g] = (x2.cpu().numpy(), s2["iterations"])
            psp.close()
    finally:
        lib.check(lib.load().tsem_set_graphs(1))
    assert out[0][1] == out[1][1] and np.array_equal(out[0][0], out[1][0])
    fresh = api.Space(ctx, 5, xyz, tet, qorder=8)
    bf = api.poisson_setup(fresh, prob2)
    xf, _ = fresh.pcg(bf, tol=1e-11, precond=4, mixed=mixed, check_every=7)
    xf = xf.cpu().numpy()
    fresh.close()
    assert np.linalg.norm(xf - out[1][0]) <= 1e-12 * np.linalg.norm(xf)
