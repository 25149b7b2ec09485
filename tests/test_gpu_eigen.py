"""GPU parity of the eigen solvers (block LOBPCG, inverse power) against the
discrete spectrum of the oracle: a dense generalized eigensolve (LAPACK via
scipy) of the oracle's eliminated S and M on the same mesh.

Bars: eigenvalues within 1e-8 relative; eigenvectors M-orthonormal to 1e-9
and zero on the eliminated DOFs; residuals ||S x - lambda M x|| small.
"""
import numpy as np
import pytest
import scipy.linalg as la

import harness
from tsem_amd import api

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[case]", *a, flush=True)


@pytest.fixture(scope="module")
def pencil(ctx):
    n, M = 3, 4
    _say("eigen setup", n, M)
    A, B, free, xyz, tet = harness.dirichlet_pencil(n, M)
    psp = api.Space(ctx, M, xyz, tet, qorder=M + 3)
    psp.set_cube_marks(0b111111, 0)
    psp.zero_bc(1)
    w = la.eigh(A, B, eigvals_only=True, subset_by_index=[0, 15])
    return psp, A, B, free, w


def _check_block(psp, A, B, free, lam, X, w, m, shift=(1.0, 0.0)):
    a, b = shift
    assert np.abs(lam - (a * w[:m] + b)).max() <= 1e-8 * (a * w[m - 1] + b)
    Xh = X.cpu().numpy()
    assert np.abs(Xh[:, ~free]).max() == 0
    Xf = Xh[:, free]
    G = Xf @ B @ Xf.T
    assert np.abs(G - np.eye(m)).max() < 1e-9
    R = Xf @ (a * A + b * B) - lam[:, None] * (Xf @ B)
    assert np.linalg.norm(R, axis=1).max() <= 1e-6 * lam.max()


# m = 10 ends between the 11 pi^2 triple and the simple 12 pi^2 eigenvalue
@pytest.mark.parametrize("precond", [0, 1, 4])
def test_lobpcg_matches_dense_eigensolve(pencil, precond):
    psp, A, B, free, w = pencil
    _say("lobpcg precond", precond)
    m = 10
    lam, X, st = psp.lobpcg(m, tol=1e-9, precond=precond, max_iter=2000)
    assert st["converged"], st
    _check_block(psp, A, B, free, lam, X, w, m)


def test_lobpcg_shifted_operator(pencil):
    psp, A, B, free, w = pencil
    _say("lobpcg shifted")
    m = 4
    lam, X, st = psp.lobpcg(m, tol=1e-9, precond=4, stiff=0.5, mass=2.0, seed=3)
    assert st["converged"], st
    _check_block(psp, A, B, free, lam, X, w, m, shift=(0.5, 2.0))


def test_lobpcg_reference_criterion(pencil):
    """Stop on the eigenvalue change (the reference rule) instead of residuals."""
    psp, A, B, free, w = pencil
    _say("lobpcg criterion 0")
    m = 7
    lam, X, st = psp.lobpcg(m, tol=1e-11, criterion=0, precond=4)
    assert st["converged"] and st["max_dlambda"] < 1e-11
    assert np.abs(lam - w[:m]).max() <= 1e-8 * w[m - 1]


def test_lobpcg_deterministic(pencil):
    psp, A, B, free, w = pencil
    l1, X1, _ = psp.lobpcg(4, tol=1e-9, precond=4, seed=5)
    l2, X2, _ = psp.lobpcg(4, tol=1e-9, precond=4, seed=5)
    assert np.array_equal(l1, l2) and np.array_equal(X1.cpu().numpy(), X2.cpu().numpy())


def test_lobpcg_vcycle_graph_matches_eager(pencil):
    """The p-multigrid V-cycles of the preconditioner replayed from one
    captured graph (hundreds of replays, eager block work between them) give
    the same eigenpairs bit for bit as the same V-cycles launched eagerly."""
    from tsem_amd import lib
    psp, A, B, free, w = pencil
    _say("lobpcg graphs on / off")
    l1, X1, _ = psp.lobpcg(4, tol=1e-9, precond=4, seed=5)
    lib.check(lib.load().tsem_set_graphs(0))
    try:
        l2, X2, _ = psp.lobpcg(4, tol=1e-9, precond=4, seed=5)
    finally:
        lib.check(lib.load().tsem_set_graphs(1))
    assert np.array_equal(l1, l2) and np.array_equal(X1.cpu().numpy(), X2.cpu().numpy())


def test_lobpcg_block_vcycle_matches_per_column(pencil):
    """The preconditioner as one block V-cycle over all residual columns
    (every kernel one launch over the columns) and as one V-cycle per
    column: the same eigenpairs bit for bit."""
    from tsem_amd import lib
    psp, A, B, free, w = pencil
    _say("lobpcg block V-cycle on / off")
    l1, X1, s1 = psp.lobpcg(6, tol=1e-9, precond=4, seed=11)
    lib.check(lib.load().tsem_set_block_vcycle(0))
    try:
        l2, X2, s2 = psp.lobpcg(6, tol=1e-9, precond=4, seed=11)
    finally:
        lib.check(lib.load().tsem_set_block_vcycle(1))
    assert s1["iterations"] == s2["iterations"]
    assert np.array_equal(l1, l2) and np.array_equal(X1.cpu().numpy(), X2.cpu().numpy())


def test_inverse_power_lowest_eigenvalue(pencil):
    psp, A, B, free, w = pencil
    _say("inverse power")
    lam, x, its = psp.inverse_power(tol=1e-13)
    assert abs(lam - w[0]) <= 1e-10 * w[0]
    xf = x.cpu().numpy()[free]
    assert abs(xf @ B @ xf - 1) < 1e-10
