"""The C++ example programs (tsem_amd/examples, over include/tsem.hpp) on the
device, against the Python path of the same library on the same mesh file.

poisson_mixbc follows examples/poisson_mixBC/poisson.cpp (u, f, g of its
lines 25-27; Dirichlet on mark 1, Neumann on mark 2 of cube.mesh, PCG to
||r||_1 < 1e-14); laplace_eig follows examples/laplaceEig/laplaceEig.cpp
(inverse power, Rayleigh quotient against 3 pi^2).  The meshes are the two
reference examples' cube.mesh files (tests/golden) and a written 3^3 Kuhn
mesh.  The programs are built by tsem_amd.build (build() of the package).
"""
import math
import os
import re
import subprocess

import numpy as np
import pytest

from tsem_amd import api, mesh

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tsem_amd", "examples", "bin")
GOLD = os.path.join(ROOT, "tests", "golden")
PI = math.pi


def _say(*a):
    print("[case]", *a, flush=True)


def _xp(p):
    if isinstance(p, np.ndarray):
        return np
    import torch
    return torch


def u(p):
    xp = _xp(p)
    return xp.cos(PI * p[..., 0]) * xp.sin(1.5 * PI * p[..., 1]) * xp.sin(0.25 * PI + 2 * PI * p[..., 2])


def f(p):
    return 7.25 * PI ** 2 * u(p)


def g(p):
    xp = _xp(p)
    return -2 * PI * xp.cos(PI * p[..., 0]) * xp.sin(1.5 * PI * p[..., 1]) * xp.cos(0.25 * PI + 2 * PI * p[..., 2])


def grad_u(p):
    import torch
    x, y, z = p[..., 0], p[..., 1], p[..., 2]
    return torch.stack([-PI * torch.sin(PI * x) * torch.sin(1.5 * PI * y) * torch.sin(0.25 * PI + 2 * PI * z),
                        1.5 * PI * torch.cos(PI * x) * torch.cos(1.5 * PI * y) * torch.sin(0.25 * PI + 2 * PI * z),
                        2 * PI * torch.cos(PI * x) * torch.sin(1.5 * PI * y) * torch.cos(0.25 * PI + 2 * PI * z)],
                       dim=-1)


def _run(prog, *args):
    exe = os.path.join(BIN, prog)
    if not os.path.exists(exe):
        raise RuntimeError("%s is not built (python -m tsem_amd.build)" % exe)
    _say(prog, *args)
    out = subprocess.run([exe] + [str(a) for a in args], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stderr
    return out.stdout


def _field(text, key):
    return float(re.search(key + r" = ([-+0-9.eE]+)", text).group(1))


def _kuhn_mesh(tmp_path):
    """A jittered 3^3 Kuhn mesh written as .mesh with Dirichlet marks on every
    plane but z = 0 (Neumann), edges and faces in reversed file order."""
    xyz, tet = mesh.kuhn_cube(3, jitter=0.1, seed=11)
    p = api.Space(None, 2, xyz, tet, qorder=4)
    ev, fv, fe, tf = p.topology()
    mv, me, mf = mesh.geometry_marks(mesh.cube_plane_marks(xyz), ev, fv, 0b101111, 0b010000)
    path = str(tmp_path / "kuhn3.mesh")
    mesh.write_afepack_mesh(path, xyz, ev[::-1], fv[::-1], fe[::-1], tet, tf, (mv, me[::-1], mf[::-1]))
    return path


def _python_poisson(ctx, path, M, tol=1e-14):
    xyz, tet, marks, (ev, fv, _, _) = mesh.read_afepack_mesh(path)
    sp = api.Space(ctx, M, xyz, tet, qorder=30)
    sp.set_file_marks(marks, ev, fv)
    b = sp.rhs(f)
    _, _, pf2 = sp.bc_points(2)
    sp.neumann(g(pf2), 2, b)
    pv, pe, pf = sp.bc_points(1)
    sp.dirichlet(u(pv), u(pe), u(pf), 1, b)
    x, st = sp.pcg(b, tol=tol, criterion=0, precond=1, max_iter=100000, check_every=1)
    err = sp.l2_error(x, u)
    _python_poisson.last = dict(err_l2g=sp.l2_error_gradient(x, grad_u), x=x.cpu().numpy())
    n = sp.ndof
    sp.close()
    return err, st["iterations"], n


@pytest.mark.parametrize("M", [4, 6, 8, 10])
def test_poisson_mixbc_example_matches_python(ctx, M, tmp_path):
    path = os.path.join(GOLD, "cube_mixbc.mesh")
    coef = str(tmp_path / "sol.coef")
    out = _run("poisson_mixbc", path, M, 1e-14, 1, coef)
    err_c, its_c, n_c = _field(out, "err_l2"), int(_field(out, "iterations")), int(_field(out, "n_dof"))
    err_p, its_p, n_p = _python_poisson(ctx, path, M)
    assert n_c == n_p
    assert abs(its_c - its_p) <= 2, (its_c, its_p)
    assert abs(err_c - err_p) <= 1e-12 + 1e-7 * err_p, (err_c, err_p)
    # calc_l2_error_gradient through the C++ interface and through Python
    g_c, g_p = _field(out, "err_l2g"), _python_poisson.last["err_l2g"]
    assert abs(g_c - g_p) <= 1e-12 + 1e-7 * g_p, (g_c, g_p)
    # calc_val_point: u_h at a point against the exact u, to the solution's error
    uh_pt = float(re.search(r"u_h\(0.3,0.4,0.5\) = ([-+0-9.eE]+)", out).group(1))
    u_pt = float(re.search(r" u\(0.3,0.4,0.5\) = ([-+0-9.eE]+)", out).group(1))
    assert abs(u_pt - u(np.array([0.3, 0.4, 0.5]))) <= 1e-15
    assert abs(uh_pt - u_pt) <= 50 * err_c, (uh_pt, u_pt, err_c)
    # write_coef: the C++ program's solution as text, read back here; the two
    # solves (host- and device-evaluated load) agree to the solver accuracy
    xyz, tet, _, _ = mesh.read_afepack_mesh(path)
    sp = api.Space(None, M, xyz, tet, qorder=5)
    xc = sp.read_coef(coef)
    xp = _python_poisson.last["x"]
    assert np.abs(xc - xp).max() <= 1e-8 * np.abs(xp).max()
    sp.close()


def test_poisson_mixbc_example_matches_oracle_errors():
    """L2 errors at orders 4..10 against the oracle's on the same mesh and
    data (tests/golden/mixbc_cube_l2.json, gen_mixbc_errors.py); the error
    ratio per two orders grows (spectral convergence)."""
    import json
    gold = json.load(open(os.path.join(GOLD, "mixbc_cube_l2.json")))["errors"]
    path = os.path.join(GOLD, "cube_mixbc.mesh")
    errs = []
    for M in (4, 6, 8, 10):
        out = _run("poisson_mixbc", path, M)
        err, n = _field(out, "err_l2"), int(_field(out, "n_dof"))
        ref = gold[str(M)]
        assert n == ref["n_dof"]
        assert abs(err - ref["err_l2"]) <= 1e-12 + 1e-8 * ref["err_l2"], (M, err, ref)
        errs.append(err)
    ratios = [a / b for a, b in zip(errs, errs[1:])]
    assert all(r > 3 for r in ratios) and ratios == sorted(ratios), (errs, ratios)


def test_poisson_mixbc_example_on_written_mesh(ctx, tmp_path):
    path = _kuhn_mesh(tmp_path)
    out = _run("poisson_mixbc", path, 5, 1e-12, 4)
    err_c = _field(out, "err_l2")
    err_p, _, _ = _python_poisson(ctx, path, 5, tol=1e-12)
    assert abs(err_c - err_p) <= 1e-12 + 1e-7 * err_p, (err_c, err_p)
    assert err_c < 1e-3


def test_laplace_eig_example(ctx):
    path = os.path.join(GOLD, "cube_dirichlet.mesh")
    errs = []
    for M in (4, 6):
        out = _run("laplace_eig", path, M)
        eig = _field(out, "eig")
        errs.append(abs(eig - 3 * PI ** 2))
        xyz, tet, marks, (ev, fv, _, _) = mesh.read_afepack_mesh(path)
        sp = api.Space(ctx, M, xyz, tet, qorder=M + 3)
        sp.set_file_marks(marks, ev, fv)
        sp.zero_bc(1)
        lam, _, _ = sp.inverse_power(tol=1e-12)
        assert abs(eig - lam) <= 1e-9 * lam, (eig, lam)
        if M == 6:
            lam4, _, st = sp.lobpcg(4, tol=1e-10)
            out4 = _run("laplace_eig", path, M, 4)
            eigs = np.array([float(v) for v in re.search(r"eig = ([-+0-9.eE ]+),", out4).group(1).split()])
            assert np.abs(eigs - lam4).max() <= 1e-8 * lam4.max(), (eigs, lam4)
        sp.close()
    assert errs[1] < errs[0] / 10, errs


def test_zero_rows_matches_marks(ctx):
    """impose_zero_boundary_condition(Vector&): exactly the DOFs of the marked
    geometries are zeroed."""
    import torch
    path = os.path.join(GOLD, "cube_mixbc.mesh")
    xyz, tet, marks, (ev, fv, _, _) = mesh.read_afepack_mesh(path)
    sp = api.Space(ctx, 5, xyz, tet, qorder=8)
    sp.set_file_marks(marks, ev, fv)
    x = torch.arange(1, sp.ndof + 1, dtype=torch.float64, device="cuda")
    sp.zero_rows(2, x)
    got = x.cpu().numpy() == 0
    # DOFs of mark-2 geometries, from the space's DOF layout
    start = sp.dof_start()
    base = [0, sp.nv, sp.nv + sp.ne]
    want = np.zeros(sp.ndof, dtype=bool)
    for dim in range(3):
        for gi in np.nonzero(sp.marks[dim] == 2)[0]:
            s = start[base[dim] + gi]
            want[s:s + sp.ndg[dim]] = True
    assert want.sum() > 0 and np.array_equal(got, want)
    sp.close()
