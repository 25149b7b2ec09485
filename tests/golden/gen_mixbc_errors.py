"""L2 errors of the oracle's mixed-BC Poisson solution on the reference
example mesh (tests/golden/cube_mixbc.mesh, examples/poisson_mixBC/cube.mesh),
orders 4..10, with the example's u, f, g (poisson.cpp:25-27) and its
quadrature order 30.  Writes tests/golden/mixbc_cube_l2.json.

    python tests/golden/gen_mixbc_errors.py
"""
import json
import math
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import harness  # noqa: E402
import oracle as O  # noqa: E402
from tsem_amd import mesh  # noqa: E402

PI = math.pi


def u(p):
    return np.cos(PI * p[..., 0]) * np.sin(1.5 * PI * p[..., 1]) * np.sin(0.25 * PI + 2 * PI * p[..., 2])


def f(p):
    return 7.25 * PI ** 2 * u(p)


def g(p):
    return -2 * PI * np.cos(PI * p[..., 0]) * np.sin(1.5 * PI * p[..., 1]) * np.cos(0.25 * PI + 2 * PI * p[..., 2])


class _Problem:
    u = staticmethod(u)


def main(orders=(4, 6, 8, 10)):
    path = os.path.join(HERE, "cube_mixbc.mesh")
    out = {"mesh": "cube_mixbc.mesh", "qorder": 30, "errors": {}}
    for M in orders:
        xyz, tet, marks, (ev_f, fv_f, _, _) = mesh.read_afepack_mesh(path)
        S = O.OracleSpace(M, xyz, tet, qorder=30)
        ev, fv, _, _ = S.topology()
        emap = {tuple(sorted(e)): m for e, m in zip(ev_f.tolist(), marks[1])}
        fmap = {tuple(sorted(x)): m for x, m in zip(fv_f.tolist(), marks[2])}
        me = np.array([emap[tuple(sorted(e))] for e in ev.tolist()], np.int32)
        mf = np.array([fmap[tuple(sorted(x))] for x in fv.tolist()], np.int32)
        S.set_marks(marks[0], me, mf)
        S.marks = (marks[0], me, mf)
        S.assemble(1.0, 0.0)
        b = S.rhs(f(S.qp_coords()))
        pv, pe, pf, _ = harness.bc_points(S)
        S.neumann(g(pf), 2, b)
        S.dirichlet(u(pv), u(pe), u(pf), 1, b, modify=True)
        x, st = S.pcg(b, tol=1e-13, criterion=1)
        assert st["converged"], st
        out["errors"][str(M)] = {"n_dof": int(S.ndof), "err_l2": harness.l2_error(S, x, _Problem)}
        print(M, out["errors"][str(M)], flush=True)
    with open(os.path.join(HERE, "mixbc_cube_l2.json"), "w") as fh:
        json.dump(out, fh, indent=1)


if __name__ == "__main__":
    main()
