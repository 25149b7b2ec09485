"""TEST INFRASTRUCTURE ONLY: mixed-BC Poisson problems solved by the oracle.

Mirrors the driver examples/poisson_mixBC/poisson.cpp:74-89 of TSEMLib:
stiffness matrix, calc_rhs(f), Neumann load on mark 2, Dirichlet imposition on
mark 1 (matrix modified), then solve_PCG.
"""
import numpy as np
import scipy.sparse

from oracle import OracleSpace
from tsem_amd import mesh as meshmod


def bc_points(S):
    """Physical points where the BC data is sampled: vertices, edge and face
    quadrature points in the orientation the reference uses."""
    o = S.orientation()
    t1, w1, p2, w2 = S.bc_quad()
    xyz = S.xyz
    s, e = xyz[o["num_node_e"][:, 0]], xyz[o["num_node_e"][:, 1]]
    pe = s[:, None, :] * (1 - t1)[None, :, None] + e[:, None, :] * t1[None, :, None]
    n0, n1, n2 = (xyz[o["num_node_f"][:, k]] for k in range(3))
    pf = (n0[:, None, :] * p2[None, :, 2, None] + n1[:, None, :] * p2[None, :, 0, None]
          + n2[:, None, :] * p2[None, :, 1, None])
    return xyz, pe, pf, o


def setup_poisson(M, xyz, tet, problem, qorder=None, dirichlet_planes=0b000001, neumann_planes=0b111110):
    """Build the oracle system A x = b of the mixed-BC Poisson problem.

    Returns (S, b) with S's CSR modified by the Dirichlet elimination."""
    S = OracleSpace(M, xyz, tet, qorder=qorder)
    ev, fv, fe, tf = S.topology()
    vp = meshmod.cube_plane_marks(xyz)
    mv, me, mf = meshmod.geometry_marks(vp, ev, fv, dirichlet_planes, neumann_planes)
    S.set_marks(mv, me, mf)
    S.marks = (mv, me, mf)
    S.assemble(1.0, 0.0)
    qp = S.qp_coords()
    b = S.rhs(problem.f(qp))
    pv, pe, pf, o = bc_points(S)
    normals = __import__("tsem_amd.problems", fromlist=["x"]).cube_face_normal(xyz[fv])
    gq = problem.dudn(pf, [normals[:, None, d] for d in range(3)])
    S.neumann(gq, 2, b)
    S.bval = S.dirichlet(problem.u(pv), problem.u(pe), problem.u(pf), 1, b, modify=True)
    return S, b


def l2_error(S, x, problem):
    """sqrt(sum_e V_e sum_q w_q (u_h - u)^2) with the space's 3-D rule."""
    pts, w = S.quad3()
    v, _ = S.basis_at(pts)
    xl = S.gather(x)
    uh = xl @ v.T
    qp = S.qp_coords()
    u = problem.u(qp)
    vol = S.volumes()
    return float(np.sqrt(np.sum(vol[:, None] * w[None, :] * (uh - u) ** 2)))


def dirichlet_pencil(n, M, jitter=0.0, seed=0):
    """Oracle S and M on the unit cube with zero Dirichlet data on all faces,
    restricted to the free DOFs (dense).  Returns (S_ff, M_ff, free mask,
    xyz, tet)."""
    xyz, tet = meshmod.kuhn_cube(n, jitter=jitter, seed=seed)
    S = OracleSpace(M, xyz, tet, qorder=M + 3)
    ev, fv, _, _ = S.topology()
    mv, me, mf = meshmod.geometry_marks(meshmod.cube_plane_marks(xyz), ev, fv, 0b111111, 0)
    S.set_marks(mv, me, mf)
    start = S.dof_start()
    nv, ne, nf = len(mv), len(me), len(mf)
    K = [1, M - 1, (M - 1) * (M - 2) // 2]
    free = np.ones(S.ndof, bool)
    for d, (marks, off) in enumerate(((mv, 0), (me, nv), (mf, nv + ne))):
        for g in np.nonzero(marks == 1)[0]:
            free[start[off + g]:start[off + g] + K[d]] = False
    mats = []
    for coef in ((1.0, 0.0), (0.0, 1.0)):
        S.assemble(*coef)
        S.zero_bc(1)
        rs, col, v = S.csr()
        mats.append(scipy.sparse.csr_matrix((v, col, rs), shape=(S.ndof, S.ndof)).toarray()[np.ix_(free, free)])
    return mats[0], mats[1], free, xyz, tet
