"""Tetrahedral meshes for the TSEM workloads.

``kuhn_cube`` builds the structured meshes the workloads are quoted on: an
n x n x n grid of hexahedra, each split into the 6 Kuhn tetrahedra that share
the hexahedron's main diagonal.  Optional seeded jitter moves interior
vertices by up to ``jitter * h`` per coordinate and is redrawn until every
tetrahedron keeps a positive orientation.  ``graded_box`` maps the unit grid
onto a box with a tanh grading toward the centre.

Meshes are plain arrays: ``xyz`` (nv, 3) float64 and ``tet`` (nt, 4) int32 in
element vertex order P0..P3.  Edges and faces (and their numbering) are derived
from ``tet`` by the library, as TSEMLib reads them from an AFEPack ``.mesh``
file (src/BuildSpace.cpp:435-560).  ``read_afepack_mesh``/``write_afepack_mesh``
convert to and from that text format.
"""
import itertools

import numpy as np

# Kuhn paths: for a permutation (a, b, c) of the axes the tetrahedron visits
# v0, v0+e_a, v0+e_a+e_b, v0+(1,1,1).
_PERMS = list(itertools.permutations(range(3)))


def _perm_parity(p):
    inv = sum(1 for i in range(3) for j in range(i + 1, 3) if p[i] > p[j])
    return inv % 2


def _orient(xyz, tet):
    p = xyz[tet]
    a = p[:, 1] - p[:, 0]
    b = p[:, 2] - p[:, 0]
    c = p[:, 3] - p[:, 0]
    return np.einsum("ij,ij->i", a, np.cross(b, c))


def kuhn_cube(n, jitter=0.0, seed=0, lo=0.0, hi=1.0):
    """Unit-cube Kuhn mesh: (n+1)^3 vertices, 6 n^3 positively oriented tets."""
    h = (hi - lo) / n
    g = np.arange(n + 1)
    I, J, K = np.meshgrid(g, g, g, indexing="ij")
    ijk = np.stack([I.ravel(), J.ravel(), K.ravel()], axis=1)
    xyz = lo + ijk.astype(np.float64) * h

    def vid(i, j, k):
        return (i * (n + 1) + j) * (n + 1) + k

    tets = []
    cells = np.stack(np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij"), -1).reshape(-1, 3)
    for perm in _PERMS:
        path = [np.zeros(3, dtype=np.int64)]
        for ax in perm:
            step = path[-1].copy()
            step[ax] += 1
            path.append(step)
        verts = [vid(cells[:, 0] + d[0], cells[:, 1] + d[1], cells[:, 2] + d[2]) for d in path]
        if _perm_parity(perm):
            verts[1], verts[2] = verts[2], verts[1]
        tets.append(np.stack(verts, axis=1))
    # order: cell-major, 6 tets of a cell adjacent (better locality)
    tet = np.stack(tets, axis=1).reshape(-1, 4).astype(np.int32)
    if jitter > 0:
        rng = np.random.default_rng(seed)
        interior = np.all((ijk > 0) & (ijk < n), axis=1)
        base = xyz.copy()
        for _ in range(100):
            d = rng.uniform(-jitter * h, jitter * h, size=(int(interior.sum()), 3))
            xyz = base.copy()
            xyz[interior] += d
            if np.all(_orient(xyz, tet) > 0):
                break
        else:
            raise RuntimeError("jitter produced inverted tetrahedra")
    assert np.all(_orient(xyz, tet) > 0)
    return xyz, tet


def graded_box(n, half=10.0, beta=2.0):
    """[-half, half]^3 Kuhn mesh graded toward the centre: the uniform grid
    s in [-1, 1] maps to x = half atanh(s tanh(beta)) / beta, whose spacing
    is smallest at 0 (by the factor tanh(beta)/beta against the uniform one)
    and grows toward the faces; the faces stay at exactly +-half."""
    s, tet = kuhn_cube(n, lo=-1.0, hi=1.0)
    x = half * np.arctanh(s * np.tanh(beta)) / beta
    x[np.abs(s) == 1.0] = np.sign(s[np.abs(s) == 1.0]) * half
    return x, tet


def cube_plane_marks(xyz, lo=0.0, hi=1.0, tol=1e-12):
    """Per-vertex bit mask of the cube's boundary planes a vertex lies on.

    bit 0: x=lo, 1: x=hi, 2: y=lo, 3: y=hi, 4: z=lo, 5: z=hi."""
    m = np.zeros(len(xyz), dtype=np.int64)
    for d in range(3):
        m |= (np.abs(xyz[:, d] - lo) < tol).astype(np.int64) << (2 * d)
        m |= (np.abs(xyz[:, d] - hi) < tol).astype(np.int64) << (2 * d + 1)
    return m


def geometry_marks(vertex_planes, edge_vtx, face_vtx, dirichlet_planes=0b000001, neumann_planes=0b111110):
    """Boundary marks per vertex, edge and face (AFEPack boundaryMark).

    A geometry lies on a plane when all its vertices do.  Mark 1 (Dirichlet)
    for geometries on a plane in ``dirichlet_planes`` (closure of those faces),
    else mark 2 (Neumann) for geometries on a plane in ``neumann_planes``,
    else 0."""
    def mark(bits):
        out = np.zeros(len(bits), dtype=np.int32)
        out[(bits & neumann_planes) != 0] = 2
        out[(bits & dirichlet_planes) != 0] = 1
        return out

    vb = vertex_planes
    eb = vb[edge_vtx[:, 0]] & vb[edge_vtx[:, 1]]
    fb = vb[face_vtx[:, 0]] & vb[face_vtx[:, 1]] & vb[face_vtx[:, 2]]
    return mark(vb), mark(eb), mark(fb)


def write_afepack_mesh(path, xyz, edge_vtx, face_vtx, face_edge, tet, tet_face, marks=None):
    """AFEPack ``.mesh`` text: points, then per dimension the geometry lists
    ``index n_vtx vtx... n_bnd bnd... mark`` (the format of
    examples/*/cube.mesh)."""
    mv, me, mf = marks if marks is not None else (None, None, None)
    nv = len(xyz)
    lines = [str(nv)]
    lines += ["%.17g\t%.17g\t%.17g" % tuple(p) for p in xyz]
    lines += ["", "", str(nv)]
    for i in range(nv):
        lines += [str(i), "1\t%d" % i, "1\t%d" % i, str(int(mv[i]) if mv is not None else 0)]
    lines += ["", str(len(edge_vtx))]
    for i, (a, b) in enumerate(edge_vtx):
        lines += [str(i), "2\t%d\t%d" % (a, b), "2\t%d\t%d" % (a, b), str(int(me[i]) if me is not None else 0)]
    lines += ["", str(len(face_vtx))]
    for i in range(len(face_vtx)):
        lines += [str(i), "3\t%d\t%d\t%d" % tuple(face_vtx[i]), "3\t%d\t%d\t%d" % tuple(face_edge[i]),
                  str(int(mf[i]) if mf is not None else 0)]
    lines += ["", str(len(tet))]
    for i in range(len(tet)):
        lines += [str(i), "4\t%d\t%d\t%d\t%d" % tuple(tet[i]), "4\t%d\t%d\t%d\t%d" % tuple(tet_face[i]), "0"]
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def read_afepack_mesh(path):
    """Parse an AFEPack ``.mesh`` file into (xyz, tet, marks, topology):
    marks per point, per file edge and per file face; topology = (edge_vtx,
    face_vtx, face_edge, tet_face) in file order, vertices as point ids."""
    toks = open(path).read().split()
    pos = 0

    def nxt():
        nonlocal pos
        t = toks[pos]
        pos += 1
        return t

    nv = int(nxt())
    xyz = np.array([[float(nxt()) for _ in range(3)] for _ in range(nv)])
    geos = []
    for dim in range(4):
        n = int(nxt())
        vtx, bnd, mk = [], [], []
        for _ in range(n):
            nxt()
            k = int(nxt())
            vtx.append([int(nxt()) for _ in range(k)])
            k = int(nxt())
            bnd.append([int(nxt()) for _ in range(k)])
            mk.append(int(nxt()))
        geos.append((vtx, bnd, np.array(mk, dtype=np.int32)))
    # dim-0 geometries name points; map through them
    pt = np.array([v[0] for v in geos[0][0]], dtype=np.int64)
    edge_vtx = pt[np.array(geos[1][0], dtype=np.int64)]
    face_vtx = pt[np.array(geos[2][0], dtype=np.int64)]
    face_edge = np.array(geos[2][1], dtype=np.int64)
    tet = pt[np.array(geos[3][0], dtype=np.int64)].astype(np.int32)
    tet_face = np.array(geos[3][1], dtype=np.int64)
    mark_pt = np.zeros(nv, dtype=np.int32)
    mark_pt[pt] = geos[0][2]
    marks = (mark_pt, geos[1][2], geos[2][2])
    return xyz, tet, marks, (edge_vtx, face_vtx, face_edge, tet_face)
