"""Host side of the product library against the oracle (CPU only, no GPU).

The product builds its index maps, DOF layout and reference matrices with its
own code (hashing topology, exact integer sort keys, long-double reference
element); a host-only space (no device context) is compared with the oracle
here: integer maps bit-exact, reference data to round-off.
"""
import os
import re

import numpy as np
import pytest

import oracle as O
from tsem_amd import api, lib, mesh

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol():
    hdr = open(os.path.join(ROOT, "include", "tsem.h")).read()
    names = sorted(set(re.findall(r"\b(tsem_[a-z0-9_]+)\s*\(", hdr)))
    L = lib.load()
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert set(lib.SIGNATURES) <= set(names)


CASES = [(1, 2, 0.0, 0), (2, 3, 0.1, 1), (2, 4, 0.1, 2), (3, 6, 0.15, 3), (2, 8, 0.1, 4), (2, 10, 0.05, 5)]


@pytest.mark.parametrize("n,M,jit,seed", CASES)
def test_host_space_matches_oracle(n, M, jit, seed):
    xyz, tet = mesh.kuhn_cube(n, jitter=jit, seed=seed)
    o = O.OracleSpace(M, xyz, tet)
    p = api.Space(None, M, xyz, tet, qorder=M + 3)
    assert (o.ndof, o.K, o.nv, o.ne, o.nf, o.nt) == (p.ndof, p.K, p.nv, p.ne, p.nf, p.nt)
    for a, b in zip(o.topology(), p.topology()):
        assert np.array_equal(a, b)
    for a, b in zip(o.maps(), p.maps()):
        assert np.array_equal(a, b)
    assert np.array_equal(o.dof_start(), p.dof_start())
    assert np.array_equal(o.sum_dof(), p.sum_dof())
    Mo, So = o.ref_matrices()
    Mp, Sp, wp = p.ref_matrices()
    assert np.abs(Mo - Mp).max() <= 1e-14 * np.abs(Mo).max()
    assert np.abs(So - Sp).max() <= 1e-13 * np.abs(So).max()
    assert np.array_equal(o.stiff_weights(), wp)
    o.assemble(1.0, 0.0)
    rs, col, v = o.csr()
    assert np.abs(v[rs[:-1]] - p.diagonal_host()).max() <= 1e-14 * np.abs(v[rs[:-1]]).max()


@pytest.mark.parametrize("n,M,jit,seed", CASES[1:])
def test_slab_scatter_matches_gather_map(n, M, jit, seed):
    """The entity-slab tables (per-element offsets, orientation words,
    transposed face rows, per-DOF counts) scatter element results like the
    transpose of the oracle's gather map; vertex order shuffled per element
    so that every face variant and edge flip occurs."""
    xyz, tet = mesh.kuhn_cube(n, jitter=jit, seed=seed)
    rng = np.random.default_rng(seed)
    tet = np.array([t[rng.permutation(4)] for t in tet], dtype=tet.dtype)
    o = O.OracleSpace(M, xyz, tet)
    p = api.Space(None, M, xyz, tet, qorder=M + 3)
    off, ind, val = o.maps()
    yl = rng.standard_normal((p.nt, p.K))
    flat = yl.ravel()
    y = np.zeros(p.ndof)
    rows = np.repeat(np.arange(len(flat)), np.diff(off))
    np.add.at(y, ind, val * flat[rows])
    ys = p.slab_apply_host(yl)
    assert np.abs(ys - y).max() <= 1e-13 * np.abs(y).max()


def test_custom_femdof_order():
    """A caller-supplied P2 DOF order (as AFEPack would give) drives the
    layout identically in oracle and product."""
    xyz, tet = mesh.kuhn_cube(2, jitter=0.1, seed=3)
    o0 = O.OracleSpace(4, xyz, tet)
    nw = o0.nv + o0.ne
    w = np.random.default_rng(7).permutation(nw).astype(np.int32)
    o = O.OracleSpace(4, xyz, tet, femdof_weight=w)
    p = api.Space(None, 4, xyz, tet, femdof_weight=w, qorder=7)
    assert np.array_equal(o.dof_start(), p.dof_start())
    for a, b in zip(o.maps(), p.maps()):
        assert np.array_equal(a, b)
    assert not np.array_equal(o.dof_start(), o0.dof_start())


def test_single_tetrahedron_and_permuted_vertices():
    xyz = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1.0]])
    for perm in ([0, 1, 2, 3], [2, 0, 3, 1], [3, 2, 1, 0]):
        tet = np.array([perm], np.int32)
        o = O.OracleSpace(5, xyz, tet)
        p = api.Space(None, 5, xyz, tet, qorder=8)
        assert p.ndof == o.ndof == 56
        for a, b in zip(o.maps(), p.maps()):
            assert np.array_equal(a, b)


def test_invalid_inputs_raise():
    xyz, tet = mesh.kuhn_cube(1)
    with pytest.raises(lib.TsemError) as e:
        api.Space(None, 0, xyz, tet)
    assert e.value.code == lib.TSEM_EINVAL
    with pytest.raises(lib.TsemError):
        api.Space(None, 13, xyz, tet)
    bad = tet.copy()
    bad[0, 0] = len(xyz) + 5
    with pytest.raises(lib.TsemError) as e:
        api.Space(None, 3, xyz, bad)
    assert e.value.code == lib.TSEM_EINVAL
    flat = xyz.copy()
    flat[:, 2] = 0.0  # every tet degenerate
    with pytest.raises(lib.TsemError) as e:
        api.Space(None, 3, flat, tet)
    assert e.value.code == lib.TSEM_EINVAL


def test_host_only_space_refuses_device_work():
    xyz, tet = mesh.kuhn_cube(1)
    p = api.Space(None, 3, xyz, tet, qorder=5)
    with pytest.raises(lib.TsemError) as e:
        import ctypes as C
        lib.check(lib.load().tsem_rhs_begin(p.h))
    assert e.value.code == lib.TSEM_ESTATE


def test_afepack_mesh_roundtrip(tmp_path):
    xyz, tet = mesh.kuhn_cube(2, jitter=0.1, seed=1)
    p = api.Space(None, 3, xyz, tet, qorder=5)
    ev, fv, fe, tf = p.topology()
    marks = mesh.geometry_marks(mesh.cube_plane_marks(xyz), ev, fv)
    path = tmp_path / "cube.mesh"
    mesh.write_afepack_mesh(str(path), xyz, ev, fv, fe, tet, tf, marks)
    xyz2, tet2, marks2, topo2 = mesh.read_afepack_mesh(str(path))
    assert np.array_equal(xyz, xyz2) and np.array_equal(tet, tet2)
    for a, b in zip(marks, marks2):
        assert np.array_equal(a, b)
    for a, b in zip((ev, fv, fe, tf), topo2):
        assert np.array_equal(a, b)


_READER = r'''
#include <cstdio>
#include "tsem.hpp"
int main(int, char** argv) {
  tsem::RegularMesh m;
  m.readMesh(argv[1]);
  std::printf("%zu %zu %zu %zu\n", m.xyz.size() / 3, m.edge_vtx.size(), m.face_vtx.size(), m.tet.size() / 4);
  for (double v : m.xyz) std::printf("%.17g ", v);
  std::printf("\n");
  for (int v : m.tet) std::printf("%d ", v);
  std::printf("\n");
  for (auto& e : m.edge_vtx) std::printf("%d %d ", e[0], e[1]);
  std::printf("\n");
  for (auto& f : m.face_vtx) std::printf("%d %d %d ", f[0], f[1], f[2]);
  std::printf("\n");
  for (auto* mk : {&m.mark_v, &m.mark_e, &m.mark_f}) {
    for (int v : *mk) std::printf("%d ", v);
    std::printf("\n");
  }
  return 0;
}
'''


@pytest.mark.parametrize("name", ["cube_mixbc.mesh", "cube_dirichlet.mesh", "written"])
def test_cxx_mesh_reader_matches_python(tmp_path, name):
    """RegularMesh::readMesh of include/tsem.hpp against mesh.read_afepack_mesh
    on the reference examples' cube meshes and on a written Kuhn mesh; the
    file marks land on the same space geometries through both mappings."""
    import subprocess
    if name == "written":
        xyz, tet = mesh.kuhn_cube(2, jitter=0.1, seed=4)
        p = api.Space(None, 3, xyz, tet, qorder=5)
        ev, fv, fe, tf = p.topology()
        marks = mesh.geometry_marks(mesh.cube_plane_marks(xyz), ev, fv, 0b101111, 0b010000)
        path = str(tmp_path / "k.mesh")
        # reverse the file's edge and face order so file and space numbering differ
        mesh.write_afepack_mesh(path, xyz, ev[::-1], fv[::-1], fe[::-1], tet, tf, (marks[0], marks[1][::-1],
                                                                                      marks[2][::-1]))
    else:
        path = os.path.join(ROOT, "tests", "golden", name)
    src = tmp_path / "reader.cpp"
    src.write_text(_READER)
    exe = str(tmp_path / "reader")
    subprocess.check_call(["g++", "-std=c++17", "-O1", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
                           "-I" + os.path.join(ROOT, "include"), str(src), "-o", exe])
    out = subprocess.check_output([exe, path], text=True).split("\n")
    nv, ne, nf, nt = (int(v) for v in out[0].split())
    xyz, tet, marks, (ev, fv, _, _) = mesh.read_afepack_mesh(path)
    assert (nv, ne, nf, nt) == (len(xyz), len(ev), len(fv), len(tet))
    assert np.array_equal(np.array(out[1].split(), dtype=np.float64).reshape(-1, 3), xyz)
    assert np.array_equal(np.array(out[2].split(), dtype=np.int64).reshape(-1, 4), tet)
    assert np.array_equal(np.array(out[3].split(), dtype=np.int64).reshape(-1, 2), ev)
    assert np.array_equal(np.array(out[4].split(), dtype=np.int64).reshape(-1, 3), fv)
    for k in range(3):
        assert np.array_equal(np.array(out[5 + k].split(), dtype=np.int64), marks[k])
    # the mapping of file marks onto the space's own edges and faces
    sp = api.Space(None, 3, xyz, tet, qorder=5)
    mv, me, mf = sp.set_file_marks(marks, ev, fv)
    sev, sfv, _, _ = sp.topology()
    emap = {tuple(sorted(e)): m for e, m in zip(ev.tolist(), marks[1])}
    fmap = {tuple(sorted(f)): m for f, m in zip(fv.tolist(), marks[2])}
    assert all(me[i] == emap[tuple(sorted(e))] for i, e in enumerate(sev.tolist()))
    assert all(mf[i] == fmap[tuple(sorted(f))] for i, f in enumerate(sfv.tolist()))
    if name == "cube_mixbc.mesh":
        assert sorted(np.bincount(mf).tolist()) == [2, 6, 10]


def test_mesh_generators():
    xyz, tet = mesh.kuhn_cube(4, jitter=0.1, seed=2026)
    assert len(tet) == 6 * 64 and len(xyz) == 125
    p = xyz[tet]
    det = np.einsum("ij,ij->i", p[:, 1] - p[:, 0], np.cross(p[:, 2] - p[:, 0], p[:, 3] - p[:, 0]))
    assert det.min() > 0
    assert abs(det.sum() / 6 - 1.0) < 1e-12  # tets tile the cube
    g, t = mesh.graded_box(6)
    assert abs(g.min() + 10) < 1e-12 and abs(g.max() - 10) < 1e-12


def test_coef_files_round_trip_and_change_order(tmp_path):
    """write_coef/read_coef (src/Miscellaneous.cpp:34-125): the same order
    round-trips to 14 digits; read into a higher order, the hierarchical
    basis gives the same function (checked at points through the oracle);
    read into a lower order, the leading modes are kept."""
    xyz, tet = mesh.kuhn_cube(2, jitter=0.1, seed=5)
    lo = api.Space(None, 4, xyz, tet, qorder=5)
    hi = api.Space(None, 6, xyz, tet, qorder=5)
    x = np.random.default_rng(2).standard_normal(lo.ndof)
    path = str(tmp_path / "x.coef")
    lo.write_coef(x, path)
    head = open(path).readline().split()
    assert head == ["4", str(lo.ndof)]
    assert np.abs(lo.read_coef(path) - x).max() <= 1e-13 * np.abs(x).max()
    xh = hi.read_coef(path)
    o_lo, o_hi = O.OracleSpace(4, xyz, tet), O.OracleSpace(6, xyz, tet)
    pts = np.random.default_rng(3).uniform(0, 0.25, size=(20, 3))
    v_lo, _ = o_lo.basis_at(pts)
    v_hi, _ = o_hi.basis_at(pts)
    u_lo = o_lo.gather(x) @ v_lo.T
    u_hi = o_hi.gather(xh) @ v_hi.T
    assert np.abs(u_lo - u_hi).max() <= 1e-12 * np.abs(u_lo).max()
    # back down: the order-6 vector read into order 4 gives x again
    path2 = str(tmp_path / "xh.coef")
    hi.write_coef(xh, path2)
    assert np.abs(lo.read_coef(path2) - x).max() <= 1e-13 * np.abs(x).max()
    # several vectors in one file
    lo.write_coef([x, 2 * x], path)
    a, b = lo.read_coef(path, count=2)
    assert np.abs(b - 2 * a).max() <= 1e-12 * np.abs(b).max()
