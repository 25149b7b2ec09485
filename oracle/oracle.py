"""TEST INFRASTRUCTURE ONLY: ctypes front end of the CPU oracle (liboracle.so).

Used by tests/, the smoke check and bench.py's cpu_baseline leg as the
checker; the product never imports it.  See tsem_oracle.cpp for the reference
files each routine restates.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
FACE_TABLE = os.path.join(os.path.dirname(HERE), "tsem_amd", "data", "face_table.txt")

_lib = None


def build(force=False):
    if force or not os.path.exists(LIB):
        subprocess.check_call(["make", "-s", "-C", HERE, "liboracle.so"])


def lib():
    global _lib
    if _lib is None:
        build()
        L = C.CDLL(LIB)
        vp, dp, ip, lp = C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int), C.POINTER(C.c_long)
        L.or_space_create.restype = vp
        L.or_space_create.argtypes = [C.c_int, C.c_int, dp, C.c_int, ip, C.c_char_p, ip, C.c_int]
        L.or_space_destroy.argtypes = [vp]
        L.or_info.argtypes = [vp, lp]
        L.or_topology.argtypes = [vp, ip, ip, ip, ip]
        L.or_maps.argtypes = [vp, ip, ip, dp]
        L.or_dof_start.argtypes = [vp, ip]
        L.or_orientation.argtypes = [vp, ip, ip, ip, ip, ip, ip]
        L.or_volumes.argtypes = [vp, dp]
        L.or_ref_matrices.argtypes = [vp, dp, dp]
        L.or_stiff_weights.argtypes = [vp, dp]
        L.or_basis_at.argtypes = [vp, C.c_int, dp, dp, dp]
        L.or_multiindex.argtypes = [vp, ip]
        L.or_apply.argtypes = [vp, C.c_double, C.c_double, dp, dp]
        L.or_gather.argtypes = [vp, dp, dp]
        L.or_assemble.restype = C.c_long
        L.or_assemble.argtypes = [vp, C.c_double, C.c_double]
        L.or_csr.argtypes = [vp, lp, ip, dp]
        L.or_csr_vmult.argtypes = [vp, dp, dp]
        L.or_set_marks.argtypes = [vp, ip, ip, ip]
        L.or_bc_quad.argtypes = [vp, ip, dp, dp, dp, dp]
        L.or_quad3.argtypes = [vp, dp, dp]
        L.or_rhs.argtypes = [vp, dp, dp]
        L.or_qp_coords.argtypes = [vp, dp]
        L.or_neumann.argtypes = [vp, dp, C.c_int, dp]
        L.or_dirichlet.argtypes = [vp, dp, dp, dp, C.c_int, C.c_int, dp, dp]
        L.or_zero_bc.argtypes = [vp, C.c_int, C.c_int, dp]
        L.or_sum_dof.argtypes = [vp, ip]
        L.or_pcg.restype = C.c_int
        L.or_pcg.argtypes = [vp, dp, dp, C.c_double, C.c_int, C.c_int, C.c_int, C.c_int, ip, C.c_int, dp]
        L.or_pcg_mt.restype = C.c_int
        L.or_pcg_mt.argtypes = [vp, dp, dp, C.c_double, C.c_int, C.c_int, C.c_int, C.c_int, ip, C.c_int, C.c_int, dp]
        _lib = L
    return _lib


def _d(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _i(a):
    return a.ctypes.data_as(C.POINTER(C.c_int))


def _l(a):
    return a.ctypes.data_as(C.POINTER(C.c_long))


class OracleSpace:
    """Oracle TSEM space on a mesh: maps, reference matrices, CSR, BCs, PCG."""

    def __init__(self, M, xyz, tet, femdof_weight=None, qorder=None):
        L = lib()
        self.M = M
        self.xyz = np.ascontiguousarray(xyz, dtype=np.float64)
        self.tet = np.ascontiguousarray(tet, dtype=np.int32)
        fw = None if femdof_weight is None else np.ascontiguousarray(femdof_weight, dtype=np.int32)
        self.qorder = qorder if qorder is not None else M + 3
        self.h = L.or_space_create(M, len(self.xyz), _d(self.xyz), len(self.tet), _i(self.tet),
                                   FACE_TABLE.encode(), _i(fw) if fw is not None else None, self.qorder)
        if not self.h:
            raise RuntimeError("oracle space creation failed")
        info = np.zeros(11, dtype=np.int64)
        L.or_info(self.h, _l(info))
        self.ndof, self.K, self.nv, self.ne, self.nf, self.nt = (int(v) for v in info[:6])
        self.ndg = [int(v) for v in info[6:10]]
        self.nnz_map = int(info[10])

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_space_destroy(self.h)
            self.h = None

    # ---- topology / maps ----
    def topology(self):
        ev = np.zeros((self.ne, 2), np.int32)
        fv = np.zeros((self.nf, 3), np.int32)
        fe = np.zeros((self.nf, 3), np.int32)
        tf = np.zeros((self.nt, 4), np.int32)
        lib().or_topology(self.h, _i(ev), _i(fv), _i(fe), _i(tf))
        return ev, fv, fe, tf

    def maps(self):
        off = np.zeros(self.nt * self.K + 1, np.int32)
        ind = np.zeros(self.nnz_map, np.int32)
        val = np.zeros(self.nnz_map, np.float64)
        lib().or_maps(self.h, _i(off), _i(ind), _d(val))
        return off, ind, val

    def dof_start(self):
        out = np.zeros(self.nv + self.ne + self.nf + self.nt, np.int32)
        lib().or_dof_start(self.h, _i(out))
        return out

    def orientation(self):
        te = np.zeros((self.nt, 6), np.int32)
        tfi = np.zeros((self.nt, 4), np.int32)
        so = np.zeros((self.nf, 3), np.int32)
        nnf = np.zeros((self.nf, 3), np.int32)
        nef = np.zeros((self.nf, 3), np.int32)
        nne = np.zeros((self.ne, 2), np.int32)
        lib().or_orientation(self.h, _i(te), _i(tfi), _i(so), _i(nnf), _i(nef), _i(nne))
        return dict(tp_edge=te, tp_face_inv=tfi, sameorder=so, num_node_f=nnf, num_edge_f=nef, num_node_e=nne)

    def volumes(self):
        v = np.zeros(self.nt)
        lib().or_volumes(self.h, _d(v))
        return v

    def ref_matrices(self):
        K = self.K
        Mr = np.zeros((K, K))
        S = np.zeros((6, K, K))
        lib().or_ref_matrices(self.h, _d(Mr), _d(S))
        return Mr, S

    def stiff_weights(self):
        w = np.zeros((self.nt, 6))
        lib().or_stiff_weights(self.h, _d(w))
        return w

    def basis_at(self, pts):
        pts = np.ascontiguousarray(pts, dtype=np.float64)
        n = len(pts)
        v = np.zeros((n, self.K))
        g = np.zeros((n, self.K, 3))
        lib().or_basis_at(self.h, n, _d(pts), _d(v), _d(g))
        return v, g

    def multiindex(self):
        out = np.zeros((self.K, 3), np.int32)
        lib().or_multiindex(self.h, _i(out))
        return out

    def sum_dof(self):
        out = np.zeros(self.ndof, np.int32)
        lib().or_sum_dof(self.h, _i(out))
        return out

    # ---- operator ----
    def apply(self, x, stiff=1.0, mass=0.0):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros(self.ndof)
        lib().or_apply(self.h, stiff, mass, _d(x), _d(y))
        return y

    def gather(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        xl = np.zeros(self.nt * self.K)
        lib().or_gather(self.h, _d(x), _d(xl))
        return xl.reshape(self.nt, self.K)

    def assemble(self, stiff=1.0, mass=0.0):
        self.nnz = int(lib().or_assemble(self.h, stiff, mass))
        return self.nnz

    def csr(self):
        rs = np.zeros(self.ndof + 1, np.int64)
        col = np.zeros(self.nnz, np.int32)
        v = np.zeros(self.nnz)
        lib().or_csr(self.h, _l(rs), _i(col), _d(v))
        return rs, col, v

    def vmult(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros(self.ndof)
        lib().or_csr_vmult(self.h, _d(x), _d(y))
        return y

    # ---- boundary / rhs ----
    def set_marks(self, mv, me, mf):
        self._marks = [np.ascontiguousarray(m, dtype=np.int32) for m in (mv, me, mf)]
        lib().or_set_marks(self.h, *(_i(m) for m in self._marks))

    def bc_quad(self):
        nq = np.zeros(3, np.int32)
        lib().or_bc_quad(self.h, _i(nq), None, None, None, None)
        t1 = np.zeros(nq[0]); w1 = np.zeros(nq[0]); p2 = np.zeros((nq[1], 3)); w2 = np.zeros(nq[1])
        lib().or_bc_quad(self.h, _i(nq), _d(t1), _d(w1), _d(p2), _d(w2))
        return t1, w1, p2, w2

    def quad3(self):
        nq = np.zeros(3, np.int32)
        lib().or_bc_quad(self.h, _i(nq), None, None, None, None)
        pts = np.zeros((nq[2], 3)); w = np.zeros(nq[2])
        lib().or_quad3(self.h, _d(pts), _d(w))
        return pts, w

    def qp_coords(self):
        nq = len(self.quad3()[1])
        out = np.zeros((self.nt, nq, 3))
        lib().or_qp_coords(self.h, _d(out))
        return out

    def rhs(self, fq):
        fq = np.ascontiguousarray(fq, dtype=np.float64)
        b = np.zeros(self.ndof)
        lib().or_rhs(self.h, _d(fq), _d(b))
        return b

    def neumann(self, gq, flag, b):
        gq = np.ascontiguousarray(gq, dtype=np.float64)
        lib().or_neumann(self.h, _d(gq), flag, _d(b))

    def dirichlet(self, uv, ue, uf, flag, b, modify=True):
        uv, ue, uf = (np.ascontiguousarray(a, dtype=np.float64) for a in (uv, ue, uf))
        bval = np.zeros(self.ndof)
        lib().or_dirichlet(self.h, _d(uv), _d(ue), _d(uf), flag, 1 if modify else 0, _d(b), _d(bval))
        return bval

    def zero_bc(self, flag, on_matrix=True, vec=None):
        lib().or_zero_bc(self.h, flag, 1 if on_matrix else 0, _d(vec) if vec is not None else None)

    def pcg(self, b, x0=None, tol=1e-10, max_step=100000, precond=1, ite_step=1, smooth=2, criterion=1, threads=1):
        """solve_PCG on the assembled CSR; threads > 1 runs the products and
        vector sweeps on that many OpenMP threads (Gauss-Seidel stays
        sequential).  Stats carry the p-MG setup and solve seconds."""
        x = np.zeros(self.ndof) if x0 is None else np.array(x0, dtype=np.float64)
        b = np.ascontiguousarray(b, dtype=np.float64)
        sd = self.sum_dof() if precond == 4 else np.zeros(1, np.int32)
        st = np.zeros(5)
        rc = lib().or_pcg_mt(self.h, _d(x), _d(b), tol, max_step, precond, ite_step, smooth, _i(sd), criterion,
                             int(threads), _d(st))
        return x, dict(converged=(rc == 0), iterations=int(st[0]), l1=st[1], relres=st[2], setup_s=st[3],
                       solve_s=st[4])
