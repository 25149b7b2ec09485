"""Python host side of the TSEM library: spaces, operators and solvers.

Device vectors are torch tensors on ``cuda:<device>`` (HIP on ROCm); the work
itself runs in libtsem_amd.so through the C ABI.  ``Space(None, ...)`` builds
a host-only space (index maps, reference matrices, diagonal) without a GPU.
"""
import ctypes as C

import numpy as np

from . import lib as L
from .mesh import cube_plane_marks, geometry_marks


class Context:
    """Device + stream (tsem_ctx_create)."""

    def __init__(self, device=0):
        self.device = device
        h = C.c_void_p()
        L.check(L.load().tsem_ctx_create(device, C.byref(h)))
        self.h = h

    def set_class(self, cls):
        """Solve class: 0 normal, 1 foreground, 2 background (a background
        PCG pauses between chunks while a foreground PCG runs)."""
        L.check(L.load().tsem_ctx_set_class(self.h, int(cls)))

    def foreground_begin(self):
        L.check(L.load().tsem_foreground_begin(self.h))

    def foreground_end(self):
        L.check(L.load().tsem_foreground_end(self.h))

    def background_wait(self, max_seconds=0.0):
        """Returns when no foreground span or solve is active, or after
        max_seconds (<= 0: no limit); at once unless this context is
        background.  The GIL is released while waiting."""
        L.check(L.load().tsem_background_wait(self.h, float(max_seconds)))

    def close(self):
        if getattr(self, "h", None):
            try:
                L.load().tsem_ctx_destroy(self.h)
            except Exception:  # interpreter shutdown: the process releases the device
                pass
            self.h = None

    def __del__(self):
        self.close()


def _ptr(t):
    return C.c_void_p(t.data_ptr())


def _ready():
    """Finish pending torch work before the library (own stream) reads a
    tensor, or writes one that torch's allocator has just handed out: a block
    freed by torch may still be written by work queued on torch's stream (by
    any thread), which is not ordered before the library's stream."""
    import torch
    torch.cuda.current_stream().synchronize()


class Space:
    """TSEM space of order M on a tetrahedral mesh (TSEM<double> + init)."""

    def __init__(self, ctx, M, xyz, tet, femdof_weight=None, qorder=30):
        lib = L.load()
        self.ctx = ctx
        self.M = M
        self.xyz = np.ascontiguousarray(xyz, dtype=np.float64)
        self.tet = np.ascontiguousarray(tet, dtype=np.int32)
        fw = None if femdof_weight is None else np.ascontiguousarray(femdof_weight, dtype=np.int32)
        h = C.c_void_p()
        L.check(lib.tsem_space_create(ctx.h if ctx is not None else None, M, len(self.xyz), L.dptr(self.xyz),
                                      len(self.tet), L.iptr(self.tet), L.iptr(fw) if fw is not None else None,
                                      qorder, C.byref(h)))
        self.h = h
        info = np.zeros(12, np.int64)
        L.check(lib.tsem_space_info(h, L.lptr(info)))
        self.ndof, self.K, self.nv, self.ne, self.nf, self.nt = (int(v) for v in info[:6])
        self.ndg = [int(v) for v in info[6:10]]
        self.nmap = int(info[10])
        self.nq3 = int(info[11])
        self.qorder = qorder
        self.marks = None

    def set_geometry(self, xyz):
        """New vertex coordinates for the same connectivity (tsem_space_set_geometry);
        boundary eliminations and the potential are cleared."""
        self.xyz = np.ascontiguousarray(xyz, dtype=np.float64)
        self._bc_cache = {}
        _ready()
        L.check(L.load().tsem_space_set_geometry(self.h, L.dptr(self.xyz)))

    def close(self):
        if getattr(self, "h", None):
            try:
                L.load().tsem_space_destroy(self.h)
            except Exception:  # interpreter shutdown
                pass
            self.h = None

    def __del__(self):
        self.close()

    @property
    def device(self):
        import torch
        return torch.device("cuda", self.ctx.device)

    # ---------------------------------------------------------------- host --
    def topology(self):
        """(edge_vtx, face_vtx, face_edge, tet_face); fixed for the space, read once."""
        if getattr(self, "_topo", None) is None:
            ev = np.zeros((self.ne, 2), np.int32)
            fv = np.zeros((self.nf, 3), np.int32)
            fe = np.zeros((self.nf, 3), np.int32)
            tf = np.zeros((self.nt, 4), np.int32)
            L.check(L.load().tsem_space_topology(self.h, L.iptr(ev), L.iptr(fv), L.iptr(fe), L.iptr(tf)))
            self._topo = (ev, fv, fe, tf)
        return tuple(a.copy() for a in self._topo)

    def maps(self):
        off = np.zeros(self.nt * self.K + 1, np.int32)
        ind = np.zeros(self.nmap, np.int32)
        val = np.zeros(self.nmap)
        L.check(L.load().tsem_space_maps(self.h, L.iptr(off), L.iptr(ind), L.dptr(val)))
        return off, ind, val

    def point_values(self, x, pts):
        """u_h at arbitrary points (host n x 3): (values, holding element or -1
        and NaN outside the mesh); TSEM::calc_val_point."""
        pts = np.ascontiguousarray(pts, dtype=np.float64).reshape(-1, 3)
        out = np.zeros(len(pts))
        elem = np.zeros(len(pts), np.int32)
        _ready()
        L.check(L.load().tsem_point_values(self.h, _ptr(x), len(pts), L.dptr(pts), L.dptr(out), L.iptr(elem)))
        return out, elem

    def slab_state(self):
        """(active, slab entries, largest chunk span, face row terms): active
        when the FP32 operator writes entity slabs (tsem_space_slab)."""
        out = np.zeros(4, np.int64)
        L.check(L.load().tsem_space_slab(self.h, L.lptr(out)))
        return bool(out[0]), int(out[1]), int(out[2]), int(out[3])

    def slab_apply_host(self, yl):
        """Scatter of element-major results yl (nt x K) through the entity-slab
        tables, on the host (a check of the layout the device kernels use)."""
        yl = np.ascontiguousarray(yl, dtype=np.float64).reshape(self.nt, self.K)
        y = np.zeros(self.ndof)
        L.check(L.load().tsem_space_slab_apply_host(self.h, L.dptr(yl), L.dptr(y)))
        return y

    def dof_start(self):
        out = np.zeros(self.nv + self.ne + self.nf + self.nt, np.int64)
        L.check(L.load().tsem_space_dof_start(self.h, L.lptr(out)))
        return out

    def sum_dof(self):
        out = np.zeros(self.ndof, np.int32)
        L.check(L.load().tsem_space_sum_dof(self.h, L.iptr(out)))
        return out

    def ref_matrices(self):
        K = self.K
        Mr = np.zeros((K, K)); S = np.zeros((6, K, K)); w = np.zeros((self.nt, 6))
        L.check(L.load().tsem_space_ref_matrices(self.h, L.dptr(Mr), L.dptr(S), L.dptr(w)))
        return Mr, S, w

    def factors(self):
        """(dfac[3][Kd][K], mfac[K][K], grad2[3][3][K][K]) of the element kernels."""
        K = self.K
        Kd = self.M * (self.M + 1) * (self.M + 2) // 6
        d = np.zeros((3, Kd, K)); m = np.zeros((K, K)); g = np.zeros((3, 3, K, K))
        L.check(L.load().tsem_space_factors(self.h, L.dptr(d), L.dptr(m), L.dptr(g)))
        return d, m, g

    def diagonal_host(self):
        d = np.zeros(self.ndof)
        L.check(L.load().tsem_space_diagonal_host(self.h, L.dptr(d)))
        return d

    def set_marks(self, mv, me, mf):
        """build_flag_bm(mesh): boundary mark per vertex, edge, face."""
        new = [np.ascontiguousarray(m, dtype=np.int32) for m in (mv, me, mf)]
        if self.marks is not None and all(np.array_equal(a, b) for a, b in zip(self.marks, new)):
            return  # unchanged: the library state and the boundary point cache stay
        self.marks = new
        self._bc_cache = {}
        L.check(L.load().tsem_set_marks(self.h, *(L.iptr(m) for m in self.marks)))

    def set_cube_marks(self, dirichlet_planes=0b000001, neumann_planes=0b111110, lo=0.0, hi=1.0):
        """Marks of an axis-aligned box [lo, hi]^3: 1 on the Dirichlet planes,
        2 on the Neumann planes (bits x=lo, x=hi, y=lo, y=hi, z=lo, z=hi)."""
        ev, fv, _, _ = self.topology()
        mv, me, mf = geometry_marks(cube_plane_marks(self.xyz, lo, hi), ev, fv, dirichlet_planes, neumann_planes)
        self.set_marks(mv, me, mf)
        return mv, me, mf

    def set_file_marks(self, marks, edge_vtx, face_vtx):
        """build_flag_bm from a mesh file's marks: per point, and per edge and
        face of the file (edge_vtx / face_vtx point ids, read_afepack_mesh)
        mapped onto this space's edges and faces by their vertex sets."""
        mv, me_file, mf_file = marks
        ev, fv, _, _ = self.topology()
        emap = {tuple(sorted(int(v) for v in e)): int(m) for e, m in zip(edge_vtx, me_file)}
        fmap = {tuple(sorted(int(v) for v in f)): int(m) for f, m in zip(face_vtx, mf_file)}
        me = np.array([emap.get(tuple(sorted(int(v) for v in e)), 0) for e in ev], dtype=np.int32)
        mf = np.array([fmap.get(tuple(sorted(int(v) for v in f)), 0) for f in fv], dtype=np.int32)
        self.set_marks(np.asarray(mv, dtype=np.int32), me, mf)
        return self.marks

    def bc_count(self, flag):
        n = np.zeros(3, np.int64)
        L.check(L.load().tsem_bc_count(self.h, flag, L.lptr(n)))
        return [int(v) for v in n]

    def bc_points(self, flag):
        """Sample points of boundary data on the geometries with mark flag
        (kept until the geometry or the marks change)."""
        cache = getattr(self, "_bc_cache", None)
        if cache is None:
            cache = self._bc_cache = {}
        if flag not in cache:
            q = self.qorder
            nv, ne, nf = self.bc_count(flag)
            pv = np.zeros((nv, 3)); pe = np.zeros((ne, q, 3)); pf = np.zeros((nf, q * q, 3))
            L.check(L.load().tsem_bc_points(self.h, flag, L.dptr(pv), L.dptr(pe), L.dptr(pf)))
            cache[flag] = (pv, pe, pf)
        return tuple(a.copy() for a in cache[flag])

    def marked_faces(self, flag):
        return np.nonzero(self.marks[2] == flag)[0]

    # -------------------------------------------------------------- device --
    def zeros(self, dtype=None):
        import torch
        return torch.zeros(self.ndof, dtype=dtype or torch.float64, device=self.device)

    def rhs(self, f, chunk=None, pause=None):
        """calc_rhs: load vector of a source term f(points) evaluated on device
        at the 3-D quadrature points, chunked over elements (chunk elements
        at a time; pause(), if given, runs between chunks)."""
        import torch
        lib = L.load()
        b = self.zeros()
        L.check(lib.tsem_rhs_begin(self.h))
        if chunk is None:
            # up to 6e8 point coordinates (4.8 GB) per piece: few pieces, each
            # wide enough for the whole device
            chunk = max(1, min(self.nt, int(6e8 // max(1, self.nq3 * 3))))
        for e0 in range(0, self.nt, chunk):
            if pause is not None and e0 > 0:
                pause()
            e1 = min(self.nt, e0 + chunk)
            pts = torch.empty((e1 - e0, self.nq3, 3), dtype=torch.float64, device=self.device)
            _ready()
            L.check(lib.tsem_qp_coords(self.h, e0, e1, _ptr(pts)))
            fq = f(pts).contiguous()
            _ready()
            L.check(lib.tsem_rhs_add(self.h, e0, e1, _ptr(fq)))
            del pts, fq
        _ready()
        L.check(lib.tsem_rhs_end(self.h, _ptr(b)))
        return b

    def rhs_qp(self, fq):
        """calc_rhs from values at the 3-D quadrature points of all elements
        (device, nt x n_qp3)."""
        b = self.zeros()
        lib = L.load()
        fq = fq.contiguous()
        _ready()
        L.check(lib.tsem_rhs_begin(self.h))
        L.check(lib.tsem_rhs_add(self.h, 0, self.nt, _ptr(fq)))
        L.check(lib.tsem_rhs_end(self.h, _ptr(b)))
        return b

    def qp_coords(self, e0=0, e1=None):
        """Physical coordinates of the 3-D quadrature points, (e1-e0, n_qp3, 3)."""
        import torch
        e1 = self.nt if e1 is None else e1
        pts = torch.empty((e1 - e0, self.nq3, 3), dtype=torch.float64, device=self.device)
        _ready()
        L.check(L.load().tsem_qp_coords(self.h, e0, e1, _ptr(pts)))
        return pts

    def qp_values(self, x, e0=0, e1=None):
        """calc_val_qp_onElement: u_h at the quadrature points, (e1-e0, n_qp3)."""
        import torch
        e1 = self.nt if e1 is None else e1
        out = torch.empty((e1 - e0, self.nq3), dtype=torch.float64, device=self.device)
        _ready()
        L.check(L.load().tsem_qp_values(self.h, _ptr(x), e0, e1, _ptr(out)))
        return out

    def density(self, psi, occ):
        """calc_density: sum_k occ_k psi_k^2 at the quadrature points; psi (k, n_dof)."""
        import torch
        psi = psi.contiguous()
        occ = np.ascontiguousarray(occ, dtype=np.float64)
        rho = torch.empty((self.nt, self.nq3), dtype=torch.float64, device=self.device)
        _ready()
        L.check(L.load().tsem_density(self.h, len(occ), _ptr(psi), L.dptr(occ), _ptr(rho)))
        return rho

    def integrate(self, f, g=None, moments=False):
        """sum_e vol_e sum_q w_q f g (and its moments 1, x, y, z, xx, xy, xz,
        yy, yz, zz when moments=True) over all elements."""
        out = np.zeros(10)
        f = f.contiguous()
        g = g.contiguous() if g is not None else None
        _ready()
        L.check(L.load().tsem_integrate(self.h, _ptr(f), _ptr(g) if g is not None else None, 10 if moments else 1,
                                        L.dptr(out)))
        return out if moments else float(out[0])

    def l2_error(self, x, u, chunk=None):
        """calc_l2_error: ||u_h - u|| with u a function of the points (device)."""
        import math
        if chunk is None:
            # up to 6e8 point coordinates (4.8 GB) per piece: few pieces, each
            # wide enough for the whole device
            chunk = max(1, min(self.nt, int(6e8 // max(1, self.nq3 * 3))))
        tot = 0.0
        s = C.c_double()
        for e0 in range(0, self.nt, chunk):
            e1 = min(self.nt, e0 + chunk)
            uq = u(self.qp_coords(e0, e1)).contiguous()
            _ready()
            L.check(L.load().tsem_l2_error_chunk(self.h, _ptr(x), e0, e1, _ptr(uq), C.byref(s)))
            tot += s.value
        return math.sqrt(tot)

    def l2_error_gradient(self, x, grad, chunk=None):
        """calc_l2_error_gradient: ||grad u_h - grad u|| with grad a function
        of the points (device) returning (..., 3)."""
        import math
        if chunk is None:
            chunk = max(1, min(self.nt, int(2e8 // max(1, self.nq3 * 6))))
        tot = 0.0
        s = C.c_double()
        for e0 in range(0, self.nt, chunk):
            e1 = min(self.nt, e0 + chunk)
            g = grad(self.qp_coords(e0, e1)).contiguous()
            _ready()
            L.check(L.load().tsem_h1_error_chunk(self.h, _ptr(x), e0, e1, _ptr(g), C.byref(s)))
            tot += s.value
        return math.sqrt(tot)

    # ---- coefficient files (write_coef / read_coef, src/Miscellaneous.cpp:34-125)
    def _geometry_layout(self):
        """(first DOF, dimension) of every geometry, in DOF-layout order."""
        start = self.dof_start()
        dims = np.repeat(np.arange(4), [self.nv, self.ne, self.nf, self.nt])
        order = np.argsort(start, kind="stable")
        return start[order], dims[order]

    def write_coef(self, x, path):
        """Text file: 'M<TAB>n_dof', then one value per line (14 significant
        digits); x (n_dof) or a list of such vectors, one after the other."""
        vecs = x if isinstance(x, (list, tuple)) else [x]
        with open(path, "w") as f:
            f.write("%d\t%d\n" % (self.M, self.ndof))
            for v in vecs:
                v = v.cpu().numpy() if hasattr(v, "cpu") else np.asarray(v)
                f.write("".join("%.14g\n" % t for t in v))

    def read_coef(self, path, count=1):
        """Vectors of this space from a coefficient file of order M' (<= or >=
        M): per geometry the first C(min(M, M') - 1, dim) coefficients are
        taken, the rest are zero (hierarchical basis).  Returns count host
        arrays (one array when count == 1)."""
        from math import comb
        toks = open(path).read().split()
        m_in, n_in = int(toks[0]), int(toks[1])
        vals = np.array(toks[2:2 + count * n_in], dtype=np.float64).reshape(count, n_in)
        start, dims = self._geometry_layout()
        nd_in = np.array([comb(m_in - 1, d) for d in range(4)])
        start_in = np.concatenate([[0], np.cumsum(nd_in[dims])[:-1]])
        keep = np.minimum(nd_in[dims], np.array([comb(self.M - 1, d) for d in range(4)])[dims])
        out = np.zeros((count, self.ndof))
        for g in np.nonzero(keep)[0]:
            k = keep[g]
            out[:, start[g]:start[g] + k] = vals[:, start_in[g]:start_in[g] + k]
        return out[0] if count == 1 else list(out)

    def lda_pz(self, rho):
        """LDA exchange (Slater) + Perdew-Zunger correlation: (v_xc, e_xc) at the points."""
        import torch
        rho = rho.contiguous()
        v = torch.empty_like(rho)
        e = torch.empty_like(rho)
        _ready()
        L.check(L.load().tsem_lda_pz(self.h, rho.numel(), _ptr(rho), _ptr(v), _ptr(e)))
        return v, e

    def set_potential(self, v):
        """Potential at the quadrature points (nt x n_qp3) weighting M_V; None clears."""
        if v is not None:
            v = v.contiguous()
        _ready()
        L.check(L.load().tsem_set_potential(self.h, _ptr(v) if v is not None else None))

    def apply_op(self, x, stiff=1.0, mass=0.0, potential=0.0, constrained=False):
        """y = (stiff S + mass M + potential M_V) x (FP64)."""
        import torch
        y = torch.empty_like(x)
        _ready()
        L.check(L.load().tsem_apply_potential(self.h, stiff, mass, potential, 1 if constrained else 0, _ptr(x),
                                              _ptr(y)))
        return y

    def neumann(self, g_face_qp, flag, b):
        g = np.ascontiguousarray(g_face_qp, dtype=np.float64)
        _ready()
        L.check(L.load().tsem_neumann(self.h, L.dptr(g), flag, _ptr(b)))

    def dirichlet(self, uv, ue, uf, flag, b):
        uv, ue, uf = (np.ascontiguousarray(a, dtype=np.float64) for a in (uv, ue, uf))
        bval = np.zeros(self.ndof)
        _ready()
        L.check(L.load().tsem_dirichlet(self.h, L.dptr(uv), L.dptr(ue), L.dptr(uf), flag, _ptr(b), L.dptr(bval)))
        return bval

    def zero_bc(self, flag):
        L.check(L.load().tsem_zero_bc(self.h, flag))

    def zero_rows(self, flag, x):
        """impose_zero_boundary_condition(Vector&, flag): x = 0 on the DOFs of
        geometries with mark == flag (in place, device)."""
        _ready()
        L.check(L.load().tsem_zero_rows(self.h, flag, _ptr(x)))

    def apply(self, x, stiff=1.0, mass=0.0, constrained=False, out=None):
        import torch
        prec = 64 if x.dtype == torch.float64 else 32
        y = torch.empty_like(x) if out is None else out
        _ready()
        L.check(L.load().tsem_apply(self.h, stiff, mass, 1 if constrained else 0, prec, _ptr(x), _ptr(y)))
        return y

    def gather(self, x):
        import torch
        xl = torch.empty((self.nt, self.K), dtype=torch.float64, device=self.device)
        _ready()
        L.check(L.load().tsem_gather(self.h, _ptr(x), _ptr(xl)))
        return xl

    def scatter(self, yl):
        y = self.zeros()
        yl = yl.contiguous()
        _ready()
        L.check(L.load().tsem_scatter(self.h, _ptr(yl), _ptr(y)))
        return y

    def diagonal(self):
        d = self.zeros()
        _ready()
        L.check(L.load().tsem_diagonal(self.h, _ptr(d)))
        return d

    def pcg(self, b, x=None, tol=1e-10, criterion=1, max_iter=100000, precond=1, mixed=False, inner_tol=1e-2,
            inner_max=100000, smooth_steps=2, check_every=1, raise_on_noconv=True):
        """solve_PCG on the eliminated operator; returns (x, stats dict).
        mixed: 0 (False) FP64 PCG; 1 (True) FP32 PCG with reliable updates
        every inner_tol of residual reduction; 2 FP64 refinement around
        restarted FP32 solves to inner_tol each."""
        x = self.zeros() if x is None else x
        o = L.PcgOpts(tol, criterion, max_iter, precond, int(mixed), inner_tol, inner_max, smooth_steps,
                      check_every)
        s = L.PcgStats()
        _ready()
        rc = L.load().tsem_pcg(self.h, _ptr(b), _ptr(x), C.byref(o), C.byref(s))
        if rc != L.TSEM_OK and (rc != L.TSEM_ENOCONV or raise_on_noconv):
            L.check(rc)
        return x, dict(iterations=s.iterations, outer=s.outer, relres=s.relres, l1=s.l1, seconds=s.seconds,
                       converged=(rc == L.TSEM_OK))


    def lobpcg(self, m, X=None, seed=7, tol=1e-8, criterion=1, max_iter=1000, precond=4, smooth_steps=4,
               stiff=1.0, mass=0.0, potential=0.0, raise_on_noconv=True):
        """solve_LOBPCG: the m lowest eigenpairs of (stiff S + mass M + potential M_V) x
        = lambda M x on the eliminated operators (M_V: set_potential).  X: (m, n_dof) float64 device tensor
        (row j = vector j), default N(0, 1) from `seed`.  smooth_steps: the
        Chebyshev degree of the preconditioning V-cycle (4: the helium SCF
        runs 10% faster than with 3, the 32-pair Laplacian the same).  Returns
        (lambda (numpy, ascending), X, stats dict)."""
        import torch
        if X is None:
            g = torch.Generator(device="cpu").manual_seed(seed)
            X = torch.randn((m, self.ndof), generator=g, dtype=torch.float64).to(self.device)
        X = X.contiguous()
        lam = np.zeros(m)
        o = L.LobpcgOpts(tol, criterion, max_iter, precond, smooth_steps, stiff, mass, potential)
        s = L.LobpcgStats()
        _ready()
        rc = L.load().tsem_lobpcg(self.h, m, _ptr(X), L.dptr(lam), C.byref(o), C.byref(s))
        if rc != L.TSEM_OK and (rc != L.TSEM_ENOCONV or raise_on_noconv):
            L.check(rc)
        return lam, X, dict(iterations=s.iterations, applies=s.applies, max_resid=s.max_resid,
                            max_dlambda=s.max_dlambda, seconds=s.seconds, converged=(rc == L.TSEM_OK))

    def inverse_power(self, x=None, seed=7, tol=1e-12, max_iter=1000, inner_tol=1e-14, precond=1):
        """solve_InversePower: lowest eigenpair of S x = lambda M x (eliminated)."""
        import torch
        if x is None:
            g = torch.Generator(device="cpu").manual_seed(seed)
            x = torch.rand(self.ndof, generator=g, dtype=torch.float64).to(self.device)
        o = L.PcgOpts(inner_tol, 1, 100000, precond, 0, 1e-4, 100000, 2, 20)
        lam = np.zeros(1)
        its = C.c_int()
        _ready()
        L.check(L.load().tsem_inverse_power(self.h, _ptr(x), tol, max_iter, C.byref(o), L.dptr(lam), C.byref(its)))
        return float(lam[0]), x, its.value


def poisson_setup(sp, problem, dirichlet_planes=0b000001, neumann_planes=0b111110, chunk=None, pause=None):
    """Mixed-BC Poisson system on the unit cube, the sequence of
    examples/poisson_mixBC/poisson.cpp:74-89: load vector, Neumann load on
    mark 2, Dirichlet elimination on mark 1.  Returns b (device).  chunk and
    pause go to Space.rhs."""
    import torch
    from .problems import cube_face_normal
    sp.set_cube_marks(dirichlet_planes, neumann_planes)
    b = sp.rhs(lambda p: problem.f(p), chunk=chunk, pause=pause)
    ev, fv, _, _ = sp.topology()
    _, _, pf2 = sp.bc_points(2)

    # boundary data evaluated on the device, handed to the library as host arrays
    def dev(a):
        return torch.from_numpy(np.ascontiguousarray(a)).to(sp.device)

    normals = dev(cube_face_normal(sp.xyz[fv[sp.marked_faces(2)]]))
    gq = problem.dudn(dev(pf2), [normals[:, None, d] for d in range(3)]).cpu().numpy()
    sp.neumann(gq, 2, b)
    pv, pe, pf = sp.bc_points(1)
    uv, ue, uf = (problem.u(dev(q)).cpu().numpy() for q in (pv, pe, pf))
    sp.bval = sp.dirichlet(uv, ue, uf, 1, b)
    # torch's own stream only: a device-wide synchronization would break a
    # graph capture running in another thread
    torch.cuda.current_stream().synchronize()
    return b
