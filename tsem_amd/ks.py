"""Kohn-Sham self-consistent field for an atom on a graded box.

The steps of the reference's helium example (examples/KohnSham-Helium/
SCFIteration.cpp) on the device library:

* Hamiltonian H = S/2 + M_V, V = V_ext + v_H + v_xc at the quadrature points
  (assemble_matrix, :15-51), applied matrix-free (Space.set_potential);
* the lowest states of H psi = eps M psi by block LOBPCG with zero Dirichlet
  data on the box (solve, :391);
* density rho = sum_k occ_k psi_k^2 at the points, linearly mixed
  rho <- rho + alpha (rho_new - rho), the kinetic energy mixed the same way
  (update_density, :53-67);
* Hartree potential from -Lap v_H = 4 pi rho with Dirichlet data from the
  multipole expansion (monopole, dipole, quadrupole) of rho on the box
  faces (calc_V_Har, :102-177), solved by mixed-precision PCG;
* energy E = E_kin + int V_ext rho + 1/2 int v_H rho + int e_xc(rho)
  (calc_energy, :232-258).

Differences from the example, by design: LDA correlation is Perdew-Zunger
1981 instead of VWN (its lines 192-230); several states are computed (the
example solves one); SCF stops after a fixed number of steps or when the
energy change falls below the tolerance.
"""
import math
import time

import numpy as np

from . import api, mesh


class KohnShamAtom:
    """One nucleus of charge Z at the origin of [-half, half]^3."""

    def __init__(self, ctx, n=24, p=5, Z=2.0, occ=(2.0, 0.0, 0.0, 0.0), half=10.0, beta=2.0, qorder=None,
                 hartree=True, xc=True, server=None):
        """server: a serve.NativeServer that takes the Hartree solve of every
        SCF step as a system request (None: solved here directly)."""
        import torch
        self.server = server
        self.Z, self.half = float(Z), float(half)
        self.occ = np.asarray(occ, dtype=np.float64)
        self.nstates = len(self.occ)
        self.hartree, self.xc = hartree, xc
        xyz, tet = mesh.graded_box(n, half, beta)
        self.sp = sp = api.Space(ctx, p, xyz, tet, qorder=qorder or p + 3)
        sp.set_cube_marks(0b111111, 0, lo=-half, hi=half)
        sp.zero_bc(1)
        pts = sp.qp_coords()
        r = torch.linalg.norm(pts, dim=-1)
        self.vext = -self.Z / r
        # initial density: occupied charge in exp(-2 r) (wave_init of the example)
        nel = float(self.occ.sum())
        rho0 = torch.exp(-4.0 * r)
        self.rho = rho0 * (nel / sp.integrate(rho0)) if nel > 0 else rho0 * 0
        self.vH = sp.zeros()
        self.vH_qp = torch.zeros_like(self.rho)
        self.X = None
        self.ekin = None
        self.bc_pts = sp.bc_points(1)
        self.history = []

    # ---------------------------------------------------------------- pieces
    def multipole(self, pts, mom):
        """Potential of a charge distribution with moments mom = int rho {1,
        x, y, z, xx, xy, xz, yy, yz, zz} at points far from it:
        q/r + p.x/r^3 + sum_ij Q_ij (3 x_i x_j - r^2 delta_ij) / (2 r^5)."""
        x, y, z = pts[..., 0], pts[..., 1], pts[..., 2]
        r2 = x * x + y * y + z * z
        r = np.sqrt(r2)
        q, px, py, pz = mom[0], mom[1], mom[2], mom[3]
        Qxx, Qxy, Qxz, Qyy, Qyz, Qzz = mom[4:10]
        quad = (Qxx * (3 * x * x - r2) + Qyy * (3 * y * y - r2) + Qzz * (3 * z * z - r2)
                + 2 * 3 * (Qxy * x * y + Qxz * x * z + Qyz * y * z)) / (2 * r2 * r2 * r)
        return q / r + (px * x + py * y + pz * z) / (r2 * r) + quad

    def solve_hartree(self, tol=1e-10):
        """v_H from -Lap v_H = 4 pi rho, multipole Dirichlet data on the box."""
        sp = self.sp
        b = sp.rhs_qp(4.0 * math.pi * self.rho)
        mom = sp.integrate(self.rho, moments=True)
        pv, pe, pf = self.bc_pts
        sp.dirichlet(self.multipole(pv, mom), self.multipole(pe, mom), self.multipole(pf, mom), 1, b)
        if self.server is not None:
            # the served Poisson solve of the step: queued among the server's
            # other requests, on this space
            rid = self.server.submit_system(sp, b, self.vH, tol=tol, mixed=1, precond=1)
            _, _, res = self.server.poll(rid)
            st = dict(iterations=res["iterations"], relres=res["relres"], served=True)
        else:
            x, st = sp.pcg(b, self.vH, tol=tol, criterion=1, precond=1, mixed=True, check_every=10)
            self.vH = x
        self.vH_qp = sp.qp_values(self.vH)
        return st

    def potential(self):
        v = self.vext.clone()
        if self.hartree:
            v += self.vH_qp
        if self.xc:
            vxc, _ = self.sp.lda_pz(self.rho)
            v += vxc
        return v

    def energy(self):
        sp = self.sp
        e_ext = sp.integrate(self.vext, self.rho)
        e_h = 0.5 * sp.integrate(self.vH_qp, self.rho) if self.hartree else 0.0
        e_xc = sp.integrate(self.sp.lda_pz(self.rho)[1]) if self.xc else 0.0
        return dict(total=self.ekin + e_ext + e_h + e_xc, kinetic=self.ekin, external=e_ext, hartree=e_h, xc=e_xc)

    def kinetic(self, X):
        """sum_k occ_k psi_k^T (S/2) psi_k."""
        tot = 0.0
        for k in range(self.nstates):
            if self.occ[k] == 0:
                continue
            sx = self.sp.apply_op(X[k], stiff=0.5, constrained=True)
            tot += self.occ[k] * float((sx * X[k]).sum())
        return tot

    # ------------------------------------------------------------------- SCF
    def run(self, steps=20, alpha=0.3, tol_energy=0.0, tol_eig=1e-8, precond=4, log=None):
        sp = self.sp
        if self.hartree:
            self.solve_hartree()
        for it in range(1, steps + 1):
            t0 = time.perf_counter()
            sp.set_potential(self.potential())
            lam, X, st = sp.lobpcg(self.nstates, X=self.X, tol=tol_eig, precond=precond, stiff=0.5,
                                   potential=1.0, max_iter=500, raise_on_noconv=False)
            self.X = X
            ek = self.kinetic(X)
            rho_new = sp.density(X, self.occ)
            self.rho = self.rho + alpha * (rho_new - self.rho)
            self.ekin = ek if self.ekin is None else self.ekin + alpha * (ek - self.ekin)
            if self.hartree:
                self.solve_hartree()
            e = self.energy()
            rec = dict(step=it, energy=e["total"], eigenvalues=lam.tolist(), lobpcg_iterations=st["iterations"],
                       lobpcg_converged=st["converged"], seconds=time.perf_counter() - t0)
            self.history.append(rec)
            if log:
                log(rec)
            if it > 1 and abs(self.history[-1]["energy"] - self.history[-2]["energy"]) < tol_energy:
                break
        return self.history[-1]
