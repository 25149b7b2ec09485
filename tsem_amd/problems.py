"""Manufactured Poisson problem on the unit cube.

u(x) = sin(pi x) sin(pi y) sin(pi z) * (1 + 0.1 g(x)),
g(x) = 1/8 sum_{m<8} cos(2 pi k_m . x + phi_m),  k_m in {1,2,3}^3, phi_m in [0, 2pi)

drawn from ``numpy.random.default_rng(seed)`` (k first, then phi).  f = -Lap u
and the Neumann data du/dn are exact.  Every function accepts numpy or torch
arrays of points with a trailing dimension of 3; float64 device tensors are
evaluated by the library in one pass (tsem_manufactured).
"""
import math

import numpy as np


class Manufactured:
    def __init__(self, seed=1234, amp=0.1, nterms=8):
        rng = np.random.default_rng(seed)
        self.k = rng.integers(1, 4, size=(nterms, 3)).astype(np.float64)
        self.phi = rng.uniform(0.0, 2 * math.pi, size=nterms)
        self.amp = amp
        self.seed = seed

    @staticmethod
    def _lib(x):
        if isinstance(x, np.ndarray):
            return np
        import torch
        return torch

    def _parts(self, p):
        xp = self._lib(p)
        pi = math.pi
        x, y, z = p[..., 0], p[..., 1], p[..., 2]
        sx, sy, sz = xp.sin(pi * x), xp.sin(pi * y), xp.sin(pi * z)
        cx, cy, cz = xp.cos(pi * x), xp.cos(pi * y), xp.cos(pi * z)
        S = sx * sy * sz
        gS = [pi * cx * sy * sz, pi * sx * cy * sz, pi * sx * sy * cz]
        G = 0.0 * x
        gG = [0.0 * x, 0.0 * x, 0.0 * x]
        LG = 0.0 * x
        nt = len(self.phi)
        for m in range(nt):
            k = self.k[m]
            th = 2 * pi * (k[0] * x + k[1] * y + k[2] * z) + self.phi[m]
            c, s = xp.cos(th), xp.sin(th)
            G = G + c / nt
            for d in range(3):
                gG[d] = gG[d] - 2 * pi * k[d] * s / nt
            LG = LG - 4 * pi * pi * float(k @ k) * c / nt
        return S, gS, G, gG, LG

    def definition(self):
        """The problem as the library's tsem_manufactured_def (cached)."""
        from . import lib
        d = getattr(self, "_def", None)
        if d is None:
            d = lib.ManufacturedDef()
            d.nterms = len(self.phi)
            for m in range(len(self.phi)):
                for j in range(3):
                    d.k[3 * m + j] = float(self.k[m][j])
                d.phi[m] = float(self.phi[m])
            d.amp = float(self.amp)
            self._def = d
        return d

    def _device(self, p, what, normal=None):
        """u, f or du/dn at device points by the library's one-pass
        evaluator (tsem_manufactured), on torch's current stream; None when p
        is not a float64 device tensor."""
        import torch
        if not (isinstance(p, torch.Tensor) and p.is_cuda and p.dtype == torch.float64):
            return None
        import ctypes as C
        from . import lib
        d = self.definition()
        pts = p.reshape(-1, 3).contiguous()
        out = torch.empty(pts.shape[0], dtype=torch.float64, device=p.device)
        if pts.shape[0] == 0:
            return out.reshape(p.shape[:-1])
        nrm = None
        if normal is not None:
            nrm = torch.stack([torch.broadcast_to(c, p.shape[:-1]) for c in normal], -1).reshape(-1, 3).contiguous()
        st = torch.cuda.current_stream(p.device).cuda_stream
        lib.check(lib.load().tsem_manufactured(C.byref(d), what, pts.shape[0], C.c_void_p(pts.data_ptr()),
                                               C.c_void_p(nrm.data_ptr()) if nrm is not None else None,
                                               C.c_void_p(out.data_ptr()), C.c_void_p(st)))
        return out.reshape(p.shape[:-1])

    def u(self, p):
        v = self._device(p, 0)
        if v is not None:
            return v
        S, _, G, _, _ = self._parts(p)
        return S * (1 + self.amp * G)

    def grad(self, p):
        S, gS, G, gG, _ = self._parts(p)
        return [gS[d] * (1 + self.amp * G) + self.amp * S * gG[d] for d in range(3)]

    def f(self, p):
        v = self._device(p, 1)
        if v is not None:
            return v
        S, gS, G, gG, LG = self._parts(p)
        lap = -3 * math.pi ** 2 * S * (1 + self.amp * G) + 2 * self.amp * sum(gS[d] * gG[d] for d in range(3)) \
            + self.amp * S * LG
        return -lap

    def dudn(self, p, normal):
        v = self._device(p, 2, normal)
        if v is not None:
            return v
        g = self.grad(p)
        return g[0] * normal[0] + g[1] * normal[1] + g[2] * normal[2]


def cube_face_normal(face_pts):
    """Outward unit normal of a unit-cube boundary face given its vertices
    (…, 3, 3): the coordinate that is constant on the face decides it."""
    xp = np if isinstance(face_pts, np.ndarray) else __import__("torch")
    spread = face_pts.max(axis=-2) - face_pts.min(axis=-2) if xp is np else \
        face_pts.max(dim=-2).values - face_pts.min(dim=-2).values
    c = face_pts[..., 0, :]
    n = xp.zeros_like(c)
    for d in range(3):
        on = spread[..., d] < 1e-12
        sign = xp.where(c[..., d] > 0.5, 1.0, -1.0) if xp is np else \
            xp.where(c[..., d] > 0.5, xp.ones_like(c[..., d]), -xp.ones_like(c[..., d]))
        n[..., d] = xp.where(on, sign, 0.0 * sign)
    return n
