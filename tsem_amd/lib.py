"""ctypes binding of libtsem_amd.so (the C ABI declared in include/tsem.h).

The library is built in-tree by ``tsem_amd.build``; importing this module
loads it and fails loudly when it is missing: there is no fallback path.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libtsem_amd.so")

TSEM_OK, TSEM_ENOCONV, TSEM_EINVAL, TSEM_ENOMEM, TSEM_EDEVICE, TSEM_ESTATE = range(6)


class TsemError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("tsem error %d: %s" % (code, msg))
        self.code = code


class PcgOpts(C.Structure):
    _fields_ = [("tol", C.c_double), ("criterion", C.c_int), ("max_iter", C.c_int), ("precond", C.c_int),
                ("mixed", C.c_int), ("inner_tol", C.c_double), ("inner_max", C.c_int),
                ("smooth_steps", C.c_int), ("check_every", C.c_int)]


class PcgStats(C.Structure):
    _fields_ = [("iterations", C.c_int), ("outer", C.c_int), ("relres", C.c_double), ("l1", C.c_double),
                ("seconds", C.c_double)]


class LobpcgOpts(C.Structure):
    _fields_ = [("tol", C.c_double), ("criterion", C.c_int), ("max_iter", C.c_int), ("precond", C.c_int),
                ("smooth_steps", C.c_int), ("stiff", C.c_double), ("mass", C.c_double), ("potential", C.c_double)]


class LobpcgStats(C.Structure):
    _fields_ = [("iterations", C.c_int), ("applies", C.c_int), ("max_resid", C.c_double),
                ("max_dlambda", C.c_double), ("seconds", C.c_double)]


class ManufacturedDef(C.Structure):
    _fields_ = [("nterms", C.c_int), ("k", C.c_double * 48), ("phi", C.c_double * 16), ("amp", C.c_double)]


# exported entry points: name -> argument types (all return int status unless noted)
_vp, _ip, _dp, _lp = C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_double), C.POINTER(C.c_int64)
SIGNATURES = {
    "tsem_ctx_create": [C.c_int, C.POINTER(C.c_void_p)],
    "tsem_ctx_destroy": [_vp],
    "tsem_space_create": [_vp, C.c_int, C.c_int, _dp, C.c_int, _ip, _ip, C.c_int, C.POINTER(C.c_void_p)],
    "tsem_space_destroy": [_vp],
    "tsem_space_diagonal_host": [_vp, _dp],
    "tsem_space_info": [_vp, _lp],
    "tsem_space_topology": [_vp, _ip, _ip, _ip, _ip],
    "tsem_space_maps": [_vp, _ip, _ip, _dp],
    "tsem_space_dof_start": [_vp, _lp],
    "tsem_space_sum_dof": [_vp, _ip],
    "tsem_space_ref_matrices": [_vp, _dp, _dp, _dp],
    "tsem_set_marks": [_vp, _ip, _ip, _ip],
    "tsem_space_factors": [_vp, _dp, _dp, _dp],
    "tsem_bc_count": [_vp, C.c_int, _lp],
    "tsem_bc_points": [_vp, C.c_int, _dp, _dp, _dp],
    "tsem_qp_coords": [_vp, C.c_int, C.c_int, _vp],
    "tsem_rhs_begin": [_vp],
    "tsem_rhs_add": [_vp, C.c_int, C.c_int, _vp],
    "tsem_rhs_end": [_vp, _vp],
    "tsem_neumann": [_vp, _dp, C.c_int, _vp],
    "tsem_dirichlet": [_vp, _dp, _dp, _dp, C.c_int, _vp, _dp],
    "tsem_zero_bc": [_vp, C.c_int],
    "tsem_ctx_set_class": [_vp, C.c_int],
    "tsem_h1_error_chunk": [_vp, _vp, C.c_int, C.c_int, _vp, C.POINTER(C.c_double)],
    "tsem_foreground_begin": [_vp],
    "tsem_foreground_end": [_vp],
    "tsem_background_wait": [_vp, C.c_double],
    "tsem_zero_rows": [_vp, C.c_int, _vp],
    "tsem_apply": [_vp, C.c_double, C.c_double, C.c_int, C.c_int, _vp, _vp],
    "tsem_apply_noalloc": [_vp, C.c_double, C.c_double, C.c_int, C.c_int, _vp, _vp],
    "tsem_manufactured": [C.POINTER(ManufacturedDef), C.c_int, C.c_int64, _vp, _vp, _vp, _vp],
    "tsem_gather": [_vp, _vp, _vp],
    "tsem_scatter": [_vp, _vp, _vp],
    "tsem_diagonal": [_vp, _vp],
    "tsem_set_apply_path": [C.c_int],
    "tsem_set_slab": [C.c_int],
    "tsem_space_slab_apply_host": [_vp, _vp, _vp],
    "tsem_space_slab": [_vp, _vp],
    "tsem_point_values": [_vp, _vp, C.c_int64, _vp, _vp, _vp],
    "tsem_serve_open": [C.c_int, _vp, _vp],
    "tsem_serve_submit": [_vp, _vp, _vp],
    "tsem_serve_poll": [_vp, C.c_int64, C.c_int, _vp, _vp, _vp],
    "tsem_serve_close": [_vp],
    "tsem_set_graphs": [C.c_int],
    "tsem_set_block_vcycle": [C.c_int],
    "tsem_timing": [C.c_int],
    "tsem_timing_read": [_dp],
    "tsem_pcg": [_vp, _vp, _vp, C.POINTER(PcgOpts), C.POINTER(PcgStats)],
    "tsem_lobpcg": [_vp, C.c_int, _vp, _dp, C.POINTER(LobpcgOpts), C.POINTER(LobpcgStats)],
    "tsem_inverse_power": [_vp, _vp, C.c_double, C.c_int, C.POINTER(PcgOpts), _dp, C.POINTER(C.c_int)],
    "tsem_qp_values": [_vp, _vp, C.c_int, C.c_int, _vp],
    "tsem_space_set_geometry": [_vp, _dp],
    "tsem_density": [_vp, C.c_int, _vp, _dp, _vp],
    "tsem_integrate": [_vp, _vp, _vp, C.c_int, _dp],
    "tsem_l2_error_chunk": [_vp, _vp, C.c_int, C.c_int, _vp, _dp],
    "tsem_lda_pz": [_vp, C.c_int64, _vp, _vp, _vp],
    "tsem_set_potential": [_vp, _vp],
    "tsem_apply_potential": [_vp, C.c_double, C.c_double, C.c_double, C.c_int, _vp, _vp],
}

_lib = None


def load():
    """Load libtsem_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError("libtsem_amd.so not built: run `python -m tsem_amd.build` (no CPU fallback exists)")
    # torch ships its own HIP runtime (same sonames as /opt/rocm's); loading
    # torch first makes this library bind to that one instance instead of
    # bringing a second runtime into the process
    import torch  # noqa: F401
    L = C.CDLL(LIB_PATH)
    for name, args in SIGNATURES.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = C.c_int
    L.tsem_last_error.restype = C.c_char_p
    L.tsem_last_error.argtypes = []
    _lib = L
    return L


def check(rc):
    if rc != TSEM_OK:
        raise TsemError(rc, load().tsem_last_error().decode())
    return rc


def dptr(a):
    return a.ctypes.data_as(_dp)


def iptr(a):
    return a.ctypes.data_as(_ip)


def lptr(a):
    return a.ctypes.data_as(_lp)
