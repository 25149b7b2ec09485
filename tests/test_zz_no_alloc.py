"""tsem_apply_noalloc (the operator with its scratch in a device-resident
array, for callers that cannot allocate) against tsem_apply through the C
ABI: the same kernels, so the same results bit for bit, on two calls in a
row with the second on a smaller space; spaces past the array's size are
refused before any work is queued."""
import ctypes as C

import numpy as np
import pytest

from tsem_amd import api, lib, mesh

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[case]", *a, flush=True)


def _call(fn, sp, x, stiff, mass, constrained):
    import torch
    prec = 64 if x.dtype == torch.float64 else 32
    y = torch.full_like(x, float("nan"))
    torch.cuda.current_stream().synchronize()
    lib.check(getattr(lib.load(), fn)(sp.h, stiff, mass, 1 if constrained else 0, prec, C.c_void_p(x.data_ptr()),
                                      C.c_void_p(y.data_ptr())))
    return y.cpu().numpy()


@pytest.mark.parametrize("path", [0, 1, 2], ids=["mfma_fp16split", "scalar", "mfma"])
def test_noalloc_apply_matches_apply(ctx, path):
    import torch
    L = lib.load()
    lib.check(L.tsem_set_apply_path(path))
    try:
        # the second space is smaller: the array keeps the first call's
        # contents, and the second call must not read any of them
        for n, M, seed in ((4, 6, 1), (2, 4, 2)):
            _say("noalloc", path, n, M)
            xyz, tet = mesh.kuhn_cube(n, jitter=0.1, seed=seed)
            sp = api.Space(ctx, M, xyz, tet, qorder=M + 3)
            sp.set_cube_marks(0b000011, 0b111100)
            sp.zero_bc(1)
            g = np.random.default_rng(seed).standard_normal(sp.ndof)
            for dtype in (torch.float64, torch.float32):
                x = torch.from_numpy(g).to(dtype).cuda()
                for stiff, mass, cons in ((1.0, 0.0, True), (0.5, -2.0, False)):
                    ref = _call("tsem_apply", sp, x, stiff, mass, cons)
                    got = _call("tsem_apply_noalloc", sp, x, stiff, mass, cons)
                    assert np.isfinite(got).all()
                    assert np.array_equal(got, ref), (n, M, dtype, stiff, mass)
            sp.close()
    finally:
        lib.check(L.tsem_set_apply_path(0))


def test_noalloc_refuses_spaces_past_its_array(ctx):
    """24^3 Kuhn cells at p = 6: the FP64 operator's scratch (results and
    gathered tiles, 111 MB) exceeds the 64 MiB array and is refused with
    TSEM_EINVAL before any work is queued; the FP32 one (60 MB) fits and
    matches tsem_apply."""
    import torch
    _say("noalloc limit")
    xyz, tet = mesh.kuhn_cube(24)
    sp = api.Space(ctx, 6, xyz, tet, qorder=9)
    g = np.random.default_rng(3).standard_normal(sp.ndof)
    x64 = torch.from_numpy(g).cuda()
    with pytest.raises(lib.TsemError) as e:
        _call("tsem_apply_noalloc", sp, x64, 1.0, 0.0, False)
    assert e.value.code == 2 and "too large" in str(e.value)
    x32 = x64.float()
    assert np.array_equal(_call("tsem_apply_noalloc", sp, x32, 1.0, 0.0, False),
                          _call("tsem_apply", sp, x32, 1.0, 0.0, False))
    sp.close()
