"""vlm_filament_vind_noalloc (device-resident inputs, no allocation)
against vlm_filament_vind on the same filaments and points: two calls in a
row, the second with smaller inputs (the workspace keeps the first call's
contents), one of them accumulating, one on the small-target path.  Device
buffers come from the HIP runtime through ctypes; only the entry point
under test must not allocate."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


class _Hip:
    def __init__(self):
        self.lib = C.CDLL("libamdhip64.so")
        self.lib.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.lib.hipFree.argtypes = [C.c_void_p]
        self.lib.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.bufs = []

    def put(self, arr):
        arr = np.ascontiguousarray(arr, dtype=np.float64)
        p = C.c_void_p()
        assert self.lib.hipMalloc(C.byref(p), max(arr.nbytes, 8)) == 0
        self.bufs.append(p)
        if arr.nbytes:
            assert self.lib.hipMemcpy(p, arr.ctypes.data, arr.nbytes, H2D) == 0
        return p

    def get(self, p, shape):
        out = np.zeros(shape)
        assert self.lib.hipMemcpy(out.ctypes.data, p, out.nbytes, D2H) == 0
        return out

    def free(self):
        for p in self.bufs:
            self.lib.hipFree(p)
        self.bufs = []


def _inputs(nfil, m, seed):
    rng = np.random.default_rng(seed)
    a, b = rng.uniform(-2, 2, size=(2, nfil, 3))
    g, rc = rng.normal(size=nfil), np.full(nfil, 0.02)
    p = rng.uniform(-2, 2, size=(m, 3))
    return a, b, g, rc, p


def _noalloc(ctx, hip, a, b, g, rc, p, v0=None):
    m = len(p)
    ptrs = [hip.put(x) for x in (a, b, g, rc, p)]
    dv = hip.put(np.zeros((m, 3)) if v0 is None else v0)
    da, db, dg, drc, dp = ptrs
    ctx._chk(ctx.lib.vlm_filament_vind_noalloc(ctx.h, len(a), da, db, dg, drc, m, dp, dv,
                                                0 if v0 is None else 1))
    ctx.synchronize()
    return hip.get(dv, (m, 3))


def test_noalloc_matches_filament_vind_on_two_calls():
    from volcanor_amd import vlm
    ctx = vlm.Context(0)
    hip = _Hip()
    try:
        for prec in (vlm.FP64, vlm.MIXED):
            ctx.set_precision(prec)
            tol = 1e-12 if prec == vlm.FP64 else 1e-7
            # first call: work-queue path; second: smaller, accumulating;
            # third: small-target path
            for nfil, m, acc in ((40000, 3000, False), (5000, 1200, True), (3000, 300, False)):
                a, b, g, rc, p = _inputs(nfil, m, nfil + m)
                v0 = np.random.default_rng(m).normal(size=(m, 3)) if acc else None
                ref = ctx.filament_vind(a, b, g, rc, p, v=v0)
                got = _noalloc(ctx, hip, a, b, g, rc, p, v0)
                scale = np.max(np.abs(ref))
                assert np.max(np.abs(got - ref)) <= tol * scale, (prec, nfil, m)
            hip.free()
    finally:
        hip.free()
        ctx.close()


def test_noalloc_rejects_sizes_beyond_the_workspace():
    from volcanor_amd import vlm
    ctx = vlm.Context(0)
    with pytest.raises(vlm.VLMError, match="MiB workspace"):
        ctx._chk(ctx.lib.vlm_filament_vind_noalloc(ctx.h, 1 << 20, 1, 1, 1, 1, 1, 1, 1, 0))
    with pytest.raises(vlm.VLMError, match="null buffer"):
        ctx._chk(ctx.lib.vlm_filament_vind_noalloc(ctx.h, 10, None, None, None, None, 5, None,
                                                   None, 0))
    ctx.close()
