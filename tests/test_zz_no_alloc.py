"""vgpu_mlp_loss_grad_noalloc (the entry point that allocates nothing): its
results through the C ABI equal vgpu_mlp_loss_grad's bit for bit, on two
calls in a row (the second with fewer samples), and sizes beyond its scratch
are rejected.  Device buffers come from the HIP runtime the engine library
itself loaded (ctypes), so the test needs nothing else on the device."""
import ctypes as C

import numpy as np
import pytest

import voro_amd

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


class _Hip:
    """hipMalloc / hipMemcpy / hipFree of the loaded HIP runtime."""

    def __init__(self):
        voro_amd.lib()      # loads libamdhip64 with the engine
        self.h = C.CDLL("libamdhip64.so.7")
        self.h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.h.hipFree.argtypes = [C.c_void_p]
        self.live = []

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        assert self.h.hipMalloc(C.byref(p), max(a.nbytes, 4)) == 0
        self.live.append(p)
        assert self.h.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), a.nbytes, H2D) == 0
        return p

    def get(self, p, like):
        out = np.empty_like(like)
        assert self.h.hipMemcpy(out.ctypes.data_as(C.c_void_p), p, out.nbytes, D2H) == 0
        return out

    def free(self):
        for p in self.live:
            self.h.hipFree(p)
        self.live = []


def _f32(p):
    return C.cast(p, C.POINTER(C.c_float))


def test_gpu_mlp_loss_grad_noalloc_equals_allocating_entry_point():
    L = voro_amd.lib()
    NP = L.vgpu_mlp_num_params()
    hip = _Hip()
    rng = np.random.default_rng(3)
    p = rng.normal(0, 0.3, NP).astype(np.float32)
    try:
        for n in (20000, 777):
            print("noalloc loss/grad n=%d" % n, flush=True)
            X = rng.uniform(0, 2, (n, 28)).astype(np.float32)
            T = rng.uniform(0, 1, (n, 2)).astype(np.float32)
            lo_ref, g_ref = voro_amd.mlp_loss_grad(p, X, T)
            nan_g = np.full(NP, np.nan, np.float32)
            nan_l = np.full(1, np.nan, np.float32)
            dp, dX, dT, dg, dl = hip.put(p), hip.put(X), hip.put(T), hip.put(nan_g), hip.put(nan_l)
            st = L.vgpu_mlp_loss_grad_noalloc(_f32(dp), _f32(dX), _f32(dT), n, _f32(dl), _f32(dg), None)
            assert st == 0, st
            g = hip.get(dg, nan_g)
            lo = float(hip.get(dl, nan_l)[0])
            assert lo == lo_ref
            assert np.array_equal(g, g_ref)
    finally:
        hip.free()


def test_gpu_mlp_loss_grad_noalloc_rejects_oversize():
    L = voro_amd.lib()
    NP = L.vgpu_mlp_num_params()
    cap = 32768
    hip = _Hip()
    try:
        dp = hip.put(np.zeros(NP, np.float32))
        dX = hip.put(np.zeros((cap + 1, 28), np.float32))
        dT = hip.put(np.zeros((cap + 1, 2), np.float32))
        dg = hip.put(np.zeros(NP, np.float32))
        st = L.vgpu_mlp_loss_grad_noalloc(_f32(dp), _f32(dX), _f32(dT), cap + 1, None, _f32(dg), None)
        assert st == voro_amd.INVALID_ARGUMENT
    finally:
        hip.free()
