"""dgk_minmax_noalloc: min/max of a device array with the reduction's
partials in the library's static device scratch, checked through the C ABI
against dgk_minmax of a solver holding the same field.

Two calls in a row, the second with a smaller input that launches fewer
blocks than the first: the array keeps the first call's partials, and the
second result must not see them.  Bit-exact (min and max are exact).  Device
memory for the inputs and outputs comes from hipMalloc of the same HIP
runtime that libdgk.so links; the library call itself allocates nothing.
"""
import ctypes as C

import numpy as np
import pytest

import dgk
from dgk import inputs

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def hip():
    dgk.lib()  # load libdgk.so first, so the runtime below is the one it uses
    h = C.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
    h.hipStreamSynchronize.argtypes = [C.c_void_p]
    h.hipStreamDestroy.argtypes = [C.c_void_p]
    return h


class DeviceArray:
    def __init__(self, h, nbytes):
        self.h, self.nbytes, self.ptr = h, nbytes, C.c_void_p()
        assert h.hipMalloc(C.byref(self.ptr), nbytes) == 0

    def upload(self, a):
        a = np.ascontiguousarray(a, np.float64)
        assert a.nbytes <= self.nbytes
        assert self.h.hipMemcpy(self.ptr, a.ctypes.data, a.nbytes, H2D) == 0

    def download(self, n):
        out = np.empty(n)
        assert self.h.hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, D2H) == 0
        return out

    def free(self):
        if self.ptr:
            assert self.h.hipFree(self.ptr) == 0
            self.ptr = C.c_void_p()


def last_error():
    return dgk.lib().dgk_last_error().decode()


def solver_minmax(ne, p, q):
    with dgk.Solver.structured("HIP", ne, ne, ne, p) as s:
        s.set_field(dgk.FIELD_Q, q)
        return s.minmax()


def noalloc(q_dev, n, res_dev, stream):
    return dgk.lib().dgk_minmax_noalloc(q_dev.ptr, n, res_dev.ptr, stream)


def test_noalloc_equals_minmax_on_two_calls_the_second_smaller():
    h = hip()
    stream = C.c_void_p()
    assert h.hipStreamCreate(C.byref(stream)) == 0
    big = 16 ** 3 * 4 ** 3                      # 16^3 elements at p = 3: 128 blocks
    d_q, d_res = DeviceArray(h, 8 * big), DeviceArray(h, 16)
    try:
        cases = [(16, 3, 5.0, 31), (3, 2, 0.25, 32)]  # second: 729 values, one block
        for ne, p, scale, seed in cases:
            Ne, Np = ne ** 3, (p + 1) ** 3
            q, _, _, _ = inputs.random_fields(ne, ne, ne, p, seed=seed)
            q = scale * q + 0.1 * scale
            ref = solver_minmax(ne, p, q)
            n = Ne * Np
            d_q.upload(q)
            assert noalloc(d_q, n, d_res, stream) == dgk.OK, last_error()
            assert h.hipStreamSynchronize(stream) == 0
            got = tuple(d_res.download(2))
            assert got == ref, (ne, p, got, ref)
        # the first call's extremes lie outside the second field's range, so
        # a fold over stale partials would have shown them
        assert cases[0][2] > 4 * cases[1][2]
    finally:
        d_q.free()
        d_res.free()
        assert h.hipStreamDestroy(stream) == 0


def test_noalloc_on_the_null_stream_and_refusals():
    h = hip()
    n = 5000
    d_q, d_res = DeviceArray(h, 8 * n), DeviceArray(h, 16)
    try:
        q = np.linspace(-2.0, 3.0, n)
        d_q.upload(q)
        assert noalloc(d_q, n, d_res, None) == dgk.OK, last_error()
        assert h.hipStreamSynchronize(None) == 0
        assert tuple(d_res.download(2)) == (-2.0, 3.0)
        lib = dgk.lib()
        host = np.zeros(n)
        cases = [
            (lambda: lib.dgk_minmax_noalloc(d_q.ptr, 0, d_res.ptr, None), "outside"),
            (lambda: lib.dgk_minmax_noalloc(d_q.ptr, (1 << 40) + 1, d_res.ptr, None), "outside"),
            (lambda: lib.dgk_minmax_noalloc(d_q.ptr, 1 << 30, d_res.ptr, None), "past its allocation"),
            (lambda: lib.dgk_minmax_noalloc(host.ctypes.data, n, d_res.ptr, None), "not device"),
            (lambda: lib.dgk_minmax_noalloc(d_q.ptr, n, host.ctypes.data, None), "not device"),
            (lambda: lib.dgk_minmax_noalloc(None, n, d_res.ptr, None), "null"),
        ]
        for call, msg in cases:
            assert call() == dgk.ERR_ARG
            assert msg in last_error(), last_error()
        # nothing was launched by the refused calls: the result is unchanged
        assert tuple(d_res.download(2)) == (-2.0, 3.0)
    finally:
        d_q.free()
        d_res.free()
