"""pb_hip_dense_solve_noalloc (the entry point that allocates nothing) against pb_hip_dense_solve, through the C ABI.

Two calls in a row on caller device buffers, the second with a smaller system, must give bitwise the solutions of
the allocating entry point (the same kernels on the same padded matrix); a system above the scratch limit is
refused with PB_ERR_INVALID. The device buffers are the test's own (hipMalloc through the HIP runtime).
"""
import ctypes as C
import sys

import numpy as np
import pytest

from plabundle_amd import _abi
from plabundle_amd.solver import SolverWorkspace

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


class _Hip:
    """The few HIP runtime calls the test needs for its own device buffers."""

    def __init__(self):
        self.lib = C.CDLL("libamdhip64.so")
        self.lib.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.lib.hipFree.argtypes = [C.c_void_p]
        self.lib.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.lib.hipDeviceSynchronize.argtypes = []

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        p = C.c_void_p()
        assert self.lib.hipMalloc(C.byref(p), C.c_size_t(max(1, a.nbytes))) == 0
        assert self.lib.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), C.c_size_t(a.nbytes), 1) == 0  # host to device
        return p

    def download(self, p, n):
        out = np.zeros(n)
        assert self.lib.hipMemcpy(out.ctypes.data_as(C.c_void_p), p, C.c_size_t(out.nbytes), 2) == 0  # to host
        return out

    def free(self, p):
        self.lib.hipFree(p)


def _spd(n, rng):
    M = rng.normal(size=(n, n // 2 + 1))
    return M @ M.T / n + np.eye(n)


def _reference(ws, A, b):
    n = A.shape[0]
    Acm = np.asfortranarray(A).reshape(-1, order="F").copy()
    x = np.zeros(n)
    info = C.c_int32(-1)
    st = _abi.load_library().pb_hip_dense_solve(ws.handle, n, 0, _abi.ptr(Acm, C.c_double),
                                                  _abi.ptr(np.ascontiguousarray(b), C.c_double),
                                                  _abi.ptr(x, C.c_double), C.byref(info), None)
    assert st == _abi.PB_OK, ws.last_error()
    return x, info.value


def test_noalloc_dense_solve_matches_the_allocating_entry_point():
    lib = _abi.load_library()
    assert lib.pb_hip_available(0), "no gfx950 device"
    ws = SolverWorkspace(0)
    hip = _Hip()
    rng = np.random.default_rng(2024)
    try:
        for n in (777, 130, 64):  # second and third calls smaller than the first
            _say(f"case no-alloc dense solve n={n}")
            A = _spd(n, rng)
            b = rng.normal(size=n)
            dA = hip.upload(np.asfortranarray(A).reshape(-1, order="F"))
            db = hip.upload(b)
            info = C.c_int32(-1)
            st = lib.pb_hip_dense_solve_noalloc(ws.handle, n, dA, db, C.byref(info))
            assert st == _abi.PB_OK, ws.last_error()
            x = hip.download(db, n)
            hip.free(dA)
            hip.free(db)
            ref, rinfo = _reference(ws, A, b)
            assert info.value == 0 and rinfo == 0
            np.testing.assert_array_equal(x, ref)
            assert np.linalg.norm(A @ x - b) <= 1e-12 * np.linalg.norm(b)
        # above the scratch limit: refused before any device work, with the limit in the message
        n = _abi.PB_NOALLOC_MAX_ROWS + 1
        st = lib.pb_hip_dense_solve_noalloc(ws.handle, n, C.c_void_p(1), C.c_void_p(1), None)
        assert st == _abi.PB_ERR_INVALID
        assert "exceeds the scratch limit" in ws.last_error()
        # an indefinite matrix is reported
        A = _spd(200, rng)
        A[120, 120] = -3.0
        dA = hip.upload(np.asfortranarray(A).reshape(-1, order="F"))
        db = hip.upload(np.ones(200))
        info = C.c_int32(-1)
        assert lib.pb_hip_dense_solve_noalloc(ws.handle, 200, dA, db, C.byref(info)) == _abi.PB_OK
        assert info.value == 1
        hip.free(dA)
        hip.free(db)
    finally:
        ws.close()
