"""The allocation-free closing (gm_close_cells_noalloc) against gm_close
through a context, on caller-owned device grids: two calls in a row, the
second with a smaller grid, then the limits (a grid beyond the scratch's
maximum is refused before any work, a radius <= 1 copies the input)."""
import ctypes

import numpy as np
import pytest

import broxy_amd as B
from _util import random_blobs

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2  # hipMemcpyKind


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def hip():
    h = ctypes.CDLL("libamdhip64.so")
    h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    h.hipFree.argtypes = [ctypes.c_void_p]
    h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    return h


def _lib():
    L = B.hip_lib()
    L.gm_close_cells_noalloc.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                                         ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
    L.gm_noalloc_last_error.restype = ctypes.c_char_p
    return L


def _context_closing(R, cells, dr, r):
    c = B.Context(0)
    try:
        bmin = np.zeros(3, np.float32)
        cell = np.float32(1.0 / R)
        c.grid_init(bmin, bmin + cell * R, R, np.asarray(dr, np.int32), float(cell))
        c.set_cells(B.GRID_INPUT, cells)
        c.scale_uniform(r)
        c.close_grid(B.SE_SPHERE)
        return c.get_cells(B.GRID_CLOSING)
    finally:
        c.close()


def _noalloc_closing(hip, L, R, cells, dr, r):
    n = R ** 3
    din, dout = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(din), n) == 0 and hip.hipMalloc(ctypes.byref(dout), n) == 0
    try:
        assert hip.hipMemcpy(din, cells.ctypes.data, n, H2D) == 0
        d = np.ascontiguousarray(dr, np.int32)
        rc = L.gm_close_cells_noalloc(0, din, dout, R, d.ctypes.data, r, None)
        assert rc == 0, L.gm_noalloc_last_error()
        out = np.zeros(n, np.uint8)
        assert hip.hipDeviceSynchronize() == 0
        assert hip.hipMemcpy(out.ctypes.data, dout, n, D2H) == 0
        return out
    finally:
        hip.hipFree(din)
        hip.hipFree(dout)


def test_noalloc_matches_context_closing(hip):
    L = _lib()
    assert L.gm_noalloc_init(0) == 0
    for R, r, seed in ((128, 5, 31), (64, 3, 32)):
        cells = random_blobs(R, seed=seed, nballs=10, rmax=9.0 if R > 64 else 5.0, margin=4)
        dr = [R, R - 3, R - 5]
        say("allocation-free closing R=%d r=%d" % (R, r))
        ref = _context_closing(R, cells, dr, r)
        got = _noalloc_closing(hip, L, R, cells, dr, r)
        assert np.array_equal(got, ref), "differs in %d cells" % int((got != ref).sum())
        assert got.any() and not np.array_equal(got, cells)


def test_noalloc_limits(hip):
    L = _lib()
    # beyond the maximum: refused before any device work (dummy pointers)
    d = np.array([144, 144, 144], np.int32)
    rc = L.gm_close_cells_noalloc(0, ctypes.c_void_p(256), ctypes.c_void_p(512), 144, d.ctypes.data, 4, None)
    assert rc == 4 and b"exceeds the allocation-free limit" in L.gm_noalloc_last_error()
    rc = L.gm_close_cells_noalloc(0, ctypes.c_void_p(256), ctypes.c_void_p(512), 64, d.ctypes.data, 32, None)
    assert rc == 5
    # radius <= 1: the closing is the input
    R = 32
    cells = random_blobs(R, seed=5, nballs=4)
    say("allocation-free closing R=32 r=1 (copy)")
    assert np.array_equal(_noalloc_closing(hip, L, R, cells, [R, R, R], 1), cells)
