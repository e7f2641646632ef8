"""ctypes bindings for the oracle library (test infrastructure only)."""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
# GM_ORACLE_LIB selects another build of the same sources (e.g. the sanitized one).
LIB_PATH = os.environ.get("GM_ORACLE_LIB") or os.path.join(ROOT, "oracle", "_build", "libgmorpho_oracle.so")

_lib = None

P = ctypes.c_void_p


def _ptr(a):
    return None if a is None else a.ctypes.data_as(P)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle library not built: run __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        L.orc_stage_slab.restype = ctypes.c_int64
        L.orc_contour_list.restype = ctypes.c_int64
        _lib = L
    return _lib


def _i3(v):
    return None if v is None else np.ascontiguousarray(v, dtype=np.int32)


def grid_attributes(P_, base_res, margin):
    bmin = np.zeros(3, np.float32)
    bmax = np.zeros(3, np.float32)
    cell = ctypes.c_float()
    dr = np.zeros(3, np.int32)
    lib().orc_grid_attributes(_ptr(P_), len(P_), base_res, ctypes.c_float(margin), _ptr(bmin), _ptr(bmax),
                              ctypes.byref(cell), _ptr(dr))
    return bmin, bmax, cell.value, dr


def voxelize(P_, T_, bmin, maxext, R):
    cells = np.zeros(R ** 3, np.uint8)
    lib().orc_voxelize(_ptr(P_), len(P_), _ptr(T_), len(T_), _ptr(bmin), ctypes.c_float(maxext), R, _ptr(cells))
    return cells


def voxelize_multislice(P_, T_, bmin, maxext, R):
    P_ = np.ascontiguousarray(P_, np.float32)
    T_ = np.ascontiguousarray(T_, np.uint32)
    bmin = np.ascontiguousarray(bmin, np.float32)
    cells = np.zeros(R ** 3, np.uint8)
    lib().orc_voxelize_multislice(_ptr(P_), len(P_), _ptr(T_), len(T_), _ptr(bmin), ctypes.c_float(maxext), R,
                                  _ptr(cells))
    return cells


def scale_from_lattice(lat, N):
    L = lat.shape[0]
    out = np.zeros(N ** 3, np.int8)
    lib().orc_scale_from_lattice(_ptr(np.ascontiguousarray(lat, np.float32)), L, N, _ptr(out))
    return out


def pyramid(cells, R, r=0, scale=None):
    n = 0
    s = R
    while s >= 1:
        n += s ** 3
        s //= 2
    out = np.zeros(n, np.uint64)
    lib().orc_pyramid(_ptr(cells), R, _ptr(scale), r, _ptr(out))
    return out


def dilate(I, R, r=0, scale=None, data_res=None, se=0, brute=0):
    D = np.zeros_like(I)
    lib().orc_dilate(_ptr(I), R, _ptr(_i3(data_res)), _ptr(scale), r, se, brute, _ptr(D))
    return D


def contour_K(D, R):
    K = np.zeros_like(D)
    lib().orc_contour_K(_ptr(D), R, _ptr(K))
    return K


def erode(I, D, R, r=0, scale=None, data_res=None, brute=0):
    C = np.zeros_like(D)
    lib().orc_erode(_ptr(I), _ptr(D), R, _ptr(_i3(data_res)), _ptr(scale), r, brute, _ptr(C))
    return C


def contour_list(G, R, which):
    n = lib().orc_contour_list(_ptr(G), R, which, None)
    out = np.zeros(n, np.uint32)
    lib().orc_contour_list(_ptr(G), R, which, _ptr(out))
    return out


def mc_counts(C, R):
    nv = ctypes.c_uint32()
    nt = ctypes.c_uint32()
    lib().orc_mc_counts(_ptr(C), R, ctypes.byref(nv), ctypes.byref(nt))
    return nv.value, nt.value


def mc(C, R, bmin, cell):
    nv, nt = mc_counts(C, R)
    V = np.zeros((nv, 3), np.float32)
    N = np.zeros((nv, 3), np.float32)
    T = np.zeros((nt, 3), np.uint32)
    lib().orc_mc(_ptr(C), R, _ptr(np.ascontiguousarray(bmin, np.float32)), ctypes.c_float(cell), _ptr(V), _ptr(N),
                 _ptr(T))
    return V, N, T


def bilateral_normals(V, N, T, iters):
    """Filtered copy of the unit normals N after `iters` iterations."""
    out = np.ascontiguousarray(N, np.float32).copy()
    lib().orc_bilateral_normals(_ptr(np.ascontiguousarray(V, np.float32)), _ptr(out),
                                _ptr(np.ascontiguousarray(T, np.uint32)), len(V), len(T), iters)
    return out


def exp_weight(x):
    x = np.ascontiguousarray(x, np.float32)
    out = np.zeros_like(x)
    lib().orc_exp_weight(_ptr(x), _ptr(out), len(x))
    return out


def scale_brush(grid, N, bmin, cell, pos, scale, support, global_scale, mode):
    """Copy of the int8 radius grid after one brush point."""
    out = np.ascontiguousarray(grid, np.int8).copy()
    f = ctypes.c_float
    lib().orc_scale_brush(_ptr(out), N, _ptr(np.ascontiguousarray(bmin, np.float32)), f(cell),
                          _ptr(np.ascontiguousarray(pos, np.float32)), f(scale), f(support), f(global_scale), mode)
    return out


def mc_tables():
    e = np.zeros(256, np.uint16)
    n = np.zeros(256, np.uint8)
    t = np.zeros(4096, np.int8)
    lib().orc_mc_tables(_ptr(e), _ptr(n), _ptr(t))
    return e, n, t.reshape(256, 16)


def stage_slab(stage, I, D, C, R, z0, z1, r=0, scale=None, data_res=None):
    out = np.zeros(R ** 3, np.uint8)
    n = lib().orc_stage_slab(stage, _ptr(I), _ptr(D), _ptr(C), R, _ptr(_i3(data_res)), _ptr(scale), r, z0, z1,
                             _ptr(out))
    return n, out


def num_threads():
    return lib().orc_num_threads()
