"""GMorpho closing path on MI355X: ctypes access to the in-tree libraries.

The product is the C-ABI library ``_lib/libgmorpho_hip.so`` (declared in
``include/gmorpho_c.h``) and its C++ facade ``include/gmorpho.hpp``. This
module only loads those libraries for Python callers (tests, bench); it has no
compute path of its own and raises when the HIP library is missing.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBDIR = os.path.join(HERE, "_lib")
HIP_LIB = os.path.join(LIBDIR, "libgmorpho_hip.so")
MESHGEN_LIB = os.path.join(LIBDIR, "libgmorpho_meshgen.so")

GRID_INPUT, GRID_DILATION, GRID_CONTOUR, GRID_CLOSING = 0, 1, 2, 3
SE_SPHERE, SE_CUBE = 0, 1
BRUSH_NONE, BRUSH_MAX, BRUSH_MIN, BRUSH_ZERO = 0, 1, 2, 3

_hip = None
_meshgen = None
_P = ctypes.c_void_p


def _ptr(a):
    return None if a is None else a.ctypes.data_as(_P)


class GMorphoError(RuntimeError):
    pass


def hip_lib():
    """Load libgmorpho_hip.so (raises if it has not been built)."""
    global _hip
    if _hip is None:
        if not os.path.exists(HIP_LIB):
            raise GMorphoError("libgmorpho_hip.so is not built; run __graft_entry__.build()")
        L = ctypes.CDLL(HIP_LIB)
        L.gm_last_error.restype = ctypes.c_char_p
        L.gm_last_error.argtypes = [_P]
        _hip = L
    return _hip


def copy_probe(nbytes, reps=3, device=0):
    """reps device copies of nbytes by the library's copy kernel (the
    known-size reference for HBM byte counters)."""
    rc = hip_lib().gm_diag_copy_probe(device, ctypes.c_uint64(nbytes), reps)
    if rc:
        raise GMorphoError("gm_diag_copy_probe -> %d" % rc)


def meshgen_lib():
    global _meshgen
    if _meshgen is None:
        if not os.path.exists(MESHGEN_LIB):
            raise GMorphoError("libgmorpho_meshgen.so is not built; run __graft_entry__.build()")
        _meshgen = ctypes.CDLL(MESHGEN_LIB)
    return _meshgen


# ---------------------------------------------------------------------------
# workload generators
def icosphere_counts(subdiv):
    nv, nf = ctypes.c_int(), ctypes.c_int()
    if meshgen_lib().gmg_icosphere_counts(subdiv, ctypes.byref(nv), ctypes.byref(nf)):
        raise ValueError("bad subdivision")
    return nv.value, nf.value


def icosphere(subdiv):
    nv, nf = icosphere_counts(subdiv)
    P = np.zeros((nv, 3), np.float32)
    T = np.zeros((nf, 3), np.uint32)
    meshgen_lib().gmg_icosphere(subdiv, _ptr(P), _ptr(T))
    return P, T


def sh_blob(subdiv, seed, lmax=8, amp=0.1):
    nv, nf = icosphere_counts(subdiv)
    P = np.zeros((nv, 3), np.float32)
    T = np.zeros((nf, 3), np.uint32)
    rc = meshgen_lib().gmg_sh_blob(subdiv, ctypes.c_uint64(seed), lmax, ctypes.c_double(amp), _ptr(P), _ptr(T))
    if rc:
        raise ValueError("gmg_sh_blob failed")
    return P, T


def sphere_soup(nspheres, subdiv, seed, cmin=0.15, cmax=0.85, rmin=0.03, rmax=0.12):
    nv, nf = icosphere_counts(subdiv)
    P = np.zeros((nspheres * nv, 3), np.float32)
    T = np.zeros((nspheres * nf, 3), np.uint32)
    rc = meshgen_lib().gmg_sphere_soup(nspheres, subdiv, ctypes.c_uint64(seed), ctypes.c_double(cmin),
                                       ctypes.c_double(cmax), ctypes.c_double(rmin), ctypes.c_double(rmax), _ptr(P),
                                       _ptr(T))
    if rc:
        raise ValueError("gmg_sphere_soup failed")
    return P, T


def scale_lattice(L, seed, lo=2.0, hi=24.0, smooth=3):
    out = np.zeros(L ** 3, np.float32)
    rc = meshgen_lib().gmg_scale_lattice(L, ctypes.c_uint64(seed), ctypes.c_double(lo), ctypes.c_double(hi), smooth,
                                         _ptr(out))
    if rc:
        raise ValueError("gmg_scale_lattice failed")
    return out


# ---------------------------------------------------------------------------
class Context:
    """One gm_ctx (one device, one stream)."""

    def __init__(self, device=0):
        self._lib = hip_lib()
        h = _P()
        rc = self._lib.gm_create(device, ctypes.byref(h))
        if rc:
            raise GMorphoError("gm_create failed (%d): no usable HIP device" % rc)
        self._h = h

    def close(self):
        if self._h:
            self._lib.gm_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _call(self, name, *args):
        rc = getattr(self._lib, name)(self._h, *args)
        if rc:
            msg = self._lib.gm_last_error(self._h)
            raise GMorphoError("%s -> %d: %s" % (name, rc, msg.decode() if msg else ""))
        return rc

    # GMorpho::Load
    def load(self, P, T, base_res, se_size):
        P = np.ascontiguousarray(P, np.float32)
        T = np.ascontiguousarray(T, np.uint32)
        self._call("gm_load", _ptr(P), None, len(P), _ptr(T), len(T), base_res, ctypes.c_float(se_size))

    def load_mesh(self, P, T):
        P = np.ascontiguousarray(P, np.float32)
        T = np.ascontiguousarray(T, np.uint32)
        self._call("gm_load_mesh", _ptr(P), None, len(P), _ptr(T), len(T))

    def voxelize(self, base_res, margin, multislice=False):
        """GM_GRID_INPUT from the loaded mesh: the conservative voxelization,
        or with multislice=True the reference's alternative MultiSlicing."""
        self._call("gm_voxelize_multislice" if multislice else "gm_voxelize", base_res, ctypes.c_float(margin))

    def grid_init(self, bmin, bmax, base_res, data_res, cell):
        bmin = np.ascontiguousarray(bmin, np.float32)
        bmax = np.ascontiguousarray(bmax, np.float32)
        dr = np.ascontiguousarray(data_res, np.int32)
        self._call("gm_grid_init", _ptr(bmin), _ptr(bmax), base_res, _ptr(dr), ctypes.c_float(cell))

    def info(self):
        bmin = np.zeros(3, np.float32)
        bmax = np.zeros(3, np.float32)
        res = np.zeros(3, np.int32)
        dr = np.zeros(3, np.int32)
        cell = ctypes.c_float()
        self._call("gm_grid_info", _ptr(bmin), _ptr(bmax), _ptr(res), _ptr(dr), ctypes.byref(cell))
        return bmin, bmax, int(res[0]), dr, cell.value

    def set_cells(self, which, cells):
        cells = np.ascontiguousarray(cells, np.uint8)
        self._call("gm_grid_set_cells", which, _ptr(cells))

    def get_cells(self, which):
        R = self.info()[2]
        out = np.zeros(R ** 3, np.uint8)
        self._call("gm_grid_get_cells", which, _ptr(out))
        return out

    def scale_uniform(self, r):
        self._call("gm_scale_uniform", int(r))

    def scale_grid(self, grid_i8):
        g = np.ascontiguousarray(grid_i8, np.int8)
        self._call("gm_scale_grid_i8", _ptr(g))

    def scale_lattice(self, lat, L):
        lat = np.ascontiguousarray(lat, np.float32)
        self._call("gm_scale_lattice_f32", _ptr(lat), L)

    # ScaleField editing: UpdateGridGlobalScale, AddPoint + UpdateGrid, GetScaleGridCPU
    def scale_fill(self, value):
        self._call("gm_scale_grid_fill", int(value))

    def scale_brush(self, pos, scale, support, global_scale=0.0, mode=BRUSH_MAX):
        p = np.ascontiguousarray(pos, np.float32)
        f = ctypes.c_float
        self._call("gm_scale_brush", _ptr(p), f(scale), f(support), f(global_scale), int(mode))

    def scale_grid_get(self):
        N = 2 * self.info()[2]
        out = np.zeros(N ** 3, np.int8)
        self._call("gm_scale_grid_get", _ptr(out))
        return out

    def dilate(self, se=SE_SPHERE):
        self._call("gm_dilate", se)

    def erode(self):
        self._call("gm_erode")

    def close_grid(self, se=SE_SPHERE):
        self._call("gm_close", se)

    def extract_mesh(self):
        self._call("gm_extract_mesh")

    def update(self, se=SE_SPHERE):
        """gm_update: close_grid then extract_mesh as one enqueued sequence."""
        self._call("gm_update", se)

    def mesh_counts(self):
        nv, nt = ctypes.c_uint32(), ctypes.c_uint32()
        self._call("gm_mesh_counts", ctypes.byref(nv), ctypes.byref(nt))
        return nv.value, nt.value

    def mesh(self):
        nv, nt = self.mesh_counts()
        V = np.zeros((nv, 3), np.float32)
        N = np.zeros((nv, 3), np.float32)
        T = np.zeros((nt, 3), np.uint32)
        self._call("gm_mesh_copy", _ptr(V), _ptr(N), _ptr(T))
        return V, N, T

    def filter_normals(self, iters):
        """Bilateral normal filtering of the extracted mesh (ExtractClosingMesh)."""
        self._call("gm_filter_normals", int(iters))

    def build_pyramid(self, which):
        self._call("gm_build_pyramid", which)
        lv = ctypes.c_int()
        tex = ctypes.c_uint64()
        self._call("gm_pyramid_size", ctypes.byref(lv), ctypes.byref(tex))
        out = np.zeros(tex.value, np.uint64)
        self._call("gm_pyramid_get", _ptr(out))
        return out

    def autotune(self, se=SE_SPHERE, reps=3, path=None):
        """Times the closing with each launch shape of the streaming passes,
        records the fastest for this grid size and radius (process-wide;
        written to path when given); returns (planes per tile, waves)."""
        z, w = ctypes.c_int(), ctypes.c_int()
        self._call("gm_autotune", int(se), int(reps), path.encode() if path else None, ctypes.byref(z),
                   ctypes.byref(w))
        return z.value, w.value

    def set_graphs(self, on=True):
        """Replay gm_close / gm_extract_mesh as captured HIP graphs (default
        on). on=True captures on the first call, "repeat" only when the same
        launch sequence comes a second time, False runs eagerly."""
        self._call("gm_set_graphs", 2 if on == "repeat" else (1 if on else 0))

    def set_timing(self, on=True):
        self._call("gm_set_timing", 1 if on else 0)

    STAGES = ("brick_summary", "dilate_coarse", "tight_contour", "dilate_fine", "contour_K", "erode_coarse",
              "base_contour", "erode_fine", "mc_count", "mc_scan", "mc_emit", "mc_fixup_normalize",
              "dilate_coarse_stream", "dilate_fine_stream", "erode_coarse_stream", "erode_fine_stream")

    def stage_times(self):
        """Device ms per stage (one kernel each, see STAGES) of the last
        close_grid / extract_mesh with timing on."""
        ms = np.zeros(len(self.STAGES), np.float32)
        self._call("gm_stage_times", _ptr(ms), len(self.STAGES))
        return ms

    def sync(self):
        self._call("gm_sync")


def tune_load(path):
    """Reads launch shapes recorded by Context.autotune into the process table."""
    rc = hip_lib().gm_tune_load(path.encode())
    if rc:
        raise GMorphoError("gm_tune_load(%s) -> %d" % (path, rc))


_DECIM = None


def decimate_lib():
    global _DECIM
    if _DECIM is None:
        path = os.path.join(LIBDIR, "libgmorpho_decimate.so")
        if not os.path.exists(path):
            raise GMorphoError("libgmorpho_decimate.so not built: run __graft_entry__.build()")
        _DECIM = ctypes.CDLL(path)
        _DECIM.gmd_last_error.restype = ctypes.c_char_p
        _DECIM.gmd_last_error.argtypes = [_P]
    return _DECIM


class Decimator:
    """Proxy decimation of a mesh (the reference's Decimator): HC smoothing,
    then bounded quadric edge collapses; host code (libgmorpho_decimate)."""

    def __init__(self, P, T, target_faces, target_error, use_linear_constraints=True, max_edge_length=1.0,
                 bbox_min=None, res=None, cell=1.0, scale_grid=None, hc_iters=3):
        self._lib = decimate_lib()
        P = np.ascontiguousarray(P, np.float32).reshape(-1, 3)
        T = np.ascontiguousarray(T, np.uint32).reshape(-1, 3)
        bm = None if bbox_min is None else np.ascontiguousarray(bbox_min, np.float32)
        rs = None if res is None else np.ascontiguousarray(res, np.int32)
        sg = None if scale_grid is None else np.ascontiguousarray(scale_grid, np.int8)
        h = _P()
        rc = self._lib.gmd_create(_ptr(P), len(P), _ptr(T), len(T), int(target_faces), ctypes.c_float(target_error),
                                  1 if use_linear_constraints else 0, ctypes.c_float(max_edge_length), _ptr(bm),
                                  _ptr(rs), ctypes.c_float(cell), _ptr(sg), int(hc_iters), ctypes.byref(h))
        if rc:
            raise GMorphoError("gmd_create -> %d" % rc)
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            self._lib.gmd_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def optimize(self, budget=0.0):
        more = ctypes.c_int()
        self._lib.gmd_optimize(self._h, ctypes.c_double(budget), ctypes.byref(more))
        return bool(more.value)

    def state(self):
        nf, nv, err = ctypes.c_int(), ctypes.c_int(), ctypes.c_float()
        self._lib.gmd_state(self._h, ctypes.byref(nf), ctypes.byref(nv), ctypes.byref(err))
        return nf.value, nv.value, err.value

    def mesh(self):
        nf, nv, _ = self.state()
        P = np.zeros((nv, 3), np.float32)
        T = np.zeros((nf, 3), np.uint32)
        self._lib.gmd_get_mesh(self._h, _ptr(P), _ptr(T))
        return P, T


def mc_tables():
    ntri = np.zeros(256, np.uint32)
    ln = np.zeros(256, np.uint32)
    tri = np.zeros(256, np.uint64)
    hip_lib().gm_mc_tables(_ptr(ntri), _ptr(ln), _ptr(tri))
    return ntri, ln, tri


class _Job(ctypes.Structure):
    _fields_ = [("P", _P), ("num_of_vertices", ctypes.c_int), ("T", _P), ("num_of_faces", ctypes.c_int),
                ("base_res", ctypes.c_int), ("se_size", ctypes.c_float), ("radius", ctypes.c_int),
                ("lattice", _P), ("lattice_res", ctypes.c_int), ("se", ctypes.c_int)]


class Engine:
    """gm_engine: a pool of `slots` contexts running closing jobs concurrently
    on one device under an HBM budget (batched and served use)."""

    def __init__(self, device=0, slots=4, hbm_budget=200 << 30):
        self._lib = hip_lib()
        self._lib.gm_engine_last_error.restype = ctypes.c_char_p
        self._lib.gm_engine_last_error.argtypes = [_P]
        h = _P()
        rc = self._lib.gm_engine_create(device, slots, ctypes.c_uint64(hbm_budget), ctypes.byref(h))
        if rc:
            raise GMorphoError("gm_engine_create failed (%d): no usable HIP device" % rc)
        self._h = h
        self._held = {}  # borrowed inputs of jobs not yet waited for

    def close(self):
        if self._h:
            self._lib.gm_engine_destroy(self._h)
            self._h = None
            self._held = {}

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _call(self, name, *args):
        rc = getattr(self._lib, name)(self._h, *args)
        if rc:
            msg = self._lib.gm_engine_last_error(self._h)
            raise GMorphoError("%s -> %d: %s" % (name, rc, msg.decode() if msg else ""))
        return rc

    def submit(self, P, T, base_res, se_size=0.3, radius=0, lattice=None, se=SE_SPHERE, borrow=False):
        """Queue one closing job (voxelize at base_res, close, mesh); returns
        its id. borrow=True hands the arrays to the engine without a copy; this
        object keeps them alive until the job is waited for, viewed, taken or
        released."""
        P = np.ascontiguousarray(P, np.float32)
        T = np.ascontiguousarray(T, np.uint32)
        L = 0
        lat = None
        if lattice is not None:
            lat = np.ascontiguousarray(lattice, np.float32)
            L = int(round(lat.size ** (1.0 / 3.0)))
        job = _Job(P.ctypes.data, len(P), T.ctypes.data, len(T), base_res, se_size, radius,
                   None if lat is None else lat.ctypes.data, L, se)
        jid = ctypes.c_uint64()
        self._call("gm_engine_submit_borrowed" if borrow else "gm_engine_submit", ctypes.byref(job),
                   ctypes.byref(jid))
        if borrow:
            self._held[jid.value] = (P, T, lat)
        return jid.value

    def wait(self, jid):
        nv = ctypes.c_uint32()
        nt = ctypes.c_uint32()
        try:
            self._call("gm_engine_wait", ctypes.c_uint64(jid), ctypes.byref(nv), ctypes.byref(nt))
        finally:
            self._held.pop(jid, None)
        return nv.value, nt.value

    def view(self, jid):
        """Mesh (V, N, T) of a finished job as arrays over the engine's pinned
        result block (no copy); valid until release(jid)."""
        V, N, T = _P(), _P(), _P()
        nv = ctypes.c_uint32()
        nt = ctypes.c_uint32()
        try:
            self._call("gm_engine_view", ctypes.c_uint64(jid), ctypes.byref(V), ctypes.byref(N), ctypes.byref(T),
                       ctypes.byref(nv), ctypes.byref(nt))
        finally:
            self._held.pop(jid, None)

        def arr(p, n, ct, dt):
            if n == 0:
                return np.zeros((0, 3), dt)
            return np.ctypeslib.as_array(ctypes.cast(p, ctypes.POINTER(ct)), shape=(n, 3))

        return (arr(V, nv.value, ctypes.c_float, np.float32), arr(N, nv.value, ctypes.c_float, np.float32),
                arr(T, nt.value, ctypes.c_uint32, np.uint32))

    def release(self, jid):
        """Forget a job; arrays from view(jid) are invalid afterwards."""
        try:
            self._call("gm_engine_release", ctypes.c_uint64(jid))
        finally:
            self._held.pop(jid, None)

    def phases(self, jid):
        """(submit, start, loaded, meshed, done) seconds on the engine clock."""
        t = np.zeros(5, np.float64)
        self._call("gm_engine_job_phases", ctypes.c_uint64(jid), _ptr(t))
        return t

    def times(self, jid):
        """(submit, start, done) seconds on the engine clock."""
        t = np.zeros(3, np.float64)
        self._call("gm_engine_job_times", ctypes.c_uint64(jid), _ptr(t))
        return t

    def now(self):
        t = ctypes.c_double()
        self._call("gm_engine_now", ctypes.byref(t))
        return t.value

    def take(self, jid):
        """Mesh (V, N, T) of a finished job; releases it."""
        nv, nt = self.wait(jid)
        V = np.zeros((nv, 3), np.float32)
        N = np.zeros((nv, 3), np.float32)
        T = np.zeros((nt, 3), np.uint32)
        self._call("gm_engine_take", ctypes.c_uint64(jid), _ptr(V), _ptr(N), _ptr(T))
        self._held.pop(jid, None)
        return V, N, T

    def stats(self):
        b = ctypes.c_uint64()
        r = ctypes.c_int()
        q = ctypes.c_int()
        self._call("gm_engine_stats", ctypes.byref(b), ctypes.byref(r), ctypes.byref(q))
        return b.value, r.value, q.value


class Batch:
    """gm_batch: closing + mesh extraction of many contexts at once, one
    kernel launch per step for the uniform-radius sphere contexts."""

    def __init__(self, device=0):
        self._lib = hip_lib()
        self._lib.gm_batch_last_error.restype = ctypes.c_char_p
        self._lib.gm_batch_last_error.argtypes = [_P]
        h = _P()
        rc = self._lib.gm_batch_create(device, ctypes.byref(h))
        if rc:
            raise GMorphoError("gm_batch_create failed (%d): no usable HIP device" % rc)
        self._h = h

    def close(self):
        if self._h:
            self._lib.gm_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_graphs(self, on=True):
        """Replay the batched launches as a captured HIP graph (default on;
        "repeat": capture when the same batch comes a second time)."""
        rc = self._lib.gm_batch_set_graphs(self._h, 2 if on == "repeat" else (1 if on else 0))
        if rc:
            raise GMorphoError("gm_batch_set_graphs -> %d" % rc)

    def close_mesh(self, contexts, se=SE_SPHERE):
        """gm_close + gm_extract_mesh for every context; meshes are then read
        with each context's mesh_counts / mesh."""
        arr = (_P * len(contexts))(*[c._h for c in contexts])
        rc = self._lib.gm_batch_close_mesh(self._h, arr, len(contexts), se)
        if rc:
            msg = self._lib.gm_batch_last_error(self._h)
            raise GMorphoError("gm_batch_close_mesh -> %d: %s" % (rc, msg.decode() if msg else ""))
