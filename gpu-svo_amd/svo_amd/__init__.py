"""ctypes binding of libsvo_amd.so, the MI355X sparse-voxel-octree builder.

The library is loaded from the package directory (gpu-svo_amd/libsvo_amd.so);
there is no CPU fallback: every build runs the HIP kernels, and importing this
module without the built library raises.

    import svo_amd
    ctx = svo_amd.Context()
    tree = ctx.build(points, tris, res=64)           # reference semantics
    voxels = tree.field("voxels")                    # sorted unique Morton codes
    nodes = tree.field("nodes_aos")                  # 168-byte SVONode records
    tree.write_obj("out", "bunny")                   # out/bunny_<depth>_tree.obj
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
PKG_ROOT = os.path.dirname(HERE)
LIB_PATH = os.environ.get("SVO_AMD_LIB", os.path.join(PKG_ROOT, "libsvo_amd.so"))

OK, E_INVALID, E_RES_LIMIT, E_OOM, E_HIP, E_IO, E_PARSE, E_EMPTY, E_BUSY, E_RANGE = (
    0, -1, -2, -3, -4, -5, -6, -7, -8, -9)
E_TOO_LARGE = -10
SEM_REFERENCE, SEM_INTENDED = 0, 1
BUILD_NO_VERTEX_EDGE = 1

FIELDS = {
    "voxels": 0, "code": 1, "is_leaf": 2, "origin": 3, "width": 4, "parent": 5, "childs": 6,
    "neighbors": 7, "vertices": 8, "edges": 9, "faces": 10, "level_nodes": 11, "nodes_aos": 12,
}
VERT_DTYPE = np.dtype([("p", "<f4", (3,)), ("owner", "<u4")])
EDGE_DTYPE = np.dtype([("a", "<f4", (3,)), ("b", "<f4", (3,)), ("owner", "<u4")])
NODE_DTYPE = np.dtype([
    ("code", "<u4"), ("is_leaf", "u1"), ("pad", "u1", (3,)), ("origin", "<f4", (3,)),
    ("width", "<f4"), ("parent", "<u4"), ("childs", "<u4", (8,)), ("neighbors", "<u4", (27,)),
])


class SvoError(RuntimeError):
    def __init__(self, status: int, what: str = ""):
        self.status = status
        super().__init__(f"{what}: {status_str(status)} ({status})")


class MeshDesc(C.Structure):
    _fields_ = [("xyz", C.c_void_p), ("n_points", C.c_uint64), ("tri", C.c_void_p),
                ("n_tris", C.c_uint64)]


class Params(C.Structure):
    _fields_ = [("res", C.c_int32 * 3), ("semantics", C.c_uint32), ("key_bits", C.c_uint32),
                ("flags", C.c_uint32)]


class TreeInfo(C.Structure):
    _fields_ = [
        ("depth", C.c_int32), ("key_bits", C.c_int32), ("semantics", C.c_int32),
        ("res", C.c_int32 * 3), ("bbox_min", C.c_float * 3), ("bbox_max", C.c_float * 3),
        ("bbox_width", C.c_float * 3), ("unit", C.c_float * 3), ("n_voxels", C.c_uint64),
        ("n_nodes", C.c_uint64), ("n_vertices", C.c_uint64), ("n_edges", C.c_uint64),
        ("n_faces", C.c_uint64), ("n_hits", C.c_uint64), ("device_bytes", C.c_uint64),
        ("level_nodes", C.c_uint64 * 32),
    ]


class ServerCfg(C.Structure):
    _fields_ = [("mem_budget", C.c_uint64), ("max_batch", C.c_uint32), ("max_queue", C.c_uint32),
                ("n_lanes", C.c_uint32), ("max_wait_us", C.c_uint32)]


class Completion(C.Structure):
    _fields_ = [("request_id", C.c_uint64), ("status", C.c_int32), ("batch_size", C.c_uint32),
                ("tree", C.c_void_p), ("t_submit_s", C.c_double), ("t_start_s", C.c_double),
                ("t_done_s", C.c_double)]


class TuneResult(C.Structure):
    _fields_ = [("small_max_cells", C.c_uint32), ("batch_lanes", C.c_uint32),
                ("best_ms", C.c_double), ("default_ms", C.c_double), ("n_trials", C.c_uint32),
                ("class_small_max", C.c_uint32 * 15), ("class_meshes", C.c_uint32 * 15)]


# Every exported entry point and its ctypes signature (kept in step with
# include/svo/svo.h; tests check the two agree).
_SIGS = {
    "svo_ctx_create": (C.c_int, [C.c_int, C.c_uint64, C.POINTER(C.c_void_p)]),
    "svo_ctx_destroy": (C.c_int, [C.c_void_p]),
    "svo_ctx_set_timing": (C.c_int, [C.c_void_p, C.c_int]),
    "svo_ctx_set_option": (C.c_int, [C.c_void_p, C.c_char_p, C.c_int64]),
    "svo_ctx_autotune": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_int,
                                   C.c_uint32, C.POINTER(TuneResult)]),
    "svo_ctx_save_tuning": (C.c_int, [C.c_void_p, C.c_char_p]),
    "svo_ctx_load_tuning": (C.c_int, [C.c_void_p, C.c_char_p]),
    "svo_ctx_get_option": (C.c_int, [C.c_void_p, C.c_char_p, C.POINTER(C.c_int64)]),
    "svo_ctx_last_stage_ms": (C.c_int, [C.c_void_p, C.POINTER(C.c_double),
                                        C.POINTER(C.c_char_p), C.c_int,
                                        C.POINTER(C.c_int)]),
    "svo_ctx_bytes_in_use": (C.c_uint64, [C.c_void_p]),
    "svo_plan_create": (C.c_int, [C.c_void_p, C.POINTER(MeshDesc), C.POINTER(Params), C.c_uint32,
                                  C.c_int, C.POINTER(C.c_void_p)]),
    "svo_plan_run": (C.c_int, [C.c_void_p, C.POINTER(C.c_void_p), C.POINTER(C.c_int),
                               C.POINTER(C.c_int)]),
    "svo_plan_destroy": (C.c_int, [C.c_void_p]),
    "svo_ctx_workspace_bytes": (C.c_uint64, [C.c_void_p]),
    "svo_ctx_device_mallocs": (C.c_uint64, [C.c_void_p]),
    "svo_device_alloc": (C.c_int, [C.c_void_p, C.c_uint64, C.POINTER(C.c_void_p)]),
    "svo_device_free": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64]),
    "svo_device_copy": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64, C.c_int]),
    "svo_load_obj": (C.c_int, [C.c_char_p, C.POINTER(C.c_void_p)]),
    "svo_mesh_view": (C.c_int, [C.c_void_p, C.POINTER(MeshDesc)]),
    "svo_mesh_free": (None, [C.c_void_p]),
    "svo_validate": (C.c_int, [C.POINTER(MeshDesc), C.POINTER(Params)]),
    "svo_build": (C.c_int, [C.c_void_p, C.POINTER(MeshDesc), C.POINTER(Params),
                            C.POINTER(C.c_void_p)]),
    "svo_build_device": (C.c_int, [C.c_void_p, C.POINTER(MeshDesc), C.POINTER(Params),
                                   C.POINTER(C.c_void_p)]),
    "svo_build_batch": (C.c_int, [C.c_void_p, C.POINTER(MeshDesc), C.POINTER(Params),
                                  C.c_uint32, C.c_int, C.POINTER(C.c_void_p),
                                  C.POINTER(C.c_int)]),
    "svo_tree_info_get": (C.c_int, [C.c_void_p, C.POINTER(TreeInfo)]),
    "svo_tree_field_bytes": (C.c_uint64, [C.c_void_p, C.c_int]),
    "svo_tree_copy_field": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_uint64]),
    "svo_tree_device_ptr": (C.c_void_p, [C.c_void_p, C.c_int]),
    "svo_tree_copy_voxels": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64]),
    "svo_tree_copy_nodes": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64]),
    "svo_tree_copy_neighbors": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64]),
    "svo_tree_copy_vertices": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64]),
    "svo_tree_copy_edges": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64]),
    "svo_tree_copy_faces": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64]),
    "svo_write_obj": (C.c_int, [C.c_void_p, C.c_char_p, C.c_char_p, C.c_char_p, C.c_size_t]),
    "svo_write_voxels_obj": (C.c_int, [C.c_void_p, C.c_char_p, C.c_char_p, C.c_char_p,
                                       C.c_size_t]),
    "svo_query_neighbors": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64,
                                      C.c_void_p]),
    "svo_tree_free": (None, [C.c_void_p]),
    "svo_status_str": (C.c_char_p, [C.c_int]),
    "svo_server_create": (C.c_int, [C.c_void_p, C.POINTER(ServerCfg), C.POINTER(C.c_void_p)]),
    "svo_server_submit": (C.c_int, [C.c_void_p, C.POINTER(MeshDesc), C.POINTER(Params),
                                    C.POINTER(C.c_uint64)]),
    "svo_server_poll": (C.c_int, [C.c_void_p, C.POINTER(Completion), C.c_uint32,
                                  C.POINTER(C.c_uint32), C.c_int64]),
    "svo_estimate_bytes": (C.c_uint64, [C.c_uint64, C.c_uint64, C.POINTER(Params)]),
    "svo_estimate_bytes_mesh": (C.c_uint64, [C.POINTER(MeshDesc), C.POINTER(Params)]),
    "svo_server_stats": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint64), C.POINTER(C.c_uint32),
                                   C.POINTER(C.c_uint32)]),
    "svo_server_destroy": (C.c_int, [C.c_void_p]),
    # include/svo/svo_kernels.h
    "svo_k_radix_sort": (C.c_int, [C.c_int, C.c_void_p, C.c_uint64, C.c_int, C.c_int]),
    "svo_k_radix_sort_noalloc": (C.c_int, [C.c_int, C.c_void_p, C.c_uint64, C.c_int, C.c_int,
                                           C.c_void_p]),
    "svo_k_ref_vertex_edge": (C.c_int, [C.c_int, C.c_void_p, C.c_void_p, C.c_uint64,
                                        C.c_void_p, C.c_void_p]),
    "svo_k_octree_level": (C.c_int, [C.c_int, C.c_void_p, C.c_uint64, C.c_int, C.c_int,
                                     C.c_void_p, C.POINTER(C.c_uint64), C.c_void_p, C.c_void_p,
                                     C.POINTER(C.c_uint64)]),
    "svo_k_voxelize": (C.c_int, [C.c_int, C.c_void_p, C.c_uint64, C.c_void_p, C.c_uint64,
                                 C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint64,
                                 C.POINTER(C.c_uint64)]),
}

_lib = None


def lib():
    """Loads libsvo_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"libsvo_amd.so not built: {LIB_PATH} (run make -C gpu-svo_amd)")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def status_str(st: int) -> str:
    return lib().svo_status_str(st).decode()


def _check(st: int, what: str):
    if st != OK:
        raise SvoError(st, what)


def _res3(res):
    if np.isscalar(res):
        return (int(res),) * 3
    r = tuple(int(x) for x in res)
    assert len(r) == 3
    return r


def make_params(res, semantics=SEM_REFERENCE, key_bits=0, flags=0) -> Params:
    p = Params()
    p.res[:] = _res3(res)
    p.semantics = semantics
    p.key_bits = key_bits
    p.flags = flags
    return p


def _host_desc(points, tris):
    pts = np.ascontiguousarray(points, dtype=np.float32).reshape(-1, 3)
    tri = np.ascontiguousarray(tris, dtype=np.int32).reshape(-1, 3)
    d = MeshDesc(pts.ctypes.data, pts.shape[0], tri.ctypes.data, tri.shape[0])
    return d, (pts, tri)


def load_obj(path: str):
    """OBJ parse with the reference's rules: (points float32[P,3], tris int32[T,3])."""
    L = lib()
    h = C.c_void_p()
    _check(L.svo_load_obj(path.encode(), C.byref(h)), f"loading {path}")
    try:
        d = MeshDesc()
        L.svo_mesh_view(h, C.byref(d))
        pts = np.zeros(0, np.float32)
        tri = np.zeros(0, np.int32)
        if d.n_points:
            pts = np.ctypeslib.as_array(C.cast(d.xyz, C.POINTER(C.c_float)),
                                        shape=(3 * d.n_points,)).copy()
        if d.n_tris:
            tri = np.ctypeslib.as_array(C.cast(d.tri, C.POINTER(C.c_int32)),
                                        shape=(3 * d.n_tris,)).copy()
        return pts.reshape(-1, 3), tri.reshape(-1, 3)
    finally:
        L.svo_mesh_free(h)


def load_obj_status(path: str) -> int:
    L = lib()
    h = C.c_void_p()
    st = L.svo_load_obj(path.encode(), C.byref(h))
    if st == OK:
        L.svo_mesh_free(h)
    return st


def validate(points, tris, res, semantics=SEM_REFERENCE, key_bits=0) -> int:
    d, _keep = _host_desc(points, tris)
    p = make_params(res, semantics, key_bits)
    return lib().svo_validate(C.byref(d), C.byref(p))


def estimate_bytes(n_points, n_tris, res, semantics=SEM_REFERENCE, flags=0) -> int:
    p = make_params(res, semantics, 0, flags)
    return int(lib().svo_estimate_bytes(n_points, n_tris, C.byref(p)))


def estimate_bytes_mesh(points, tris, res, semantics=SEM_REFERENCE, key_bits=0, flags=0) -> int:
    """Upper bound on the device bytes one build of the mesh takes."""
    d, _keep = _host_desc(points, tris)
    p = make_params(res, semantics, key_bits, flags)
    return int(lib().svo_estimate_bytes_mesh(C.byref(d), C.byref(p)))


_FIELD_DT = {
    "is_leaf": (np.uint8, None), "origin": (np.float32, 3), "width": (np.float32, None),
    "parent": (np.uint32, None), "childs": (np.uint32, 8), "neighbors": (np.uint32, 27),
    "vertices": (VERT_DTYPE, None), "edges": (EDGE_DTYPE, None), "faces": (EDGE_DTYPE, None),
    "level_nodes": (np.uint64, None), "nodes_aos": (NODE_DTYPE, None),
}


def k_radix_sort(keys: np.ndarray, bits: int | None = None, device: int = 0) -> np.ndarray:
    """Sorts uint32 / uint64 keys with the library's LSD radix sort (a copy)."""
    a = np.ascontiguousarray(keys).copy()
    assert a.dtype in (np.uint32, np.uint64)
    nb = 8 * a.dtype.itemsize if bits is None else bits
    _check(lib().svo_k_radix_sort(device, a.ctypes.data, a.size, a.dtype.itemsize, nb),
           "radix sort")
    return a


def k_ref_vertex_edge(origin: np.ndarray, width: np.ndarray, device: int = 0):
    """Reference-semantics vertex and edge records of nodes
    (svo_k_ref_vertex_edge): (vertices [8n], edges [12n]) structured arrays."""
    o = np.ascontiguousarray(origin, np.float32).reshape(-1, 3)
    w = np.ascontiguousarray(width, np.float32).reshape(-1)
    assert o.shape[0] == w.size
    n = w.size
    verts = np.empty(8 * n, VERT_DTYPE)
    edges = np.empty(12 * n, EDGE_DTYPE)
    _check(lib().svo_k_ref_vertex_edge(device, o.ctypes.data if n else None,
                                       w.ctypes.data if n else None, n,
                                       verts.ctypes.data if n else None,
                                       edges.ctypes.data if n else None), "vertex/edge records")
    return verts, edges


def k_octree_level(sorted_codes: np.ndarray, code_bits: int, device: int = 0):
    """One level step on sorted codes (svo_k_octree_level): (unique codes,
    parent codes, index of each parent group's first member in unique)."""
    a = np.ascontiguousarray(sorted_codes)
    assert a.dtype in (np.uint32, np.uint64)
    n = a.size
    uniq = np.empty(max(n, 1), a.dtype)
    par = np.empty(max(n, 1), a.dtype)
    fc = np.empty(max(n, 1), np.uint32)
    nu, npar = C.c_uint64(), C.c_uint64()
    _check(lib().svo_k_octree_level(device, a.ctypes.data if n else None, n, a.dtype.itemsize,
                                    code_bits, uniq.ctypes.data, C.byref(nu), par.ctypes.data,
                                    fc.ctypes.data, C.byref(npar)), "octree level")
    return uniq[:nu.value].copy(), par[:npar.value].copy(), fc[:npar.value].copy()


def k_voxelize(points, tris, res, key_bits=32, device: int = 0):
    """Raw voxel hits of one mesh (emission order, duplicates kept): (codes, n_hits)."""
    d, (pts, tri) = _host_desc(points, tris)
    r = np.asarray(_res3(res), dtype=np.int32)
    n = C.c_uint64()
    _check(lib().svo_k_voxelize(device, pts.ctypes.data, pts.shape[0], tri.ctypes.data,
                                tri.shape[0], r.ctypes.data, key_bits, None, 0, C.byref(n)),
           "voxelize")
    out = np.empty(n.value, dtype=np.uint64 if key_bits == 64 else np.uint32)
    _check(lib().svo_k_voxelize(device, pts.ctypes.data, pts.shape[0], tri.ctypes.data,
                                tri.shape[0], r.ctypes.data, key_bits, out.ctypes.data,
                                out.size, C.byref(n)), "voxelize")
    return out, n.value


class Tree:
    """A built octree resident in device memory."""

    def __init__(self, handle, ctx: "Context"):
        self._h = C.c_void_p(handle)
        self._ctx = ctx
        info = TreeInfo()
        _check(lib().svo_tree_info_get(self._h, C.byref(info)), "tree info")
        self.info = info

    @property
    def depth(self) -> int:
        return self.info.depth

    @property
    def handle(self):
        return self._h

    def field(self, name: str) -> np.ndarray:
        fid = FIELDS[name]
        L = lib()
        nbytes = int(L.svo_tree_field_bytes(self._h, fid))
        if name in ("voxels", "code"):
            dt, width = (np.uint64 if self.info.key_bits == 64 else np.uint32), None
        else:
            dt, width = _FIELD_DT[name]
        dt = np.dtype(dt)
        a = np.empty(nbytes // dt.itemsize, dtype=dt)
        if nbytes:
            _check(L.svo_tree_copy_field(self._h, fid, a.ctypes.data, nbytes), f"copy {name}")
        return a.reshape(-1, width) if width else a

    def device_ptr(self, name: str) -> int:
        return lib().svo_tree_device_ptr(self._h, FIELDS[name]) or 0

    def field_rows(self, name: str, start: int, count: int) -> np.ndarray:
        """Rows [start, start + count) of a node or record field (a slice of
        what field() returns), copied straight from device memory."""
        if name in ("voxels", "code"):
            dt, width = (np.uint64 if self.info.key_bits == 64 else np.uint32), None
        else:
            dt, width = _FIELD_DT[name]
            if name in ("level_nodes", "nodes_aos"):
                raise ValueError(f"{name} has no device rows")
        dt = np.dtype(dt)
        row = dt.itemsize * (width or 1)
        L = lib()
        nbytes = int(L.svo_tree_field_bytes(self._h, FIELDS[name]))
        if name in ("voxels", "code"):
            n = self.info.n_voxels if name == "voxels" else self.info.n_nodes
            if nbytes != n * row:
                raise ValueError(f"{name}: stored width differs from key_bits")
        total = nbytes // row
        start = max(0, int(start))
        count = max(0, min(int(count), total - start))
        a = np.empty(count * (width or 1), dtype=dt)
        if count:
            src = self.device_ptr(name) + start * row
            _check(L.svo_device_copy(self._ctx.handle, a.ctypes.data, C.c_void_p(src),
                                     count * row, 1), f"copy rows of {name}")
        return a.reshape(-1, width) if width else a

    def write_obj(self, directory: str, stem: str) -> str:
        buf = C.create_string_buffer(4096)
        _check(lib().svo_write_obj(self._h, directory.encode(), stem.encode(), buf, 4096),
               "write obj")
        return buf.value.decode()

    def write_voxels_obj(self, directory: str, stem: str) -> str:
        """writeVoxel: <directory>/<stem>_1_voxel.obj, one cube per occupied voxel."""
        buf = C.create_string_buffer(4096)
        _check(lib().svo_write_voxels_obj(self._h, directory.encode(), stem.encode(), buf, 4096),
               "write voxel obj")
        return buf.value.decode()

    def query_neighbors(self, nodes, slots) -> np.ndarray:
        n = np.ascontiguousarray(nodes, dtype=np.uint32)
        s = np.ascontiguousarray(slots, dtype=np.uint8)
        out = np.empty(n.shape[0], np.uint32)
        _check(lib().svo_query_neighbors(self._h, n.ctypes.data, s.ctypes.data, n.shape[0],
                                         out.ctypes.data), "query neighbors")
        return out

    def free(self):
        if self._h is not None and self._h.value:
            lib().svo_tree_free(self._h)
        self._h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Context:
    """One device; builds run on its pipeline lanes."""

    def __init__(self, device: int = 0, mem_budget: int = 0):
        L = lib()
        h = C.c_void_p()
        _check(L.svo_ctx_create(device, mem_budget, C.byref(h)), "context")
        self._h = h

    @property
    def handle(self):
        return self._h

    def close(self):
        if self._h is not None and self._h.value:
            lib().svo_ctx_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_timing(self, on: bool = True):
        lib().svo_ctx_set_timing(self._h, 1 if on else 0)

    def set_option(self, name: str, value: int):
        """Tuning option (svo_ctx_set_option); set between builds."""
        _check(lib().svo_ctx_set_option(self._h, name.encode(), int(value)), f"option {name}")

    def get_option(self, name: str) -> int:
        v = C.c_int64()
        _check(lib().svo_ctx_get_option(self._h, name.encode(), C.byref(v)), f"option {name}")
        return v.value

    def stage_ms(self) -> dict:
        n = C.c_int()
        ms = (C.c_double * 64)()
        names = (C.c_char_p * 64)()
        lib().svo_ctx_last_stage_ms(self._h, ms, names, 64, C.byref(n))
        return {names[i].decode(): ms[i] for i in range(min(n.value, 64))}

    def bytes_in_use(self) -> int:
        return int(lib().svo_ctx_bytes_in_use(self._h))

    def workspace_bytes(self) -> int:
        """The part of bytes_in_use() held by the lanes' workspaces."""
        return int(lib().svo_ctx_workspace_bytes(self._h))

    def device_mallocs(self) -> int:
        """hipMalloc calls the context has made (its block cache's misses)."""
        return int(lib().svo_ctx_device_mallocs(self._h))

    def upload_mesh(self, points, tris) -> "DeviceMesh":
        """Copies a mesh to device memory (kept until DeviceMesh.free)."""
        return DeviceMesh(self, points, tris)

    def build(self, points, tris, res, semantics=SEM_REFERENCE, key_bits=0, flags=0) -> Tree:
        d, _keep = _host_desc(points, tris)
        p = make_params(res, semantics, key_bits, flags)
        h = C.c_void_p()
        _check(lib().svo_build(self._h, C.byref(d), C.byref(p), C.byref(h)), "build")
        return Tree(h.value, self)

    def build_device(self, points_ptr: int, n_points: int, tris_ptr: int, n_tris: int, res,
                     semantics=SEM_REFERENCE, key_bits=0, flags=0) -> Tree:
        d = MeshDesc(points_ptr, n_points, tris_ptr, n_tris)
        p = make_params(res, semantics, key_bits, flags)
        h = C.c_void_p()
        _check(lib().svo_build_device(self._h, C.byref(d), C.byref(p), C.byref(h)),
               "build_device")
        return Tree(h.value, self)

    @staticmethod
    def _batch_args(meshes, semantics, key_bits, flags, on_device):
        n = len(meshes)
        descs = (MeshDesc * n)()
        params = (Params * n)()
        keep = []
        for i, m in enumerate(meshes):
            if on_device:
                pp, npnt, tp, ntri, res = m
                descs[i] = MeshDesc(pp, npnt, tp, ntri)
            else:
                pts, tri, res = m
                d, k = _host_desc(pts, tri)
                keep.append(k)
                descs[i] = d
            sem = semantics[i] if isinstance(semantics, (list, tuple)) else semantics
            fl = flags[i] if isinstance(flags, (list, tuple)) else flags
            params[i] = make_params(res, sem, key_bits, fl)
        return descs, params, keep

    def build_batch(self, meshes, semantics=SEM_REFERENCE, key_bits=0, flags=0,
                    on_device=False):
        """meshes: list of (points, tris, res) (host arrays) or, with on_device,
        (points_ptr, n_points, tris_ptr, n_tris, res); semantics and flags are
        one value or one per mesh.  Returns (trees, statuses)."""
        n = len(meshes)
        descs, params, keep = self._batch_args(meshes, semantics, key_bits, flags, on_device)
        out = (C.c_void_p * n)()
        st = (C.c_int * n)()
        lib().svo_build_batch(self._h, descs, params, n, 1 if on_device else 0, out, st)
        trees = [Tree(out[i], self) if st[i] == OK and out[i] else None for i in range(n)]
        return trees, [st[i] for i in range(n)]

    def plan(self, meshes, semantics=SEM_REFERENCE, key_bits=0, flags=0,
             on_device=False) -> "BatchPlan":
        """A batch recorded once and replayed as one HIP graph (svo_plan_*);
        meshes as for build_batch."""
        return BatchPlan(self, meshes, semantics, key_bits, flags, on_device)

    def autotune(self, meshes, semantics=SEM_REFERENCE, key_bits=0, flags=0, on_device=False,
                 reps=3) -> dict:
        """svo_ctx_autotune on a sample batch (same forms as build_batch); the
        fastest options stay set on the context."""
        n = len(meshes)
        descs, params, keep = self._batch_args(meshes, semantics, key_bits, flags, on_device)
        r = TuneResult()
        _check(lib().svo_ctx_autotune(self._h, descs, params, n, 1 if on_device else 0, reps,
                                      C.byref(r)), "autotune")
        out = {k: getattr(r, k) for k, _ in TuneResult._fields_[:5]}
        out["classes"] = {k: {"meshes": r.class_meshes[k], "small_max_cells": r.class_small_max[k]}
                          for k in range(15) if r.class_meshes[k]}
        return out

    def save_tuning(self, path: str):
        _check(lib().svo_ctx_save_tuning(self._h, path.encode()), "save tuning")

    def load_tuning(self, path: str):
        _check(lib().svo_ctx_load_tuning(self._h, path.encode()), "load tuning")


class BatchPlan:
    """svo_plan_create / svo_plan_run / svo_plan_destroy.  run() returns
    (trees, statuses, replayed); the trees stay valid until the next run()."""

    def __init__(self, ctx: Context, meshes, semantics, key_bits, flags, on_device):
        self.ctx = ctx
        self.n = len(meshes)
        descs, params, keep = ctx._batch_args(meshes, semantics, key_bits, flags, on_device)
        h = C.c_void_p()
        _check(lib().svo_plan_create(ctx.handle, descs, params, self.n, 1 if on_device else 0,
                                     C.byref(h)), "plan")
        self._h = h

    def run(self):
        out = (C.c_void_p * self.n)()
        st = (C.c_int * self.n)()
        rp = C.c_int()
        lib().svo_plan_run(self._h, out, st, C.byref(rp))
        trees = [Tree(out[i], self.ctx) if st[i] == OK and out[i] else None
                 for i in range(self.n)]
        return trees, [st[i] for i in range(self.n)], bool(rp.value)

    def close(self):
        if self._h is not None and self._h.value:
            lib().svo_plan_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceMesh:
    """A mesh resident in device memory of a context."""

    def __init__(self, ctx: Context, points, tris):
        L = lib()
        pts = np.ascontiguousarray(points, dtype=np.float32).reshape(-1, 3)
        tri = np.ascontiguousarray(tris, dtype=np.int32).reshape(-1, 3)
        self.ctx = ctx
        self.n_points, self.n_tris = pts.shape[0], tri.shape[0]
        self._bytes = (pts.nbytes, tri.nbytes)
        p, t = C.c_void_p(), C.c_void_p()
        _check(L.svo_device_alloc(ctx.handle, max(1, pts.nbytes), C.byref(p)), "device alloc")
        _check(L.svo_device_alloc(ctx.handle, max(1, tri.nbytes), C.byref(t)), "device alloc")
        self.points_ptr, self.tris_ptr = p.value, t.value
        _check(L.svo_device_copy(ctx.handle, p, pts.ctypes.data, pts.nbytes, 0), "upload")
        _check(L.svo_device_copy(ctx.handle, t, tri.ctypes.data, tri.nbytes, 0), "upload")

    def desc(self, res):
        return (self.points_ptr, self.n_points, self.tris_ptr, self.n_tris, res)

    def free(self):
        if self.points_ptr:
            lib().svo_device_free(self.ctx.handle, C.c_void_p(self.points_ptr), max(1, self._bytes[0]))
            lib().svo_device_free(self.ctx.handle, C.c_void_p(self.tris_ptr), max(1, self._bytes[1]))
        self.points_ptr = self.tris_ptr = 0


@dataclass
class Done:
    request_id: int
    status: int
    batch_size: int
    tree: Tree | None
    t_submit: float
    t_start: float
    t_done: float


class Server:
    """Continuous batching under a device-memory budget (svo_server_*)."""

    def __init__(self, ctx: Context, mem_budget=0, max_batch=64, max_queue=1 << 16, n_lanes=2,
                 max_wait_us=500):
        cfg = ServerCfg(mem_budget, max_batch, max_queue, n_lanes, max_wait_us)
        h = C.c_void_p()
        _check(lib().svo_server_create(ctx.handle, C.byref(cfg), C.byref(h)), "server")
        self._h = h
        self._ctx = ctx

    def submit(self, points, tris, res, semantics=SEM_REFERENCE, key_bits=0, flags=0) -> int:
        d, _keep = _host_desc(points, tris)
        p = make_params(res, semantics, key_bits, flags)
        rid = C.c_uint64()
        _check(lib().svo_server_submit(self._h, C.byref(d), C.byref(p), C.byref(rid)), "submit")
        return rid.value

    def poll(self, cap=256, timeout_us=0):
        buf = (Completion * cap)()
        n = C.c_uint32()
        _check(lib().svo_server_poll(self._h, buf, cap, C.byref(n), timeout_us), "poll")
        out = []
        for i in range(n.value):
            c = buf[i]
            t = Tree(c.tree, self._ctx) if c.tree else None
            out.append(Done(c.request_id, c.status, c.batch_size, t, c.t_submit_s, c.t_start_s,
                            c.t_done_s))
        return out

    def stats(self):
        b = C.c_uint64()
        p = C.c_uint32()
        f = C.c_uint32()
        lib().svo_server_stats(self._h, C.byref(b), C.byref(p), C.byref(f))
        return {"bytes": b.value, "pending": p.value, "in_flight": f.value}

    def close(self):
        if self._h is not None and self._h.value:
            lib().svo_server_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
