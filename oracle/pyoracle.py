"""ctypes wrapper of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and by
bench.py's cpu_baseline leg, as the checker or as the reported CPU baseline.
The product library (gpu-svo_amd/) never imports this module.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

SEM_REFERENCE = 0
SEM_INTENDED = 1

F_VOXELS, F_CODE, F_ISLEAF, F_ORIGIN, F_WIDTH, F_PARENT, F_CHILDS, F_NEIGHBORS = range(8)
F_VERTICES, F_EDGES, F_FACES, F_LEVEL_NODES, F_NODES_AOS = 8, 9, 10, 11, 12

VERT_DTYPE = np.dtype([("p", "<f4", (3,)), ("owner", "<u4")])
EDGE_DTYPE = np.dtype([("a", "<f4", (3,)), ("b", "<f4", (3,)), ("owner", "<u4")])
NODE_DTYPE = np.dtype([
    ("code", "<u4"), ("is_leaf", "u1"), ("pad", "u1", (3,)), ("origin", "<f4", (3,)),
    ("width", "<f4"), ("parent", "<u4"), ("childs", "<u4", (8,)), ("neighbors", "<u4", (27,)),
])
assert VERT_DTYPE.itemsize == 16 and EDGE_DTYPE.itemsize == 28 and NODE_DTYPE.itemsize == 168


class OracleInfo(C.Structure):
    _fields_ = [
        ("status", C.c_int32), ("depth", C.c_int32), ("key_bits", C.c_int32),
        ("semantics", C.c_int32), ("res", C.c_int32 * 3), ("pad0", C.c_int32),
        ("bbox_min", C.c_float * 3), ("bbox_max", C.c_float * 3),
        ("bbox_width", C.c_float * 3), ("unit", C.c_float * 3),
        ("n_voxels", C.c_uint64), ("n_nodes", C.c_uint64), ("n_vertices", C.c_uint64),
        ("n_edges", C.c_uint64), ("n_faces", C.c_uint64), ("level_nodes", C.c_uint64 * 32),
        ("build_seconds", C.c_double),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing: {LIB_PATH} (run make -C oracle)")
        L = C.CDLL(LIB_PATH)
        P = C.c_void_p
        L.oracle_load_obj.argtypes = [C.c_char_p, C.POINTER(C.POINTER(C.c_float)),
                                      C.POINTER(C.c_uint64), C.POINTER(C.POINTER(C.c_int32)),
                                      C.POINTER(C.c_uint64)]
        L.oracle_load_obj.restype = C.c_int
        L.oracle_free_buf.argtypes = [P]
        L.oracle_build.argtypes = [P, C.c_uint64, P, C.c_uint64, P, C.c_int32, C.c_int32,
                                   C.c_int32, C.POINTER(C.c_int32)]
        L.oracle_build.restype = P
        L.oracle_tree_free.argtypes = [P]
        L.oracle_tree_info.argtypes = [P, C.POINTER(OracleInfo)]
        L.oracle_tree_field_bytes.argtypes = [P, C.c_int32]
        L.oracle_tree_field_bytes.restype = C.c_uint64
        L.oracle_tree_get_field.argtypes = [P, C.c_int32, P, C.c_uint64]
        L.oracle_write_obj.argtypes = [P, C.c_char_p]
        L.oracle_write_voxel_obj.argtypes = [P, C.c_char_p]
        L.oracle_morton_encode_u16.argtypes = [C.c_uint32] * 3
        L.oracle_morton_encode_u16.restype = C.c_uint32
        L.oracle_morton_encode_u64.argtypes = [C.c_uint32] * 3
        L.oracle_morton_encode_u64.restype = C.c_uint64
        L.oracle_morton_decode.argtypes = [C.c_uint64, C.POINTER(C.c_uint32),
                                           C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]
        L.oracle_luts.argtypes = [P] * 10
        L.oracle_voxelize.argtypes = [P, C.c_uint64, P, C.c_uint64, P, C.c_int32, P, C.c_uint64,
                                      C.POINTER(C.c_int32)]
        L.oracle_voxelize.restype = C.c_uint64
        L.oracle_set_arith_variant.argtypes = [C.c_int32]
        L.oracle_voxelize_alloc.argtypes = [P, C.c_uint64, P, C.c_uint64, P, C.c_int32,
                                            C.c_int32, C.POINTER(C.c_uint64),
                                            C.POINTER(C.c_int32)]
        L.oracle_voxelize_alloc.restype = C.POINTER(C.c_uint64)
        _lib = L
    return _lib


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def load_obj(path: str):
    """BaseModel::loadOBJ rules; returns (status, points float32[P,3], tris int32[T,3])."""
    L = lib()
    pp = C.POINTER(C.c_float)()
    tp = C.POINTER(C.c_int32)()
    npnt = C.c_uint64()
    ntri = C.c_uint64()
    st = L.oracle_load_obj(path.encode(), C.byref(pp), C.byref(npnt), C.byref(tp), C.byref(ntri))
    if st != 0:
        return st, None, None
    pts = np.ctypeslib.as_array(pp, shape=(max(1, npnt.value * 3),))[: npnt.value * 3].copy()
    tri = np.ctypeslib.as_array(tp, shape=(max(1, ntri.value * 3),))[: ntri.value * 3].copy()
    L.oracle_free_buf(C.cast(pp, C.c_void_p))
    L.oracle_free_buf(C.cast(tp, C.c_void_p))
    return 0, pts.reshape(-1, 3), tri.reshape(-1, 3)


@dataclass
class OracleTree:
    status: int
    info: OracleInfo | None = None
    arrays: dict = field(default_factory=dict)

    @property
    def depth(self):
        return self.info.depth

    def __getitem__(self, k):
        return self.arrays[k]


def _as_res(res):
    if np.isscalar(res):
        res = (int(res),) * 3
    return np.asarray(res, dtype=np.int32)


def build(points, tris, res, semantics=SEM_REFERENCE, key_bits=32, threads=1, keep_handle=False,
          fields=None):
    """Run the oracle build and return every output array as numpy."""
    L = lib()
    pts = np.ascontiguousarray(points, dtype=np.float32).reshape(-1, 3)
    tri = np.ascontiguousarray(tris, dtype=np.int32).reshape(-1, 3)
    r = _as_res(res)
    st = C.c_int32()
    h = L.oracle_build(_ptr(pts), pts.shape[0], _ptr(tri), tri.shape[0], _ptr(r), semantics,
                       key_bits, threads, C.byref(st))
    try:
        info = OracleInfo()
        L.oracle_tree_info(h, C.byref(info))
        out = OracleTree(status=st.value, info=info)
        if st.value != 0:
            return out
        kdt = np.uint64 if key_bits == 64 else np.uint32
        specs = {
            "voxels": (F_VOXELS, kdt, None), "code": (F_CODE, kdt, None),
            "is_leaf": (F_ISLEAF, np.uint8, None), "origin": (F_ORIGIN, np.float32, 3),
            "width": (F_WIDTH, np.float32, None), "parent": (F_PARENT, np.uint32, None),
            "childs": (F_CHILDS, np.uint32, 8), "neighbors": (F_NEIGHBORS, np.uint32, 27),
            "vertices": (F_VERTICES, VERT_DTYPE, None), "edges": (F_EDGES, EDGE_DTYPE, None),
            "faces": (F_FACES, EDGE_DTYPE, None), "level_nodes": (F_LEVEL_NODES, np.uint64, None),
        }
        if key_bits == 32:
            specs["nodes_aos"] = (F_NODES_AOS, NODE_DTYPE, None)
        for name, (fid, dt, width) in specs.items():
            if fields is not None and name not in fields:
                continue
            nbytes = L.oracle_tree_field_bytes(h, fid)
            dt = np.dtype(dt)
            a = np.empty(nbytes // dt.itemsize, dtype=dt)
            if nbytes:
                L.oracle_tree_get_field(h, fid, _ptr(a), nbytes)
            if width:
                a = a.reshape(-1, width)
            out.arrays[name] = a
        if keep_handle:
            out.arrays["_handle"] = h
            h = None
        return out
    finally:
        if h is not None:
            L.oracle_tree_free(h)


def write_obj(points, tris, res, path, semantics=SEM_REFERENCE, key_bits=32, voxels=False):
    """writeTree (or writeVoxel with voxels=True) to `path`; returns (status, depth)."""
    L = lib()
    pts = np.ascontiguousarray(points, dtype=np.float32).reshape(-1, 3)
    tri = np.ascontiguousarray(tris, dtype=np.int32).reshape(-1, 3)
    r = _as_res(res)
    st = C.c_int32()
    h = L.oracle_build(_ptr(pts), pts.shape[0], _ptr(tri), tri.shape[0], _ptr(r), semantics,
                       key_bits, 1, C.byref(st))
    try:
        if st.value != 0:
            return st.value, 0
        info = OracleInfo()
        L.oracle_tree_info(h, C.byref(info))
        w = L.oracle_write_voxel_obj if voxels else L.oracle_write_obj
        return w(h, path.encode()), info.depth
    finally:
        L.oracle_tree_free(h)


def voxelize(points, tris, res, key_bits=32, threads=None):
    """Sorted unique level-1 codes (uint64); threads=None runs the
    single-threaded count-then-copy form, an int the OpenMP form."""
    L = lib()
    if threads is not None:
        pts = np.ascontiguousarray(points, dtype=np.float32).reshape(-1, 3)
        tri = np.ascontiguousarray(tris, dtype=np.int32).reshape(-1, 3)
        r = _as_res(res)
        st = C.c_int32()
        n = C.c_uint64()
        p = L.oracle_voxelize_alloc(_ptr(pts), pts.shape[0], _ptr(tri), tri.shape[0], _ptr(r),
                                    key_bits, int(threads), C.byref(n), C.byref(st))
        if not p:
            return st.value, np.zeros(0, np.uint64)
        try:
            out = np.ctypeslib.as_array(p, shape=(max(1, n.value),))[: n.value].copy()
        finally:
            L.oracle_free_buf(C.cast(p, C.c_void_p))
        return st.value, out
    pts = np.ascontiguousarray(points, dtype=np.float32).reshape(-1, 3)
    tri = np.ascontiguousarray(tris, dtype=np.int32).reshape(-1, 3)
    r = _as_res(res)
    st = C.c_int32()
    n = L.oracle_voxelize(_ptr(pts), pts.shape[0], _ptr(tri), tri.shape[0], _ptr(r), key_bits,
                          None, 0, C.byref(st))
    out = np.empty(n, dtype=np.uint64)
    if n:
        L.oracle_voxelize(_ptr(pts), pts.shape[0], _ptr(tri), tri.shape[0], _ptr(r), key_bits,
                          _ptr(out), n, C.byref(st))
    return st.value, out


def set_arith_variant(variant: int):
    """0 = pinned restatement; bit 0 left-to-right 3-term dots; bit 1 FMA contraction."""
    lib().oracle_set_arith_variant(int(variant))


def morton_encode_u16(x, y, z):
    return lib().oracle_morton_encode_u16(x, y, z)


def morton_encode_u64(x, y, z):
    return lib().oracle_morton_encode_u64(x, y, z)


def morton_decode(code):
    x, y, z = C.c_uint32(), C.c_uint32(), C.c_uint32()
    lib().oracle_morton_decode(code, C.byref(x), C.byref(y), C.byref(z))
    return x.value, y.value, z.value


def luts():
    ex, ey, ez = (np.zeros(256, np.uint32) for _ in range(3))
    dx, dy, dz = (np.zeros(512, np.uint8) for _ in range(3))
    nbp = np.zeros((8, 27), np.int16)
    nbc = np.zeros((8, 27), np.int16)
    vs = np.zeros((8, 8), np.int16)
    es = np.zeros((12, 4), np.int16)
    lib().oracle_luts(*(_ptr(a) for a in (ex, ey, ez, dx, dy, dz, nbp, nbc, vs, es)))
    return dict(enc_x=ex, enc_y=ey, enc_z=ez, dec_x=dx, dec_y=dy, dec_z=dz, nb_parent=nbp,
                nb_child=nbc, vert_shared=vs, edge_shared=es)
