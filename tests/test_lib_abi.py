"""The C ABI: every declared entry point is exported, the ctypes signatures
and struct layouts match the headers, and the host-only checks return the
documented statuses without a device."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

import svo_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = [os.path.join(ROOT, "include", "svo", h) for h in ("svo.h", "svo_kernels.h")]


def _declared():
    names = set()
    for h in HEADERS:
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"\b(svo_[a-z0-9_]+)\s*\(", text))
    return names


def test_every_declared_symbol_is_exported():
    lib = svo_amd.lib()
    declared = _declared()
    assert len(declared) >= 30
    for name in sorted(declared):
        assert hasattr(lib, name), name


def test_bindings_cover_the_headers():
    declared = _declared()
    missing = declared - set(svo_amd._SIGS) - {"svo_ctx_last_pass_stats"}
    assert not missing, missing


def test_struct_layouts_match_c(tmp_path):
    src = tmp_path / "sizes.c"
    src.write_text(
        '#include <stdio.h>\n#include <stddef.h>\n#include "svo/svo.h"\n'
        "int main(void){printf(\"%zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu\\n\","
        "sizeof(svo_mesh_desc), sizeof(svo_params), sizeof(svo_tree_info),"
        "sizeof(svo_server_cfg), sizeof(svo_completion), sizeof(svo_vertex), sizeof(svo_edge),"
        "sizeof(svo_node_rec), offsetof(svo_tree_info, level_nodes), offsetof(svo_node_rec, neighbors),"
        "sizeof(svo_tune_result), sizeof(svo_pass_stats));"
        "return 0;}\n")
    exe = tmp_path / "sizes"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)],
                   check=True)
    got = [int(x) for x in subprocess.check_output([str(exe)]).split()]
    want = [C.sizeof(svo_amd.MeshDesc), C.sizeof(svo_amd.Params), C.sizeof(svo_amd.TreeInfo),
            C.sizeof(svo_amd.ServerCfg), C.sizeof(svo_amd.Completion),
            svo_amd.VERT_DTYPE.itemsize, svo_amd.EDGE_DTYPE.itemsize,
            svo_amd.NODE_DTYPE.itemsize, svo_amd.TreeInfo.level_nodes.offset,
            svo_amd.NODE_DTYPE.fields["neighbors"][1], C.sizeof(svo_amd.TuneResult),
            11 * 8 + 4 * 4 + 4 * 8]
    assert got == want
    assert got[5:8] == [16, 28, 168]


def test_status_strings():
    for st in range(-9, 1):
        assert svo_amd.status_str(st) != "unknown status"


def test_validate_statuses():
    P = np.random.default_rng(0).random((3, 3)).astype(np.float32)
    T = np.array([[0, 1, 2]], np.int32)
    assert svo_amd.validate(P, T, 64) == svo_amd.OK
    assert svo_amd.validate(P, T, 1024, key_bits=32) == svo_amd.OK
    assert svo_amd.validate(P, T, 1025, key_bits=32) == svo_amd.E_RES_LIMIT
    assert svo_amd.validate(P, T, 2048) == svo_amd.OK           # picks 64-bit keys
    assert svo_amd.validate(P, T, 65537, key_bits=64) == svo_amd.E_RES_LIMIT
    assert svo_amd.validate(P, T, 0) == svo_amd.E_INVALID
    assert svo_amd.validate(P, T, (64, 64, -1)) == svo_amd.E_INVALID
    assert svo_amd.validate(P, T, 64, key_bits=48) == svo_amd.E_INVALID
    assert svo_amd.validate(P, T, 64, semantics=7) == svo_amd.E_INVALID
    assert svo_amd.validate(P[:0], T, 64) == svo_amd.E_EMPTY
    assert svo_amd.validate(P, T[:0], 64) == svo_amd.E_EMPTY


def test_estimate_bytes_monotone():
    e = [svo_amd.estimate_bytes(1000, 2000, r) for r in (64, 128, 256, 512)]
    assert all(b > a for a, b in zip(e, e[1:]))
    assert svo_amd.estimate_bytes(1000, 2000, 128, flags=svo_amd.BUILD_NO_VERTEX_EDGE) < e[1]
    assert svo_amd.estimate_bytes(10 ** 6, 2 * 10 ** 6, 64) > e[0]


def test_context_without_device_reports_hip_error():
    # on a host without a GPU, context creation fails with SVO_E_HIP instead
    # of terminating the process
    import ctypes
    h = ctypes.c_void_p()
    st = svo_amd.lib().svo_ctx_create(0, 0, ctypes.byref(h))
    if st == svo_amd.OK:
        svo_amd.lib().svo_ctx_destroy(h)
        pytest.skip("a device is present")
    assert st == svo_amd.E_HIP
    assert svo_amd.lib().svo_ctx_create(-1, 0, ctypes.byref(h)) == svo_amd.E_HIP
