"""Pins the chunked full-size checks (tests/fullsize.py) against the oracle and
against the closed form the oracle is pinned by, on small meshes, so that the
GPU tests at full size check against something independently fixed."""
import json
import os

import numpy as np
import pytest

import fullsize as fs
import pyoracle as po
from test_oracle_levels import closed_form

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_decode_matches_oracle():
    rng = np.random.default_rng(0)
    codes = np.concatenate([rng.integers(0, 1 << 30, 2000), rng.integers(0, 1 << 33, 2000),
                            [0, (1 << 30) - 1, (1 << 33) - 1]]).astype(np.uint64)
    got = fs.decode(codes)
    for c, (x, y, z) in zip(codes[:500].tolist() + codes[-503:].tolist(),
                            np.concatenate([got[:500], got[-503:]]).tolist()):
        assert po.morton_decode(int(c)) == (x, y, z)


def _cases():
    from svo_amd import meshgen
    P, T = meshgen.icosphere(3)
    yield "ico3@32", P, T, 32, 32
    P, T = meshgen.torus(0.4, 0.1, 40, 20)
    yield "torus@64", P, T, 64, 32
    P, T = meshgen.soup(400, seed=4, median=1 / 32)
    yield "soup@128", P, T, 128, 32
    P, T = meshgen.icosphere(2)
    yield "ico2@(20,13,7)", P, T, (20, 13, 7), 32
    P, T = meshgen.soup(300, seed=9, median=1 / 64)
    yield "soup@2048-u64", P, T, 2048, 64


@pytest.mark.parametrize("case", range(5))
def test_levels_rows_equal_oracle(case):
    name, P, T, res, kb = list(_cases())[case]
    r = po.build(P, T, res, key_bits=kb)
    assert r.status == 0
    lv = fs.Levels(r["voxels"], r.info.unit[0], r.info.bbox_min)
    assert lv.depth == r.depth and lv.N == r.info.n_nodes
    assert np.array_equal(lv.level_nodes(), r["level_nodes"])
    if kb == 32:
        cf = closed_form(r["voxels"], r.info.unit[0], r.info.bbox_min)
        assert np.array_equal(cf["off"], lv.off)
    # odd chunk sizes cut levels and parent groups in the middle
    for chunk in (lv.N, 997, 64):
        for start in range(0, lv.N, chunk):
            got = lv.rows(start, chunk)
            sl = slice(start, min(lv.N, start + chunk))
            assert np.array_equal(got["code"], r["code"][sl].astype(np.uint64)), (name, start)
            assert np.array_equal(got["is_leaf"], r["is_leaf"][sl]), (name, start)
            assert np.array_equal(got["parent"], r["parent"][sl]), (name, start)
            assert np.array_equal(got["childs"], r["childs"][sl]), (name, start)
            assert np.array_equal(got["width"].view(np.uint32), r["width"][sl].view(np.uint32))
            assert np.array_equal(got["origin"].view(np.uint32), r["origin"][sl].view(np.uint32))
        if lv.N > 20000:
            break


def test_record_points_equal_oracle():
    with open(os.path.join(GOLDEN, "luts.json")) as f:
        d = json.load(f)
    ofs = np.array(d["edge_endpoints"], np.int64)
    bare = np.array(d["edge_endpoint_bare"], bool)
    from svo_amd import meshgen
    P, T = meshgen.torus(0.4, 0.1, 40, 20)
    r = po.build(P, T, 64)
    N = r.info.n_nodes
    e = fs.edge_points(r["origin"], r["width"], ofs, bare).reshape(12 * N, 2, 3)
    assert np.array_equal(r["edges"]["a"].view(np.uint32), e[:, 0].view(np.uint32))
    assert np.array_equal(r["edges"]["b"].view(np.uint32), e[:, 1].view(np.uint32))
    v = fs.vertex_points(r["width"]).reshape(8 * N, 3)
    assert np.array_equal(r["vertices"]["p"].view(np.uint32), v.view(np.uint32))
