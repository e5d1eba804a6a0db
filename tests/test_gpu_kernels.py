"""Single-kernel entry points (include/svo/svo_kernels.h) on the GPU.

svo_k_radix_sort is checked against numpy's stable sort on the same bits
(tile edges 4095/4096/4097, runs of equal keys, presorted and reversed input,
partial-bit sorts whose stability shows in the high bits); svo_k_voxelize's
raw hits, deduplicated, against the oracle's voxel set; svo_k_octree_level
against numpy's unique on random sorted codes with duplicates, and chained
from the oracle's voxels up to the root against the oracle's node codes of
every level; svo_k_ref_vertex_edge on the oracle's node origins and widths
against the oracle's vertex and edge records.
"""
import numpy as np
import pytest

import pyoracle as po
import svo_amd

pytestmark = pytest.mark.gpu


def _expect(keys, bits):
    w = 8 * keys.dtype.itemsize
    if bits >= w:
        return np.sort(keys, kind="stable")
    mask = keys.dtype.type((1 << bits) - 1)
    order = np.argsort(keys & mask, kind="stable")
    return keys[order]


def _sort_case(keys, bits):
    print(f"sort case: n {keys.size} dtype {keys.dtype} bits {bits}", flush=True)
    got = svo_amd.k_radix_sort(keys, bits)
    exp = _expect(keys, bits)
    assert got.dtype == keys.dtype and got.shape == keys.shape
    assert np.array_equal(got, exp)


@pytest.mark.parametrize("n", [1, 63, 64, 4095, 4096, 4097, 100_003, 3_000_000])
@pytest.mark.parametrize("dt", [np.uint32, np.uint64])
def test_sort_random(n, dt):
    rng = np.random.default_rng(n)
    keys = rng.integers(0, np.iinfo(dt).max, size=n, dtype=dt, endpoint=True)
    _sort_case(keys, 8 * np.dtype(dt).itemsize)


@pytest.mark.parametrize("bits", [1, 7, 8, 13, 30])
def test_sort_partial_bits_stable(bits):
    rng = np.random.default_rng(bits)
    keys = rng.integers(0, 2**32 - 1, size=200_000, dtype=np.uint32)
    _sort_case(keys, bits)


def test_sort_patterns():
    n = 1 << 20
    _sort_case(np.full(n, 7, dtype=np.uint32), 32)
    _sort_case(np.arange(n, dtype=np.uint32), 32)
    _sort_case(np.arange(n, dtype=np.uint64)[::-1].copy() << np.uint64(20), 64)
    rng = np.random.default_rng(5)
    runs = np.repeat(rng.integers(0, 1 << 30, size=n // 64, dtype=np.uint32), 64)
    _sort_case(runs, 30)


def test_sort_voxel_keys_u64():
    # keys shaped like composite batch keys: mesh index above a 33-bit code
    rng = np.random.default_rng(11)
    mesh = rng.integers(0, 300, size=2_000_000).astype(np.uint64)
    code = rng.integers(0, 1 << 33, size=mesh.size, dtype=np.uint64)
    _sort_case((mesh << np.uint64(33)) | code, 42)


def test_sort_empty():
    out = svo_amd.k_radix_sort(np.empty(0, dtype=np.uint32))
    assert out.size == 0


@pytest.mark.parametrize("res,kb", [(64, 32), (256, 32), (300, 64)])
def test_voxelize_matches_oracle(bunny, res, kb):
    P, T = bunny
    print(f"voxelize case: bunny res {res} key_bits {kb}", flush=True)
    hits, n = svo_amd.k_voxelize(P, T, res, key_bits=kb)
    assert hits.size == n
    st, ref = po.voxelize(P, T, res, key_bits=kb)
    assert st == 0
    assert np.array_equal(np.unique(hits.astype(np.uint64)), np.unique(ref))


def _level_expect(keys):
    uniq = np.unique(keys)
    par_all = uniq >> uniq.dtype.type(3)
    parents, first = np.unique(par_all, return_index=True)
    return uniq, parents, first.astype(np.uint32)


@pytest.mark.parametrize("n,dt,bits", [(1, np.uint32, 30), (5000, np.uint32, 30),
                                       (1_000_003, np.uint32, 30), (300_000, np.uint64, 45),
                                       (300_000, np.uint64, 63)])
def test_octree_level_random(n, dt, bits):
    rng = np.random.default_rng(n + bits)
    # duplicates and dense runs: draw from a range about n wide, then sort
    lo = int(rng.integers(0, (1 << bits) - 4 * n))
    keys = np.sort((lo + rng.integers(0, 4 * n, size=n)).astype(dt))
    print(f"octree level case: n {n} dtype {np.dtype(dt).name} code bits {bits}", flush=True)
    uniq, par, fc = svo_amd.k_octree_level(keys, bits)
    eu, ep, ef = _level_expect(keys)
    assert uniq.dtype == keys.dtype and par.dtype == keys.dtype
    assert np.array_equal(uniq, eu)
    assert np.array_equal(par, ep)
    assert np.array_equal(fc, ef)


def test_octree_level_chain_matches_oracle_levels(bunny):
    """Level steps from the oracle's voxels up to the root give the parent codes
    of the oracle's tree, level by level (its level-l nodes are the eight
    children 8p..8p+7 of every level-(l+1) parent p)."""
    P, T = bunny
    ref = po.build(P, T, 128, key_bits=32, fields={"voxels", "code", "level_nodes"})
    assert ref.status == 0
    codes, counts = ref.arrays["code"], ref.arrays["level_nodes"].astype(np.int64)
    cur = ref.arrays["voxels"]
    off = 0
    for l, nl in enumerate(counts[:-1]):
        print(f"octree level chain: level {l + 1}, {cur.size} codes", flush=True)
        uniq, par, fc = svo_amd.k_octree_level(cur, 30)
        assert np.array_equal(uniq, cur)  # already unique
        level_codes = codes[off:off + nl]
        assert np.array_equal(np.unique(level_codes >> np.uint32(3)), par)
        assert np.array_equal(fc, np.searchsorted(uniq >> np.uint32(3), par).astype(np.uint32))
        off += nl
        cur = par
    assert cur.size == 1


def test_octree_level_rejects_bad_input():
    with pytest.raises(svo_amd.SvoError):
        svo_amd.k_octree_level(np.array([5, 3], dtype=np.uint32), 30)  # unsorted
    with pytest.raises(svo_amd.SvoError):
        svo_amd.k_octree_level(np.array([1, 1 << 30], dtype=np.uint32), 30)  # past code_bits
    with pytest.raises(svo_amd.SvoError):
        svo_amd.k_octree_level(np.array([1], dtype=np.uint32), 32)  # code_bits too large
    uniq, par, fc = svo_amd.k_octree_level(np.empty(0, dtype=np.uint64), 40)
    assert uniq.size == par.size == fc.size == 0


@pytest.mark.parametrize("res", [64, 100])
def test_ref_vertex_edge_matches_oracle(bunny, res):
    """The oracle tree's node origins and widths through svo_k_ref_vertex_edge
    give the oracle's vertex and edge records bit for bit (reference
    semantics, where the consecutive unique removes nothing on the bunny)."""
    P, T = bunny
    ref = po.build(P, T, res, key_bits=32, fields={"origin", "width", "vertices", "edges"})
    assert ref.status == 0
    o, w = ref.arrays["origin"], ref.arrays["width"]
    print(f"vertex/edge records case: bunny res {res}, {w.size} nodes", flush=True)
    verts, edges = svo_amd.k_ref_vertex_edge(o, w)
    assert verts.size == ref.arrays["vertices"].size == 8 * w.size
    assert edges.size == ref.arrays["edges"].size == 12 * w.size
    assert np.array_equal(verts.view(np.uint32), ref.arrays["vertices"].view(np.uint32))
    assert np.array_equal(edges.view(np.uint32), ref.arrays["edges"].view(np.uint32))


def test_ref_vertex_edge_empty():
    v, e = svo_amd.k_ref_vertex_edge(np.empty((0, 3), np.float32), np.empty(0, np.float32))
    assert v.size == 0 and e.size == 0
