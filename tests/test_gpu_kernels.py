"""Single-kernel entry points (include/svo/svo_kernels.h) on the GPU.

svo_k_radix_sort is checked against numpy's stable sort on the same bits
(tile edges 4095/4096/4097, runs of equal keys, presorted and reversed input,
partial-bit sorts whose stability shows in the high bits); svo_k_voxelize's
raw hits, deduplicated, against the oracle's voxel set.
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
