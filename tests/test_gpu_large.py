"""Trees too large for the oracle: size-independent properties.

A 1,000,000-triangle soup at res 2048 (64-bit keys) gives several hundred
million nodes, so the neighbor table alone holds more than 2^32 words and
every kernel that covers it must index past 32 bits.  Checked without the
oracle: reference-mode neighbors are UINT_MAX except root[13] = root, on a
random sample of nodes and on the last nodes of the array; level counts add
up to the node count.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NONE = 0xFFFFFFFF


def test_soup_2048_beyond_32bit_indexing(gpu_ctx):
    import svo_amd
    from svo_amd import meshgen
    P, T = meshgen.soup(1_000_000, seed=2)
    print("large case: soup 1M triangles res 2048 key_bits 64", flush=True)
    tree = gpu_ctx.build(P, T, 2048, key_bits=64, flags=svo_amd.BUILD_NO_VERTEX_EDGE)
    try:
        info = tree.info
        N = int(info.n_nodes)
        print(f"nodes {N} voxels {info.n_voxels} depth {info.depth}", flush=True)
        assert 27 * N > 2**32, "case too small to cross 32-bit indexing"
        lv = [int(x) for x in info.level_nodes[: info.depth]]
        assert sum(lv) == N and lv[-1] == 1
        rng = np.random.default_rng(0)
        nodes = np.concatenate([rng.integers(0, N, 20000), np.arange(N - 2000, N)])
        nodes = nodes.astype(np.uint32)
        for s in range(27):
            got = tree.query_neighbors(nodes, np.full(nodes.size, s, np.uint8))
            exp = np.full(nodes.size, NONE, np.uint32)
            if s == 13:
                exp[nodes == N - 1] = N - 1
            assert np.array_equal(got, exp), f"slot {s}"
    finally:
        tree.free()
