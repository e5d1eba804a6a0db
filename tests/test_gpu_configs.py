"""GPU builds at the benchmark configurations, checked against the oracle.

The icosphere-4 at res 64 (the plain single-mesh case) is checked bit-exact in
test_gpu_parity.py.  Here:

- icosphere-7 + 3-octave value noise (327,680 triangles) at res 512: every
  output field bit-exact against the oracle;
- the 2,000,000-triangle soup at res 1024 (uint32 keys, with the vertex and
  edge records) and at res 2048 (uint64 keys, no records): the voxel set
  bit-exact against the oracle's voxelizer (all of it), the grid against the
  oracle's, level counts and node fields against the closed form over that
  voxel set (tests/fullsize.py, pinned against the oracle on small meshes):
  every node at res 1024, evenly spread node ranges plus every level's first
  and last nodes at res 2048; reference neighbors (UINT_MAX except the root's
  centre slot) on every node at res 1024; vertex / edge records on node
  ranges at the start, middle and end;
- the 256-mesh batch of the benchmark (icospheres and tori, res 64-256,
  seed 3): batched == plain == oracle for every mesh and every field;
- a served Poisson stream (2000 req/s, res 64-512, 64 GB budget): sampled
  results bit-exact against the oracle, 1000 neighbor queries per tree
  against its own neighbor array.
"""
import json
import os
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import fullsize as fs
import pyoracle as po
from parity import assert_same, compare

pytestmark = pytest.mark.gpu

NONE = 0xFFFFFFFF
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
ALL = ["voxels", "code", "is_leaf", "origin", "width", "parent", "childs", "neighbors",
       "vertices", "edges", "faces", "level_nodes"]


def _threads():
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def _edge_table():
    with open(os.path.join(GOLDEN, "luts.json")) as f:
        d = json.load(f)
    return np.array(d["edge_endpoints"], np.int64), np.array(d["edge_endpoint_bare"], bool)


def _same_bits(name, got, want):
    assert got.shape == want.shape, (name, got.shape, want.shape)
    g = np.ascontiguousarray(got).view(np.uint8).reshape(got.shape[0], -1)
    w = np.ascontiguousarray(want).view(np.uint8).reshape(want.shape[0], -1)
    if not np.array_equal(g, w):
        rows = np.nonzero((g != w).any(axis=1))[0]
        raise AssertionError(f"{name}: {rows.size} rows differ, first {rows[0]}: "
                             f"got {got[rows[0]]!r} want {want[rows[0]]!r}")


def _check_grid(tree, P, res, key_bits):
    """bbox cube and unit against the oracle's grid (it depends on the points
    and the resolution only: one triangle is enough)."""
    r = po.build(P, np.array([[0, 1, 2]], np.int32), res, key_bits=key_bits, fields=("voxels",))
    assert r.status == 0
    for k in ("bbox_min", "bbox_max", "bbox_width", "unit"):
        assert np.array(getattr(tree.info, k), np.float32).tobytes() == \
            np.array(getattr(r.info, k), np.float32).tobytes(), k


def _check_node_rows(tree, lv, start, count, origin=True):
    exp = lv.rows(start, count, want_origin=origin)
    n = exp["code"].size
    _same_bits(f"code@{start}", tree.field_rows("code", start, n).astype(np.uint64), exp["code"])
    _same_bits(f"is_leaf@{start}", tree.field_rows("is_leaf", start, n), exp["is_leaf"])
    _same_bits(f"width@{start}", tree.field_rows("width", start, n), exp["width"])
    _same_bits(f"parent@{start}", tree.field_rows("parent", start, n), exp["parent"])
    _same_bits(f"childs@{start}", tree.field_rows("childs", start, n), exp["childs"])
    if origin:
        _same_bits(f"origin@{start}", tree.field_rows("origin", start, n), exp["origin"])


def _check_neighbor_rows(tree, start, count):
    nb = tree.field_rows("neighbors", start, count)
    N = tree.info.n_nodes
    bad = nb != NONE
    if start + nb.shape[0] == N:
        assert nb[-1, 13] == N - 1, "root centre slot"
        bad[-1, 13] = False
    assert not bad.any(), f"neighbors@{start}: {np.count_nonzero(bad)} slots set"


def _check_record_rows(tree, start, count, ofs, bare):
    """Reference records of nodes [start, start+count): 8 vertices and 12
    edges per node (no adjacent duplicates at these widths), owned by the
    node (only the root has a neighbor, itself)."""
    org = tree.field_rows("origin", start, count)
    wid = tree.field_rows("width", start, count)
    n = wid.size
    own = np.arange(start, start + n, dtype=np.uint32)
    v = tree.field_rows("vertices", 8 * start, 8 * n)
    _same_bits(f"vertices@{start}", v["p"], fs.vertex_points(wid).reshape(8 * n, 3))
    assert np.array_equal(v["owner"], np.repeat(own, 8)), f"vertex owners@{start}"
    e = tree.field_rows("edges", 12 * start, 12 * n)
    pts = fs.edge_points(org, wid, ofs, bare).reshape(12 * n, 2, 3)
    _same_bits(f"edges.a@{start}", e["a"], pts[:, 0])
    _same_bits(f"edges.b@{start}", e["b"], pts[:, 1])
    assert np.array_equal(e["owner"], np.repeat(own, 12)), f"edge owners@{start}"


# ------------------------------------------------------------------ icosphere-7 @ 512
def test_noisy_icosphere7_res512_bit_exact(gpu_ctx):
    from svo_amd import meshgen
    P, T = meshgen.noisy_icosphere(7, seed=1)
    print(f"config: icosphere-7 + noise, {T.shape[0]} triangles, res 512", flush=True)
    tree = gpu_ctx.build(P, T, 512)
    try:
        ref = po.build(P, T, 512, threads=_threads())
        assert ref.status == 0
        print(f"nodes {tree.info.n_nodes} voxels {tree.info.n_voxels}", flush=True)
        compare(tree, ref)
    finally:
        tree.free()


# ------------------------------------------------------------------ soup @ 1024 / 2048
CHUNK = 1 << 24


class Soup:
    def __init__(self, ctx, res, key_bits, flags):
        import svo_amd
        from svo_amd import meshgen
        self.P, self.T = meshgen.soup(2_000_000, seed=2)
        print(f"config: soup 2M triangles res {res} key_bits {key_bits} flags {flags}",
              flush=True)
        self.tree = ctx.build(self.P, self.T, res, key_bits=key_bits, flags=flags)
        info = self.tree.info
        print(f"nodes {info.n_nodes} voxels {info.n_voxels} depth {info.depth}", flush=True)
        t0 = time.time()
        st, self.ref_voxels = po.voxelize(self.P, self.T, res, key_bits=key_bits,
                                          threads=_threads())
        assert st == 0
        print(f"oracle voxelizer: {self.ref_voxels.size} voxels in {time.time() - t0:.1f} s",
              flush=True)
        self.lv = fs.Levels(self.ref_voxels, info.unit[0], info.bbox_min)
        self.res, self.key_bits = res, key_bits
        self.svo_amd = svo_amd

    def free(self):
        self.tree.free()


def _check_voxels_levels(s):
    tree, lv = s.tree, s.lv
    _check_grid(tree, s.P, s.res, s.key_bits)
    got = tree.field("voxels").astype(np.uint64)
    _same_bits("voxels", got, s.ref_voxels)
    assert tree.depth == lv.depth
    assert tree.info.n_nodes == lv.N
    assert np.array_equal(tree.field("level_nodes"), lv.level_nodes())


# Class-scoped trees: each is freed before the next class builds (the two
# soups take about 150 GB and 125 GB of device memory).
class TestSoup1024:
    @pytest.fixture(scope="class")
    def soup(self, gpu_ctx):
        s = Soup(gpu_ctx, 1024, 32, 0)
        yield s
        s.free()

    def test_voxels_and_levels(self, soup):
        _check_voxels_levels(soup)

    def test_every_node(self, soup):
        tree, lv = soup.tree, soup.lv
        starts = list(range(0, lv.N, CHUNK))
        for j, a in enumerate(starts):
            # origins (a float32 decode per node) on every third chunk and the last
            _check_node_rows(tree, lv, a, CHUNK, origin=(j % 3 == 0 or a == starts[-1]))

    def test_neighbors_every_node(self, soup):
        for a in range(0, soup.lv.N, CHUNK):
            _check_neighbor_rows(soup.tree, a, CHUNK)

    def test_vertex_edge_records(self, soup):
        tree, N = soup.tree, soup.lv.N
        assert tree.info.n_vertices == 8 * N and tree.info.n_edges == 12 * N
        ofs, bare = _edge_table()
        k = 1 << 20
        for a in (0, N // 2 - k // 2, N - k):
            _check_record_rows(tree, a, k, ofs, bare)


class TestSoup2048:
    @pytest.fixture(scope="class")
    def soup(self, gpu_ctx):
        import svo_amd
        s = Soup(gpu_ctx, 2048, 64, svo_amd.BUILD_NO_VERTEX_EDGE)
        yield s
        s.free()

    def test_voxels_and_levels(self, soup):
        _check_voxels_levels(soup)
        # the neighbor table alone holds more than 2^32 words: the kernels that
        # cover it index past 32 bits
        assert 27 * soup.lv.N > 2 ** 32

    def test_node_ranges(self, soup):
        tree, lv = soup.tree, soup.lv
        k = 1 << 21
        for a in np.linspace(0, lv.N - k, 12).astype(np.int64):
            _check_node_rows(tree, lv, int(a), k)
        for i in range(lv.L):  # every level's first and last nodes
            _check_node_rows(tree, lv, int(lv.off[i]), 4096)
            _check_node_rows(tree, lv, max(0, int(lv.off[i + 1]) - 4096), 4096)

    def test_neighbor_ranges(self, soup):
        N = soup.lv.N
        k = 1 << 22
        for a in [int(x) for x in np.linspace(0, N - k, 10).astype(np.int64)] + [N - k]:
            _check_neighbor_rows(soup.tree, a, k)


# ------------------------------------------------------------------ 256-mesh batch
class TestBenchBatch:
    @pytest.fixture(scope="class")
    def batch(self, gpu_ctx):
        from svo_amd import meshgen
        meshes = meshgen.mixed_batch(256, seed=3)
        print(f"config: 256-mesh batch, {sum(T.shape[0] for _, T, _ in meshes)} triangles",
              flush=True)
        dms = [gpu_ctx.upload_mesh(P, T) for P, T, _ in meshes]
        trees, st = gpu_ctx.build_batch([d.desc(r) for d, (_, _, r) in zip(dms, meshes)],
                                        on_device=True)
        assert all(s == 0 for s in st), st
        yield meshes, dms, trees
        for t in trees:
            t.free()
        for d in dms:
            d.free()

    @pytest.mark.parametrize("part", range(8))
    def test_batched_equals_plain_and_oracle(self, gpu_ctx, batch, part):
        meshes, dms, trees = batch
        idx = list(range(part * 32, part * 32 + 32))
        print(f"batch meshes {idx[0]}-{idx[-1]}", flush=True)
        with ThreadPoolExecutor(_threads()) as ex:
            refs = list(ex.map(lambda i: po.build(*meshes[i]), idx))
        for i, ref in zip(idx, refs):
            P, T, res = meshes[i]
            assert ref.status == 0, i
            plain = gpu_ctx.build_device(*dms[i].desc(res))
            try:
                compare(trees[i], ref)
                for f in ALL + ["nodes_aos"]:
                    assert_same(f"mesh {i} {f} batched vs plain", trees[i].field(f),
                                plain.field(f))
            finally:
                plain.free()


# ------------------------------------------------------------------ served stream
def test_served_stream_against_oracle(gpu_ctx):
    import svo_amd
    from svo_amd import meshgen
    stream = meshgen.request_stream(240, rate=2000.0, seed=4)
    cache = {}
    reqs = []
    for t_arr, s in stream:
        key = (s.kind, s.params)
        if key not in cache:
            cache[key] = s.build()
        reqs.append((t_arr, cache[key][0], cache[key][1], s.res))
    # oracle sample: every 4th request up to 20,480 triangles, and the first
    # six at res 512
    big = [i for i, r in enumerate(reqs) if r[3] == 512][:6]
    sample = sorted(set([i for i in range(0, len(reqs), 4) if reqs[i][2].shape[0] <= 20480]
                        + big))
    print(f"served stream: {len(reqs)} requests, res {sorted({r[3] for r in reqs})}, "
          f"oracle sample {len(sample)}", flush=True)
    srv = svo_amd.Server(gpu_ctx, mem_budget=64 << 30, max_batch=64, n_lanes=2,
                         max_wait_us=300)
    rng = np.random.default_rng(5)
    got = {}
    ids = {}
    try:
        t0 = time.perf_counter()
        i = 0
        done = 0
        while done < len(reqs):
            now = time.perf_counter() - t0
            while i < len(reqs) and reqs[i][0] <= now:
                _, P, T, res = reqs[i]
                ids[srv.submit(P, T, res)] = i
                i += 1
            for d in srv.poll(cap=64, timeout_us=200):
                k = ids[d.request_id]
                assert d.status == 0, (k, d.status)
                tr = d.tree
                n = tr.info.n_nodes
                nodes = rng.integers(0, n, 1000).astype(np.uint32)
                slots = rng.integers(0, 27, 1000).astype(np.uint8)
                q = tr.query_neighbors(nodes, slots)
                assert np.array_equal(q, tr.field_rows("neighbors", 0, n)[nodes, slots])
                if k in sample:
                    got[k] = {f: tr.field(f) for f in ALL}
                    got[k]["_info"] = (tr.info.depth, tr.info.n_voxels, tr.info.n_nodes)
                tr.free()
                done += 1
            assert time.perf_counter() - t0 < 100, f"{done} of {len(reqs)} served"
    finally:
        srv.close()
    with ThreadPoolExecutor(_threads()) as ex:
        futs = {k: ex.submit(po.build, reqs[k][1], reqs[k][2], reqs[k][3]) for k in sample}
        for k in sample:
            ref = futs[k].result()
            assert ref.status == 0
            g = got.pop(k)
            assert g["_info"] == (ref.info.depth, ref.info.n_voxels, ref.info.n_nodes), k
            for f in ALL:
                if f in ref.arrays:
                    assert_same(f"request {k} {f}", g[f], ref[f])
