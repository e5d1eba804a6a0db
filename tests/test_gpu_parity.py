"""HIP build vs the CPU oracle, every output field bit-exact.

Cases cover the reference's sample mesh (bunny.obj), the plumbing config
(icosphere subdivision 4 at res 64), tori, triangle soups, non-power-of-two
and non-cubic grids, the widened 64-bit key path, and the failure statuses.
"""
import numpy as np
import pytest

import pyoracle as po
from parity import compare

pytestmark = pytest.mark.gpu


def _meshgen():
    from svo_amd import meshgen
    return meshgen


def _case(ctx, P, T, res, sem=0, kb=32):
    print(f"case: {P.shape[0]} points {T.shape[0]} tris res {res} sem {sem} key_bits {kb}",
          flush=True)
    tree = ctx.build(P, T, res, semantics=sem, key_bits=kb)
    ref = po.build(P, T, res, semantics=sem, key_bits=kb)
    assert ref.status == 0
    compare(tree, ref)
    return tree, ref


@pytest.mark.parametrize("sem", [0, 1])
@pytest.mark.parametrize("res", [64, 32, 100])
def test_bunny(gpu_ctx, bunny, res, sem):
    P, T = bunny
    _case(gpu_ctx, P, T, res, sem)


def test_bunny_noncubic(gpu_ctx, bunny):
    P, T = bunny
    _case(gpu_ctx, P, T, (64, 32, 48))
    _case(gpu_ctx, P, T, (17, 90, 33), sem=1)


@pytest.mark.parametrize("sem", [0, 1])
def test_icosphere4_res64(gpu_ctx, sem):
    P, T = _meshgen().icosphere(4)
    tree, ref = _case(gpu_ctx, P, T, 64, sem)
    assert T.shape[0] == 5120 and P.shape[0] == 2562


@pytest.mark.parametrize("res", [128, 256])
def test_torus(gpu_ctx, res):
    P, T = _meshgen().torus(0.4, 0.12, 96, 48)
    _case(gpu_ctx, P, T, res)


@pytest.mark.parametrize("seed,n,res", [(11, 500, 64), (12, 3000, 256), (13, 200, 1024)])
def test_soup(gpu_ctx, seed, n, res):
    P, T = _meshgen().soup(n, seed=seed, median=1.0 / 64)
    _case(gpu_ctx, P, T, res)
    _case(gpu_ctx, P, T, res, sem=1)


def test_big_triangles(gpu_ctx):
    # a few triangles spanning most of the grid take the big-triangle path
    rng = np.random.default_rng(5)
    P = rng.random((12, 3)).astype(np.float32)
    T = np.arange(12, dtype=np.int32).reshape(4, 3)
    _case(gpu_ctx, P, T, 128)
    _case(gpu_ctx, P, T, 300, sem=1)


def _slivers(n, seed):
    # long thin axis-following triangles: few box cells even at high resolution
    rng = np.random.default_rng(seed)
    a = rng.random((n, 3)) * 0.6
    d = np.zeros((n, 3))
    d[np.arange(n), rng.integers(0, 3, n)] = 0.3 + 0.1 * rng.random(n)
    b = a + d + 0.001 * rng.random((n, 3))
    c = a + 0.002 * rng.random((n, 3))
    P = np.stack([a, b, c], 1).reshape(-1, 3).astype(np.float32)
    return P, np.arange(3 * n, dtype=np.int32).reshape(n, 3)


def test_key64(gpu_ctx):
    P, T = _meshgen().icosphere(3)
    _case(gpu_ctx, P, T, 64, kb=64)
    P, T = _slivers(20, 3)
    _case(gpu_ctx, P, T, 2048, kb=64)
    _case(gpu_ctx, P, T, 4096, kb=64, sem=1)


def test_dense_bitmap_tiles(gpu_ctx):
    """Most cells occupied: a round of the level-1 emission (256 nonzero
    bitmap words) then holds more unique keys than its LDS stage (2048, i.e.
    more than eight per word), and its keys are stored directly."""
    # 32 stacked unit squares (two triangles each), one per 1/32 in z: at
    # res 32 every cell is a voxel; at res 48 most are
    P, T = [], []
    for k in range(32):
        z = (k + 0.5) / 32
        P += [[0, 0, z], [1, 0, z], [1, 1, z], [0, 1, z]]
        T += [[4 * k, 4 * k + 1, 4 * k + 2], [4 * k, 4 * k + 2, 4 * k + 3]]
    P, T = np.array(P, np.float32), np.array(T, np.int32)
    for path in (1, 2):  # the bitmap path, and the sort path on the same input
        gpu_ctx.set_option("dedup_path", path)
        try:
            for res in (32, 48):
                tree, _ = _case(gpu_ctx, P, T, res)
                assert tree.info.n_voxels > 0.5 * res ** 3, "case too sparse to fill the stage"
                tree.free()
        finally:
            gpu_ctx.set_option("dedup_path", 0)


def test_single_triangle(gpu_ctx):
    P = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0.5]], np.float32)
    T = np.array([[0, 1, 2]], np.int32)
    for res in (1, 2, 5, 16):
        _case(gpu_ctx, P, T, res)


def test_statuses(gpu_ctx):
    import svo_amd
    P = np.zeros((3, 3), np.float32)
    T = np.array([[0, 1, 2]], np.int32)
    with pytest.raises(svo_amd.SvoError) as e:
        gpu_ctx.build(P, T, 16)  # zero-extent box
    assert e.value.status == svo_amd.E_INVALID
    P = np.random.default_rng(0).random((3, 3)).astype(np.float32)
    with pytest.raises(svo_amd.SvoError) as e:
        gpu_ctx.build(P, np.array([[0, 1, 3]], np.int32), 16)
    assert e.value.status == svo_amd.E_INVALID
    with pytest.raises(svo_amd.SvoError) as e:
        gpu_ctx.build(P, T, 2048, key_bits=32)
    assert e.value.status == svo_amd.E_RES_LIMIT
    P2 = P.copy()
    P2[1, 0] = np.nan
    with pytest.raises(svo_amd.SvoError) as e:
        gpu_ctx.build(P2, T, 16)
    assert e.value.status == svo_amd.E_INVALID


@pytest.mark.parametrize("res,kb", [(64, 32), ((40, 24, 33), 32), (128, 64)])
def test_obj_files_byte_exact(gpu_ctx, bunny, tmp_path, res, kb):
    """writeTree and writeVoxel files from the GPU tree equal the oracle's bytes."""
    P, T = bunny
    print(f"obj file case: bunny res {res} key_bits {kb}", flush=True)
    tree = gpu_ctx.build(P, T, res, key_bits=kb)
    d = tmp_path / "gpu"
    d.mkdir()
    path = tree.write_obj(str(d), "bunny")
    vpath = tree.write_voxels_obj(str(d), "bunny")
    depth = tree.info.depth
    assert path.endswith(f"bunny_{depth}_tree.obj") and vpath.endswith("bunny_1_voxel.obj")
    ref_t = str(tmp_path / "ref_tree.obj")
    ref_v = str(tmp_path / "ref_voxel.obj")
    st, rdepth = po.write_obj(P, T, res, ref_t, key_bits=kb)
    assert st == 0 and rdepth == depth
    st, _ = po.write_obj(P, T, res, ref_v, key_bits=kb, voxels=True)
    assert st == 0
    assert open(path, "rb").read() == open(ref_t, "rb").read()
    assert open(vpath, "rb").read() == open(ref_v, "rb").read()
    tree.free()


@pytest.mark.parametrize("path", [1, 2], ids=["bitmap", "sort"])
def test_dedup_paths(gpu_ctx, bunny, path):
    """Both de-duplication paths (occupancy bitmap, radix sort of the hits)
    give the oracle's tree.  The builds share one context, so the bitmap
    path reuses one bitmap buffer at sizes that grow and shrink: a bit left
    set by an earlier build (the level-1 pass clears what it reads) would
    show as an extra voxel."""
    gpu_ctx.set_option("dedup_path", path)
    assert gpu_ctx.get_option("dedup_path") == path
    try:
        P, T = bunny
        _case(gpu_ctx, P, T, 64)
        _case(gpu_ctx, P, T, (40, 24, 33), sem=1)
        Ps, Ts = _meshgen().soup(4000, seed=7, median=1.0 / 32.0)
        _case(gpu_ctx, Ps, Ts, 256)
        _case(gpu_ctx, Ps, Ts, 100, kb=64)
        P1 = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0.5]], np.float32)
        T1 = np.array([[0, 1, 2]], np.int32)
        for res in (1, 2, 7):
            _case(gpu_ctx, P1, T1, res)
        Pl, Tl = _slivers(20, 3)
        _case(gpu_ctx, Pl, Tl, 2048, kb=64)
    finally:
        gpu_ctx.set_option("dedup_path", 0)


def test_options_rejected(gpu_ctx):
    import svo_amd
    for name, value in (("dedup_path", 3), ("no_such_option", 1), ("sort_blocks", 0)):
        with pytest.raises(svo_amd.SvoError) as e:
            gpu_ctx.set_option(name, value)
        assert e.value.status == svo_amd.E_INVALID


def test_named_readouts(gpu_ctx, bunny):
    """svo_tree_copy_{voxels,nodes,neighbors,vertices,edges,faces} return the
    same bytes as the corresponding field copies."""
    import svo_amd
    P, T = bunny
    tree = gpu_ctx.build(P, T, 32, semantics=1)
    L = svo_amd.lib()
    try:
        for name, field in (("voxels", "voxels"), ("nodes", "nodes_aos"),
                            ("neighbors", "neighbors"), ("vertices", "vertices"),
                            ("edges", "edges"), ("faces", "faces")):
            ref = tree.field(field)
            buf = np.zeros_like(ref)
            st = getattr(L, f"svo_tree_copy_{name}")(tree._h, buf.ctypes.data, buf.nbytes)
            assert st == 0, name
            assert buf.tobytes() == ref.tobytes(), name
    finally:
        tree.free()


def test_reference_unique_removes_records(gpu_ctx):
    """The reference's consecutive unique (uniqueVert / uniqueEdge,
    SVO.cu:679-720) removes edge records where node corners collapse in
    float32: icospheres far from the origin (o + w == o for widths below
    2^-8 at 1e5).  The build takes the count / scan / write path there;
    alone and in a batch beside ordinary meshes it must give the oracle's
    arrays (the oracle's removals are pinned in test_oracle_pins.py)."""
    from test_oracle_pins import far_mesh
    cases = [(0.25, 256), (0.05, 64), (0.02, 100)]
    batch = []
    for radius, res in cases:
        P, T = far_mesh(radius=radius)
        print(f"far mesh radius {radius} res {res}", flush=True)
        ref = po.build(P, T, res)
        assert ref.status == 0
        assert ref.info.n_edges < 12 * ref.info.n_nodes  # records were removed
        tree = gpu_ctx.build(P, T, res)
        try:
            compare(tree, ref)
        finally:
            tree.free()
        batch.append((P, T, res, ref))
    Pn, Tn = _meshgen().icosphere(3)
    batch.insert(1, (Pn, Tn, 64, po.build(Pn, Tn, 64)))
    print(f"far meshes in a batch of {len(batch)}", flush=True)
    trees, st = gpu_ctx.build_batch([(P, T, r) for P, T, r, _ in batch])
    try:
        assert all(s == 0 for s in st), st
        for t, (_, _, _, ref) in zip(trees, batch):
            compare(t, ref)
    finally:
        for t in trees:
            t.free()
