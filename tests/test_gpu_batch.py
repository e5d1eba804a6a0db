"""Batched pass (composite (mesh, Morton) keys) == plain per-mesh builds == oracle."""
import numpy as np
import pytest

import pyoracle as po
from parity import compare, assert_same

pytestmark = pytest.mark.gpu

ALL = ["voxels", "code", "is_leaf", "origin", "width", "parent", "childs", "neighbors",
       "vertices", "edges", "faces", "level_nodes"]


def _meshes(n, seed, res_choices=(64, 128)):
    from svo_amd import meshgen
    specs = meshgen.mixed_specs(n, seed=seed, res_choices=res_choices)
    out = []
    for s in specs:
        if s.n_tris > 20480:  # keep the oracle cheap
            continue
        P, T = s.build()
        out.append((P, T, s.res))
    return out


@pytest.mark.parametrize("path", [1, 2], ids=["bitmap", "sort"])
@pytest.mark.parametrize("sem", [0, 1])
def test_batch_equals_plain_and_oracle(gpu_ctx, sem, path):
    meshes = _meshes(40, seed=21)
    print(f"batch of {len(meshes)} meshes, semantics {sem}, dedup path {path}", flush=True)
    gpu_ctx.set_option("dedup_path", path)
    try:
        trees, st = gpu_ctx.build_batch(meshes, semantics=sem)
    finally:
        gpu_ctx.set_option("dedup_path", 0)
    assert all(s == 0 for s in st), st
    for i, ((P, T, res), tb) in enumerate(zip(meshes, trees)):
        tp = gpu_ctx.build(P, T, res, semantics=sem)
        assert tb.depth == tp.depth
        for f in ALL:
            assert_same(f"mesh {i} {f}", tb.field(f), tp.field(f))
        assert_same(f"mesh {i} nodes_aos", tb.field("nodes_aos"), tp.field("nodes_aos"))
        if i % 5 == 0:
            compare(tb, po.build(P, T, res, semantics=sem))
        tp.free()
    for t in trees:
        t.free()


def test_batch_noncubic_grids(gpu_ctx):
    """Bitmap regions whose word counts are not whole tiles (non-cubic and
    small grids) are padded to tile boundaries; the level-1 spans and
    depth-1 roots then come from the tiles' offsets.  A batch of such meshes,
    with a one-voxel mesh among them, must equal the plain builds and the
    oracle."""
    from svo_amd import meshgen
    specs = [sp for sp in meshgen.mixed_specs(16, seed=41, res_choices=(64,))
             if sp.n_tris <= 20480]  # keep the oracle cheap
    grids = [(17, 90, 33), (64, 32, 48), (5, 7, 3), (100, 37, 64), (1, 1, 1), (33, 33, 33),
             (8, 120, 9), (64, 64, 65), (3, 3, 200), (31, 17, 29), (2, 2, 2), (70, 20, 90)]
    meshes = [sp.build() + (g,) for sp, g in zip(specs, grids)]
    assert len(meshes) == len(grids)
    print(f"batch of {len(meshes)} meshes on non-cubic grids", flush=True)
    gpu_ctx.set_option("dedup_path", 1)
    try:
        trees, st = gpu_ctx.build_batch(meshes)
    finally:
        gpu_ctx.set_option("dedup_path", 0)
    assert all(s == 0 for s in st), st
    try:
        for i, ((P, T, res), tb) in enumerate(zip(meshes, trees)):
            tp = gpu_ctx.build(P, T, res)
            assert tb.depth == tp.depth, (i, res)
            for f in ALL:
                assert_same(f"mesh {i} {res} {f}", tb.field(f), tp.field(f))
            compare(tb, po.build(P, T, res))
            tp.free()
    finally:
        for t in trees:
            t.free()


def test_batch_device_inputs(gpu_ctx):
    meshes = _meshes(24, seed=22, res_choices=(64, 128, 256))
    dms = [gpu_ctx.upload_mesh(P, T) for P, T, _ in meshes]
    try:
        trees, st = gpu_ctx.build_batch([d.desc(r) for d, (_, _, r) in zip(dms, meshes)],
                                        on_device=True)
        assert all(s == 0 for s in st), st
        for (P, T, res), tb in zip(meshes, trees):
            tp = gpu_ctx.build(P, T, res)
            for f in ("voxels", "code", "origin", "parent", "childs", "vertices", "edges"):
                assert_same(f, tb.field(f), tp.field(f))
            tp.free()
            tb.free()
    finally:
        for d in dms:
            d.free()


def test_batch_mixed_status(gpu_ctx):
    import svo_amd
    from svo_amd import meshgen
    P, T = meshgen.icosphere(2)
    flat = np.zeros((3, 3), np.float32)
    bad_idx = np.array([[0, 1, 999]], np.int32)
    meshes = [(P, T, 32), (flat, np.array([[0, 1, 2]], np.int32), 32), (P, bad_idx, 32),
              (P, T, 64), (P, T, 2000)]
    trees, st = gpu_ctx.build_batch(meshes, key_bits=32)
    assert st[0] == 0 and st[3] == 0
    assert st[1] == svo_amd.E_INVALID and st[2] == svo_amd.E_INVALID
    assert st[4] == svo_amd.E_RES_LIMIT
    for i in (0, 3):
        compare(trees[i], po.build(*meshes[i]))
    for t in trees:
        if t is not None:
            t.free()


def test_batch_mixed_semantics_and_flags(gpu_ctx):
    """One call with both semantics and both flag sets: every mesh builds (one
    set of passes per (semantics, flags), the larger sets split over lanes)
    and equals its plain build."""
    import svo_amd
    meshes = _meshes(80, seed=23)
    n = len(meshes)
    assert n >= 50
    # 36+ meshes of (reference, no flags): enough for two concurrent passes
    sems = [0 if i < 36 else i % 2 for i in range(n)]
    flags = [svo_amd.BUILD_NO_VERTEX_EDGE if i >= 36 and i % 3 == 2 else 0 for i in range(n)]
    print(f"mixed batch of {n} meshes", flush=True)
    trees, st = gpu_ctx.build_batch(meshes, semantics=sems, flags=flags)
    assert all(s == 0 for s in st), st
    for i, ((P, T, res), tb) in enumerate(zip(meshes, trees)):
        tp = gpu_ctx.build(P, T, res, semantics=sems[i], flags=flags[i])
        assert tb.info.semantics == sems[i]
        fields = ALL if flags[i] == 0 else [f for f in ALL if f not in ("vertices", "edges", "faces")]
        for f in fields:
            assert_same(f"mesh {i} {f}", tb.field(f), tp.field(f))
        if flags[i]:
            assert tb.info.n_vertices == 0 and tb.info.n_edges == 0
        tp.free()
        tb.free()


@pytest.mark.parametrize("path", [1, 2], ids=["bitmap", "sort"])
def test_batch_odd_resolutions(gpu_ctx, path):
    """Resolutions whose bitmap regions are no whole number of 16-byte pieces
    or level-1 tiles: mesh boundaries fall at odd word offsets inside tiles."""
    from svo_amd import meshgen
    P3, T3 = meshgen.icosphere(3)
    P4, T4 = meshgen.icosphere(4)
    meshes = [(P3, T3, 5), (P4, T4, 17), (P3, T3, 33), (P4, T4, 47), (P3, T3, 100),
              (P4, T4, (40, 24, 33)), (P3, T3, 3), (P4, T4, 129)]
    gpu_ctx.set_option("dedup_path", path)
    try:
        trees, st = gpu_ctx.build_batch(meshes)
    finally:
        gpu_ctx.set_option("dedup_path", 0)
    assert all(s == 0 for s in st), st
    for m, t in zip(meshes, trees):
        print(f"mesh res {m[2]}", flush=True)
        compare(t, po.build(*m))
        t.free()


def test_batch_single_voxel_meshes(gpu_ctx):
    # meshes whose voxel set is a single cell (depth-1 trees) next to normal ones
    tri = np.array([[0, 1, 2]], np.int32)
    tiny = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
    from svo_amd import meshgen
    P, T = meshgen.icosphere(3)
    meshes = [(tiny, tri, 1), (P, T, 16), (tiny, tri, 1), (P, T, 8), (tiny, tri, 2)]
    trees, st = gpu_ctx.build_batch(meshes)
    assert all(s == 0 for s in st), st
    for m, t in zip(meshes, trees):
        compare(t, po.build(*m))
        t.free()


def test_autotune_keeps_results_and_round_trips(gpu_ctx, tmp_path):
    """The tuning search picks options on a sample batch; trees built under the
    chosen options equal the default context's; options survive a save/load."""
    import svo_amd
    meshes = _meshes(36, seed=23)
    print(f"autotune on {len(meshes)} meshes", flush=True)
    ctx = svo_amd.Context(0)
    try:
        r = ctx.autotune(meshes, reps=1)
        assert r["n_trials"] > 1 and r["best_ms"] <= r["default_ms"] + 1e-9
        assert ctx.get_option("small_max_cells") == r["small_max_cells"]
        assert ctx.get_option("batch_lanes") == r["batch_lanes"]
        path = str(tmp_path / "tuning.txt")
        ctx.save_tuning(path)
        ctx2 = svo_amd.Context(0)
        try:
            ctx2.load_tuning(path)
            for name in ("small_max_cells", "batch_lanes", "dedup_path", "bitmap_max_bytes"):
                assert ctx2.get_option(name) == ctx.get_option(name)
        finally:
            ctx2.close()
        trees, st = ctx.build_batch(meshes)
        ref, st2 = gpu_ctx.build_batch(meshes)
        assert st == st2 and all(s == 0 for s in st)
        for i, (a, b) in enumerate(zip(trees, ref)):
            for f in ALL:
                assert_same(f"mesh {i} {f}", a.field(f), b.field(f))
            a.free()
            b.free()
    finally:
        ctx.close()


def test_sorted_batch_with_device_flagged_mesh(gpu_ctx):
    """A device-resident batch of more than 65,536 triangles (voxelized in
    size-sorted order) holding a mesh with an out-of-range index: that mesh
    fails with E_INVALID (found on the device, where its triangles are
    classed to walk nothing), every other mesh equals its plain build."""
    import svo_amd
    meshes = _meshes(36, seed=25, res_choices=(64, 128))
    assert sum(T.shape[0] for _, T, _ in meshes) > 65536
    k = 5
    P, T, _ = meshes[k]
    t_bad = T.copy()
    t_bad[T.shape[0] // 2, 1] = P.shape[0] + 7
    dms = [gpu_ctx.upload_mesh(P_, t_bad if i == k else T_)
           for i, (P_, T_, _) in enumerate(meshes)]
    try:
        print(f"sorted batch of {len(meshes)} device meshes, mesh {k} with a bad index",
              flush=True)
        trees, st = gpu_ctx.build_batch([d.desc(r) for d, (_, _, r) in zip(dms, meshes)],
                                        on_device=True)
        assert st[k] == svo_amd.E_INVALID and trees[k] is None
        assert all(s == 0 for i, s in enumerate(st) if i != k), st
        for i, ((P_, T_, res), tb) in enumerate(zip(meshes, trees)):
            if i == k:
                continue
            tp = gpu_ctx.build(P_, T_, res)
            for f in ("voxels", "code", "origin", "parent", "childs", "vertices", "edges"):
                assert_same(f"mesh {i} {f}", tb.field(f), tp.field(f))
            tp.free()
            tb.free()
    finally:
        for d in dms:
            d.free()
