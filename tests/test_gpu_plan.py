"""Batch plans (svo_plan_*): the batched build recorded once and replayed as one
HIP graph.  Every replay must equal the eager batched build bit for bit, with
eager work and synchronizations of the program between replays; a replay
over meshes whose sizes changed must fall back to an eager build (and say
so), and replay again once the meshes are back."""
import ctypes as C

import numpy as np
import pytest

from parity import assert_same

pytestmark = pytest.mark.gpu

ALL = ["voxels", "code", "is_leaf", "origin", "width", "parent", "childs", "neighbors",
       "vertices", "edges", "faces", "level_nodes", "nodes_aos"]


def _meshes(n, seed, res_choices=(64, 128, 256), max_tris=40000):
    from svo_amd import meshgen
    out = []
    for s in meshgen.mixed_specs(n, seed=seed, res_choices=res_choices):
        if s.n_tris > max_tris:
            continue
        P, T = s.build()
        out.append((P, T, s.res))
    return out


def _fields(trees):
    return [{f: t.field(f) for f in ALL} if t else None for t in trees]


def _same(tag, got, want):
    assert len(got) == len(want)
    for i, (g, w) in enumerate(zip(got, want)):
        assert (g is None) == (w is None), (tag, i)
        if g is None:
            continue
        for f in ALL:
            assert_same(f"{tag} mesh {i} {f}", g[f], w[f])


@pytest.mark.parametrize("sem", [0, 1])
def test_replays_equal_eager(gpu_ctx, sem):
    meshes = _meshes(48, seed=51)
    print(f"plan of {len(meshes)} host meshes, semantics {sem}", flush=True)
    eager, st = gpu_ctx.build_batch(meshes, semantics=sem)
    assert all(s == 0 for s in st), st
    want = _fields(eager)
    for t in eager:
        t.free()
    plan = gpu_ctx.plan(meshes, semantics=sem)
    try:
        for rep in range(3):
            trees, st, replayed = plan.run()
            print(f"replay {rep}: replayed {replayed}", flush=True)
            assert replayed and all(s == 0 for s in st), st
            _same(f"replay {rep}", _fields(trees), want)
            for t in trees:
                t.free()
            # the program's own work between replays
            P, T, res = meshes[rep]
            gpu_ctx.build(P, T, res, semantics=sem).free()
    finally:
        plan.close()


def test_replay_of_changed_meshes_falls_back(gpu_ctx):
    import svo_amd
    L = svo_amd.lib()
    meshes = _meshes(40, seed=52, res_choices=(64, 128))
    dms = [gpu_ctx.upload_mesh(P, T) for P, T, _ in meshes]
    descs = [d.desc(r) for d, (_, _, r) in zip(dms, meshes)]
    try:
        eager, st = gpu_ctx.build_batch(descs, on_device=True)
        want = _fields(eager)
        for t in eager:
            t.free()
        plan = gpu_ctx.plan(descs, on_device=True)
        try:
            trees, st, replayed = plan.run()
            assert replayed
            _same("first replay", _fields(trees), want)
            for t in trees:
                t.free()
            # squash mesh 3 along z in place: its voxel and node counts change
            P3, T3, r3 = meshes[3]
            Q = P3.copy()
            Q[:, 2] *= np.float32(0.25)
            Q = np.ascontiguousarray(Q, np.float32)
            svo_amd._check(L.svo_device_copy(gpu_ctx.handle, C.c_void_p(dms[3].points_ptr),
                                             Q.ctypes.data, Q.nbytes, 0), "upload")
            changed, st = gpu_ctx.build_batch(descs, on_device=True)
            want_changed = _fields(changed)
            for t in changed:
                t.free()
            assert want_changed[3]["voxels"].size != want[3]["voxels"].size
            trees, st, replayed = plan.run()
            print(f"changed mesh: replayed {replayed}", flush=True)
            assert not replayed and all(s == 0 for s in st)
            _same("fallback", _fields(trees), want_changed)
            for t in trees:
                t.free()
            # restored: the recorded sizes hold again and the graph replays
            P3c = np.ascontiguousarray(P3, np.float32)
            svo_amd._check(L.svo_device_copy(gpu_ctx.handle, C.c_void_p(dms[3].points_ptr),
                                             P3c.ctypes.data, P3c.nbytes, 0), "upload")
            trees, st, replayed = plan.run()
            assert replayed
            _same("restored replay", _fields(trees), want)
            for t in trees:
                t.free()
        finally:
            plan.close()
    finally:
        for d in dms:
            d.free()


def test_single_mesh_plan(gpu_ctx, bunny):
    P, T = bunny
    eager = gpu_ctx.build(P, T, 128)
    want = _fields([eager])
    eager.free()
    plan = gpu_ctx.plan([(P, T, 128)])
    try:
        for rep in range(2):
            trees, st, replayed = plan.run()
            assert replayed and st == [0]
            _same(f"bunny replay {rep}", _fields(trees), want)
            trees[0].free()
    finally:
        plan.close()


def test_plan_rejects_mixed_semantics(gpu_ctx):
    import svo_amd
    meshes = _meshes(4, seed=53, res_choices=(32,))
    with pytest.raises(svo_amd.SvoError) as e:
        gpu_ctx.plan(meshes, semantics=[0, 1, 0, 1][:len(meshes)])
    assert e.value.status == svo_amd.E_INVALID


def test_plan_with_64bit_keys_and_no_records(gpu_ctx):
    """64-bit keys (composite keys of a batch past 32 bits) and builds without
    vertex / edge records replay bit-exact as well."""
    import svo_amd
    meshes = _meshes(20, seed=54, res_choices=(128, 256))
    for kb, flags in ((64, 0), (0, svo_amd.BUILD_NO_VERTEX_EDGE)):
        print(f"plan key_bits {kb} flags {flags}", flush=True)
        eager, st = gpu_ctx.build_batch(meshes, key_bits=kb, flags=flags)
        assert all(s == 0 for s in st), st
        want = _fields(eager)
        for t in eager:
            t.free()
        plan = gpu_ctx.plan(meshes, key_bits=kb, flags=flags)
        try:
            for rep in range(2):
                trees, st, replayed = plan.run()
                assert replayed and all(s == 0 for s in st), st
                _same(f"kb {kb} flags {flags} replay {rep}", _fields(trees), want)
                for t in trees:
                    t.free()
        finally:
            plan.close()
