"""One context shared by several host threads: builds and batches issued at
the same time (ctypes releases the GIL during each call) take the context's
lanes as they free up, and every tree equals the same mesh built alone."""
import threading

import pytest

from parity import assert_same

pytestmark = pytest.mark.gpu

FIELDS = ["voxels", "code", "is_leaf", "origin", "width", "parent", "childs", "neighbors",
          "vertices", "edges", "level_nodes"]


def _meshes():
    from svo_amd import meshgen
    out = []
    for s in meshgen.mixed_specs(24, seed=41, res_choices=(32, 64, 128)):
        if s.n_tris > 20480:
            continue
        P, T = s.build()
        out.append((P, T, s.res))
    return out


def test_threads_share_a_context(gpu_ctx):
    meshes = _meshes()
    n = len(meshes)
    assert n >= 12
    want = []
    for P, T, res in meshes:
        t = gpu_ctx.build(P, T, res)
        want.append({f: t.field(f) for f in FIELDS})
        t.free()
    got = [None] * n
    errors = []

    def singles(idx):
        try:
            for i in idx:
                P, T, res = meshes[i]
                t = gpu_ctx.build(P, T, res)
                got[i] = {f: t.field(f) for f in FIELDS}
                t.free()
        except Exception as e:  # reported by the main thread
            errors.append(repr(e))

    def batch(idx):
        try:
            trees, st = gpu_ctx.build_batch([meshes[i] for i in idx])
            assert all(s == 0 for s in st), st
            for i, t in zip(idx, trees):
                got[i] = {f: t.field(f) for f in FIELDS}
                t.free()
        except Exception as e:
            errors.append(repr(e))

    half = n // 2
    jobs = [threading.Thread(target=singles, args=(list(range(0, half, 2)),)),
            threading.Thread(target=singles, args=(list(range(1, half, 2)),)),
            threading.Thread(target=batch, args=(list(range(half, n, 2)),)),
            threading.Thread(target=batch, args=(list(range(half + 1, n, 2)),))]
    print(f"{n} meshes over 4 threads", flush=True)
    for j in jobs:
        j.start()
    for j in jobs:
        j.join(timeout=100)
    assert not any(j.is_alive() for j in jobs), "a thread did not finish"
    assert not errors, errors
    for i in range(n):
        for f in FIELDS:
            assert_same(f"mesh {i} {f}", got[i][f], want[i][f])
