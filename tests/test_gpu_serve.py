"""Served path: continuous batching under a device-memory budget.

Requests that arrive while others are on the device are admitted as lanes
and memory free up; every result must equal the same mesh built alone,
bit for bit, and resident trees answer neighbor queries.
"""
import time

import numpy as np
import pytest

from parity import assert_same

pytestmark = pytest.mark.gpu

FIELDS = ["voxels", "code", "is_leaf", "origin", "width", "parent", "childs", "neighbors",
          "vertices", "edges"]


def _requests(n, seed):
    from svo_amd import meshgen
    out = []
    for s in meshgen.mixed_specs(n, seed=seed, res_choices=(32, 64, 128)):
        if s.n_tris > 20480:
            continue
        P, T = s.build()
        out.append((P, T, s.res))
    return out


def _drain(srv, n, timeout_s=120):
    done = {}
    t0 = time.time()
    while len(done) < n and time.time() - t0 < timeout_s:
        for d in srv.poll(cap=64, timeout_us=20000):
            done[d.request_id] = d
    assert len(done) == n, f"{len(done)} of {n} requests completed"
    return done


@pytest.mark.parametrize("sem", [0, 1, "mixed"])
def test_served_equals_alone(gpu_ctx, sem):
    # "mixed": requests alternate the two semantics, so batches hold both
    import svo_amd
    reqs = _requests(48, seed=31)
    sems = [(i % 2 if sem == "mixed" else sem) for i in range(len(reqs))]
    srv = svo_amd.Server(gpu_ctx, mem_budget=32 << 30, max_batch=8, n_lanes=2, max_wait_us=300)
    try:
        ids = []
        for i, (P, T, res) in enumerate(reqs):
            ids.append(srv.submit(P, T, res, semantics=sems[i]))
            if i % 7 == 0:
                time.sleep(0.002)  # arrivals spread over time
        done = _drain(srv, len(ids))
    finally:
        srv.close()
    sizes = [done[i].batch_size for i in ids]
    print(f"batch sizes {sorted(set(sizes))}", flush=True)
    assert max(sizes) > 1  # requests were batched together
    for rid, (P, T, res), sm in zip(ids, reqs, sems):
        d = done[rid]
        assert d.status == 0
        assert d.tree.info.semantics == sm
        alone = gpu_ctx.build(P, T, res, semantics=sm)
        for f in FIELDS:
            assert_same(f"request {rid} {f}", d.tree.field(f), alone.field(f))
        # neighbor queries on the resident tree
        rng = np.random.default_rng(rid)
        n = d.tree.info.n_nodes
        nodes = rng.integers(0, n, 200).astype(np.uint32)
        slots = rng.integers(0, 27, 200).astype(np.uint8)
        got = d.tree.query_neighbors(nodes, slots)
        assert np.array_equal(got, d.tree.field("neighbors")[nodes, slots])
        assert d.t_submit <= d.t_start <= d.t_done
        alone.free()
        d.tree.free()


def test_served_budget_admission(gpu_ctx):
    # a budget that fits only a few requests at a time still completes all of
    # them while the caller keeps freeing finished trees
    import svo_amd
    reqs = _requests(12, seed=32)
    est = max(svo_amd.estimate_bytes_mesh(P, T, r) for P, T, r in reqs)
    base = gpu_ctx.bytes_in_use()
    srv = svo_amd.Server(gpu_ctx, mem_budget=base + 3 * est + (64 << 20), max_batch=4, n_lanes=2,
                         max_wait_us=100)
    try:
        for P, T, res in reqs:
            srv.submit(P, T, res)
        got = 0
        t0 = time.time()
        while got < len(reqs) and time.time() - t0 < 120:
            for d in srv.poll(cap=16, timeout_us=20000):
                assert d.status == 0
                d.tree.free()
                got += 1
        assert got == len(reqs)
    finally:
        srv.close()


def test_server_rejects_invalid_and_full_queue(gpu_ctx):
    import svo_amd
    srv = svo_amd.Server(gpu_ctx, max_batch=2, max_queue=2, n_lanes=1, max_wait_us=200000)
    try:
        P = np.random.default_rng(0).random((3, 3)).astype(np.float32)
        T = np.array([[0, 1, 2]], np.int32)
        with pytest.raises(svo_amd.SvoError) as e:
            srv.submit(P, T, 4096, key_bits=32)
        assert e.value.status == svo_amd.E_RES_LIMIT
        srv.submit(P, T, 8)
        srv.submit(P, T, 8)
        # the worker may already hold the first two; fill until the queue is full
        with pytest.raises(svo_amd.SvoError) as e:
            for _ in range(8):
                srv.submit(P, T, 8)
        assert e.value.status == svo_amd.E_BUSY
        for d in _drain_some(srv):
            if d.tree:
                d.tree.free()
    finally:
        srv.close()


def _drain_some(srv, timeout_s=10):
    out = []
    t0 = time.time()
    while time.time() - t0 < timeout_s:
        got = srv.poll(cap=16, timeout_us=50000)
        out += got
        if not got and srv.stats()["pending"] == 0 and srv.stats()["in_flight"] == 0:
            break
    return out


def test_estimate_bounds_device_bytes():
    """svo_estimate_bytes_mesh is at least what a build takes on a fresh
    context (its tree plus the lane's workspace, allocator rounding
    included), across res 64-512 meshes of the served mix, a large noisy
    icosphere and the 2,000,000-triangle soup at res 1024."""
    import svo_amd
    from svo_amd import meshgen
    cases = []
    for s in meshgen.mixed_specs(24, seed=41, res_choices=(64, 128, 256, 512))[::3]:
        P, T = s.build()
        cases.append((f"{s.kind}{s.params} res {s.res}", P, T, s.res, 0))
    P, T = meshgen.noisy_icosphere(6, seed=1)
    cases.append(("icosphere-6 + noise res 512 intended", P, T, 512, 1))
    P, T = meshgen.icosphere(2)
    cases.append(("icosphere-2 res 512", P, T, 512, 0))
    P, T = meshgen.soup(2_000_000, seed=2)
    cases.append(("soup 2M res 1024", P, T, 1024, 0))
    for name, P, T, res, sem in cases:
        bound = svo_amd.estimate_bytes_mesh(P, T, res, semantics=sem)
        ctx = svo_amd.Context(0)
        try:
            base = ctx.bytes_in_use()
            print(f"estimate: {name}", flush=True)
            tree = ctx.build(P, T, res, semantics=sem)
            used = ctx.bytes_in_use() - base
            print(f"  bound {bound / 1e9:.3f} GB, used {used / 1e9:.3f} GB "
                  f"({bound / used:.2f}x)", flush=True)
            assert bound >= used, name
            tree.free()
        finally:
            ctx.close()


def test_served_with_context_budget_completes():
    """A context budget that holds only a few trees: passes of split batches
    run out of memory and are retried (or complete with SVO_E_OOM); every
    request completes, and once the trees are freed the context holds only
    its lanes' workspaces (no tree of a failed pass leaks)."""
    import svo_amd
    reqs = _requests(64, seed=33)
    est = max(svo_amd.estimate_bytes_mesh(P, T, r) for P, T, r in reqs)
    ctx = svo_amd.Context(0, mem_budget=6 * est)
    try:
        srv = svo_amd.Server(ctx, max_batch=40, n_lanes=2, max_wait_us=2000)
        try:
            for P, T, res in reqs:
                srv.submit(P, T, res)
            done = _drain(srv, len(reqs))
        finally:
            srv.close()
        st = [d.status for d in done.values()]
        print(f"statuses {sorted(set(st))}, ok {st.count(0)} of {len(st)}", flush=True)
        assert set(st) <= {0, svo_amd.E_OOM}
        assert st.count(0) > 0
        for d in done.values():
            if d.tree:
                d.tree.free()
        print(f"in use {ctx.bytes_in_use()}, workspaces {ctx.workspace_bytes()}", flush=True)
        assert ctx.bytes_in_use() == ctx.workspace_bytes()
    finally:
        ctx.close()
