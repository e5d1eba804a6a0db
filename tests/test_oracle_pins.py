"""Pins the oracle's grid set-up, edge endpoints, intended-mode records and the
reference's consecutive unique with checks that do not restate its code.

- Grid (BaseModel.h:24-49, 64-73; SVO.cu:348-351): hand-computed power-of-two
  boxes with each axis in turn the longest, and properties the padding fixes on
  random asymmetric meshes: every point strictly inside, every axis as wide as
  the longest one plus 2/10001 of it, shorter axes centred on the points, and
  unit = width / res per axis.  Points that no face uses still widen the box.
- Edge endpoints (SVO.cu:638-663): the 12 (offset, offset) pairs and the bare
  `origin` terms extracted from the reference's source text into
  tests/golden/luts.json; every edge record must be its node's origin plus
  those offsets times the width.
- Intended semantics: every vertex / edge / face record (coordinates, owner,
  order) against a brute force over the lattice points of each level.
- Reference semantics (SVO.cu:679-720): the consecutive thrust::unique does
  remove records when node corners collapse in float32 (a mesh far from the
  origin, o + w == o); the oracle's arrays must equal a consecutive unique
  applied here to the 8N / 12N records built from the node fields.
"""
import json
import os

import numpy as np
import pytest

import pyoracle as po

NONE = 0xFFFFFFFF
F32 = np.float32
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _luts():
    with open(os.path.join(GOLDEN, "luts.json")) as f:
        return json.load(f)


def _grid(P, res, T=None):
    P = np.asarray(P, np.float32)
    if T is None:
        T = np.array([[0, 1, 2]], np.int32)
    r = po.build(P, T, res, fields=("voxels",))
    assert r.status == 0
    i = r.info
    return tuple(np.array(getattr(i, k), np.float32)
                 for k in ("bbox_min", "bbox_max", "bbox_width", "unit"))


# ---------------------------------------------------------------- grid
@pytest.mark.parametrize("longest", [0, 1, 2])
def test_grid_hand_computed(longest):
    # spans 1, 1/2, 1/4 (rotated so that axis `longest` spans 1): the shorter
    # axes are padded by half the difference, -1/4..3/4 and -3/8..5/8, so every
    # axis spans exactly 1 before the 1/10001 padding
    span = np.roll(np.array([1.0, 0.5, 0.25]), longest)
    P = np.array([[0, 0, 0], span, [0.25 * span[0], 0.5 * span[1], 0]], np.float32)
    res = (16, 8, 4)
    bmin, bmax, width, unit = _grid(P, res)
    lo = np.roll(np.array([0.0, -0.25, -0.375], np.float32), longest)
    hi = np.roll(np.array([1.0, 0.75, 0.625], np.float32), longest)
    eps = F32(1.0) / F32(10001.0)
    assert np.array_equal(bmin, (lo - eps).astype(np.float32))
    assert np.array_equal(bmax, (hi + eps).astype(np.float32))
    assert np.array_equal(width, (bmax - bmin).astype(np.float32))
    assert np.array_equal(unit, (width / np.array(res, np.float32)).astype(np.float32))


def _ulp(x):
    return np.spacing(np.abs(np.asarray(x, np.float32))).astype(np.float64)


@pytest.mark.parametrize("seed", range(12))
def test_grid_properties(seed):
    rng = np.random.default_rng(seed)
    scale = rng.uniform(0.01, 10.0, 3)
    scale[seed % 3] *= 3.0  # each axis in turn the longest
    shift = rng.uniform(-50, 50, 3)
    P = (rng.random((40, 3)) * scale + shift).astype(np.float32)
    T = rng.integers(0, 40, (10, 3)).astype(np.int32)
    res = tuple(int(v) for v in rng.integers(1, 300, 3))
    bmin, bmax, width, unit = _grid(P, res, T)
    mn, mx = P.min(0), P.max(0)
    # every point strictly inside the padded cube
    assert np.all(bmin < mn) and np.all(bmax > mx)
    assert np.array_equal(width, (bmax - bmin).astype(np.float32))
    # every axis is as wide as the longest one plus 1/10001 of it on each side
    L = float(np.max((mx - mn).astype(np.float32)))
    want = L * (1.0 + 2.0 / 10001.0)
    assert np.all(np.abs(width.astype(np.float64) - want) <= 6 * _ulp(want)), (width, want)
    # shorter axes are centred on the points (the longest one trivially)
    c_box = (bmin.astype(np.float64) + bmax) / 2
    c_pts = (mn.astype(np.float64) + mx) / 2
    tol = 4 * _ulp(np.maximum(np.abs(bmin), np.abs(bmax)))
    assert np.all(np.abs(c_box - c_pts) <= tol), (c_box, c_pts)
    # voxel size per axis (SVO.cu:348-350)
    assert np.array_equal(unit, (width / np.array(res, np.float32)).astype(np.float32))


def test_grid_counts_unreferenced_points():
    # BaseModel's box is over all points (BaseModel.h:68-72), not only the
    # ones a face uses
    P = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [5, 3, 7]], np.float32)
    bmin, bmax, width, _ = _grid(P, 8)
    assert np.all(bmax > np.array([5, 3, 7], np.float32))
    assert width[0] > 7.0


# ---------------------------------------------------------------- edge endpoints
def _edge_table():
    d = _luts()
    ofs = np.array(d["edge_endpoints"], np.int64)  # [12][2][3]
    bare = np.array(d["edge_endpoint_bare"], bool)  # [12][2]
    return ofs, bare


def test_edge_table_shape():
    ofs, bare = _edge_table()
    assert ofs.shape == (12, 2, 3) and bare.shape == (12, 2)
    # each edge runs along one axis, one width long, and the 12 are distinct
    d = np.abs(ofs[:, 1] - ofs[:, 0]).sum(axis=1)
    assert np.all(d == 1)
    keys = {tuple(sorted(map(tuple, e))) for e in ofs.tolist()}
    assert len(keys) == 12
    # the bare `origin` term is exactly the zero offset
    assert np.array_equal(bare, (ofs == 0).all(axis=2))


def _edge_points(origin, width, ofs, bare):
    """[N, 12, 2, 3] float32 endpoints: origin (bare) or origin + offset * w."""
    o = origin.astype(np.float32)
    w = width.astype(np.float32)
    out = np.empty((o.shape[0], 12, 2, 3), np.float32)
    for k in range(12):
        for s in range(2):
            if bare[k, s]:
                out[:, k, s] = o
            else:
                out[:, k, s] = o + ofs[k, s].astype(np.float32)[None, :] * w[:, None]
    return out


def _owner_last_valid(nb, table):
    """SVO.cu:600-603 / 669-672: the last shared slot holding a node wins,
    else the node itself."""
    N = nb.shape[0]
    own = np.repeat(np.arange(N, dtype=np.uint32)[:, None], len(table), axis=1)
    for k, row in enumerate(table):
        for s in row:
            v = nb[:, s]
            own[:, k] = np.where(v != NONE, v, own[:, k])
    return own


def _meshes():
    from svo_amd import meshgen
    P, T = meshgen.icosphere(3)
    yield "ico3@32", P, T, 32
    P, T = meshgen.torus(0.4, 0.1, 40, 20)
    yield "torus@64", P, T, 64
    P, T = meshgen.soup(300, seed=4, median=1 / 32)
    yield "soup@128", P, T, 128
    P, T = meshgen.icosphere(2)
    yield "ico2@(20,13,7)", P, T, (20, 13, 7)


@pytest.mark.parametrize("case", range(4))
def test_reference_edges_follow_table(case):
    name, P, T, res = list(_meshes())[case]
    r = po.build(P, T, res)
    assert r.status == 0
    ofs, bare = _edge_table()
    N = r.info.n_nodes
    pts = _edge_points(r["origin"], r["width"], ofs, bare)
    e = r["edges"]
    assert e.size == 12 * N, name
    got = np.stack([e["a"], e["b"]], axis=1).reshape(N, 12, 2, 3)
    assert np.array_equal(got.view(np.uint32), pts.view(np.uint32)), name
    es = np.array(_luts()["edge_shared"]).reshape(12, 4)
    own = _owner_last_valid(r["neighbors"], es)
    assert np.array_equal(e["owner"].reshape(N, 12), own), name
    vs = np.array(_luts()["vert_shared"]).reshape(8, 8)
    assert np.array_equal(r["vertices"]["owner"].reshape(N, 8),
                          _owner_last_valid(r["neighbors"], vs)), name


# ---------------------------------------------------------------- intended records
def _decode(codes):
    codes = np.asarray(codes, dtype=np.uint64)
    out = np.zeros((codes.size, 3), np.int64)
    for i in range(21):
        for a in range(3):
            out[:, a] |= ((codes >> np.uint64(3 * i + a)) & np.uint64(1)).astype(np.int64) << i
    return out


def _brute_intended_records(r):
    """Vertex, edge and face records of intended semantics from lattice
    geometry: a feature of a level is kept once, by the smallest-index node of
    that level whose cell touches it, and written by that node in node order
    (corners 0-7, the reference's 12 edges, faces -x +x -y +y -z +z)."""
    ofs, bare = _edge_table()
    N = r.info.n_nodes
    off = np.concatenate([[0], np.cumsum(r["level_nodes"])]).astype(np.int64)
    level = np.searchsorted(off, np.arange(N), side="right") - 1
    coords = _decode(r["code"])
    cells = [dict() for _ in range(r.depth)]
    for t in range(N):
        cells[level[t]][tuple(coords[t])] = t
    root = N - 1

    def node_at(t, c):
        if r.depth >= 2 and t == root:
            return NONE  # the root's only neighbor is itself
        return cells[level[t]].get(c, NONE)

    def owner(t, touching):
        return min([t] + [node_at(t, c) for c in touching])

    o, w = r["origin"], r["width"]

    def point(t, d):
        if not any(d):
            return o[t].copy()
        return (o[t] + np.array(d, np.float32) * w[t]).astype(np.float32)

    V, E, Fc = [], [], []
    for t in range(N):
        x, y, z = coords[t]
        for j in range(8):
            b = (j & 1, (j >> 1) & 1, (j >> 2) & 1)
            touching = [(x + b[0] - 1 + i, y + b[1] - 1 + k, z + b[2] - 1 + q)
                        for q in (0, 1) for k in (0, 1) for i in (0, 1)]
            if owner(t, touching) == t:
                V.append((point(t, b), t))
        for k in range(12):
            a0, a1 = ofs[k, 0], ofs[k, 1]
            run = int(np.nonzero(a0 != a1)[0][0])
            other = [ax for ax in range(3) if ax != run]
            touching = []
            for d0 in (a0[other[0]] - 1, a0[other[0]]):
                for d1 in (a0[other[1]] - 1, a0[other[1]]):
                    c = [x, y, z]
                    c[other[0]] += d0
                    c[other[1]] += d1
                    touching.append(tuple(c))
            if owner(t, touching) == t:
                E.append((point(t, a0), point(t, a1), t))
        for axis in range(3):
            for side in (0, 1):
                c = [x, y, z]
                c[axis] += 2 * side - 1
                if owner(t, [tuple(c)]) == t:
                    lo = [0, 0, 0]
                    hi = [1, 1, 1]
                    lo[axis] = hi[axis] = side
                    Fc.append((point(t, lo), point(t, hi), t))
    return V, E, Fc


@pytest.mark.parametrize("case", [0, 1, 3])
def test_intended_records_bruteforce(case):
    name, P, T, res = list(_meshes())[case]
    r = po.build(P, T, res, semantics=po.SEM_INTENDED)
    assert r.status == 0
    V, E, Fc = _brute_intended_records(r)
    v = r["vertices"]
    assert v.size == len(V), name
    assert np.array_equal(v["p"].view(np.uint32), np.array([p for p, _ in V]).view(np.uint32))
    assert np.array_equal(v["owner"], np.array([t for _, t in V], np.uint32))
    for key, recs in (("edges", E), ("faces", Fc)):
        got = r[key]
        assert got.size == len(recs), (name, key)
        assert np.array_equal(got["a"].view(np.uint32),
                              np.array([a for a, _, _ in recs]).view(np.uint32)), (name, key)
        assert np.array_equal(got["b"].view(np.uint32),
                              np.array([b for _, b, _ in recs]).view(np.uint32)), (name, key)
        assert np.array_equal(got["owner"], np.array([t for _, _, t in recs], np.uint32))


# ---------------------------------------------------------------- reference unique
def far_mesh(offset=1.0e5, radius=0.25, subdiv=2):
    """An icosphere far from the origin: float32 spacing at `offset` is
    2^-7, so node corners o + w collapse onto o for widths below 2^-8."""
    from svo_amd import meshgen
    P, T = meshgen.icosphere(subdiv)
    return (P * np.float32(radius) + np.float32(offset)).astype(np.float32), T


def _consecutive_unique(keep_eq, n):
    keep = np.ones(n, bool)
    keep[1:] = ~keep_eq
    return keep


def _eq3(a, b):
    return (a == b).all(axis=-1)


@pytest.mark.parametrize("radius,res", [(0.25, 256), (0.05, 64), (0.02, 100)])
def test_reference_unique_removes_collapsed_records(radius, res):
    P, T = far_mesh(radius=radius)
    r = po.build(P, T, res)
    assert r.status == 0
    N = r.info.n_nodes
    ofs, bare = _edge_table()
    # 12N edge records of determineNodeEdge in node-major order
    pts = _edge_points(r["origin"], r["width"], ofs, bare).reshape(12 * N, 2, 3)
    a, b = pts[:, 0], pts[:, 1]
    same = (_eq3(a[1:], a[:-1]) & _eq3(b[1:], b[:-1])) | (_eq3(a[1:], b[:-1]) & _eq3(b[1:], a[:-1]))
    keep = _consecutive_unique(same, 12 * N)
    e = r["edges"]
    assert keep.sum() < 12 * N  # the unique does remove records here
    assert e.size == keep.sum()
    assert np.array_equal(e["a"].view(np.uint32), a[keep].view(np.uint32))
    assert np.array_equal(e["b"].view(np.uint32), b[keep].view(np.uint32))
    es = np.array(_luts()["edge_shared"]).reshape(12, 4)
    own = _owner_last_valid(r["neighbors"], es).reshape(-1)
    assert np.array_equal(e["owner"], own[keep])
    # corners w * (x, y, z) never collapse for w > 0 (SVO.cu:598 omits the origin)
    bits = np.array([[j & 1, (j >> 1) & 1, (j >> 2) & 1] for j in range(8)], np.float32)
    vp = (r["width"][:, None, None] * bits[None]).reshape(-1, 3).astype(np.float32)
    vsame = _eq3(vp[1:], vp[:-1])
    assert not vsame.any()
    assert np.array_equal(r["vertices"]["p"].view(np.uint32), vp.view(np.uint32))
