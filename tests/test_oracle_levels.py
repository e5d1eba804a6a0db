"""Pins the oracle's octree levels, nodes, neighbors and vertex/edge records.

Reference semantics (SVO.cu:383-728) against a closed form written here with
numpy set operations: level l+1 codes are the distinct parents (code >> 3) of
level l; each distinct parent owns eight consecutive nodes (code = parent*8 +
k); occupied nodes above level 1 link to the eight nodes of their group one
level down; every node below the root links to the node holding its parent
code; origins are gridOrigin + width*coord in float32 with width doubling per
level; only root.neighbors[13] is stored; vertex / edge records are 8 and 12
per node, owned by the node.

Intended semantics against brute force by coordinates: a neighbor is the
same-level node at (x+dx, y+dy, z+dz); a vertex / edge / face is kept by the
smallest-index same-level node that has it.
"""
import numpy as np
import pytest

import pyoracle as po

NONE = 0xFFFFFFFF


def _decode_vec(codes):
    codes = np.asarray(codes, dtype=np.uint64)
    out = np.zeros((codes.size, 3), np.int64)
    for i in range(21):
        for a in range(3):
            out[:, a] |= ((codes >> np.uint64(3 * i + a)) & np.uint64(1)).astype(np.int64) << i
    return out


def closed_form(voxels, unit_x, origin):
    S = [np.asarray(voxels, dtype=np.uint64)]
    while S[-1].size > 1:
        S.append(np.unique(S[-1] >> np.uint64(3)))
    L = len(S)
    counts = [8 * S[i + 1].size for i in range(L - 1)] + [1]
    off = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
    N = int(off[-1])
    code = np.zeros(N, np.uint64)
    leaf = np.ones(N, np.uint8)
    width = np.zeros(N, np.float32)
    parent = np.full(N, NONE, np.uint32)
    childs = np.full((N, 8), NONE, np.uint32)
    w = np.float32(unit_x)
    widths = []
    for i in range(L):
        widths.append(w)
        w = np.float32(w * np.float32(2.0))
    for i in range(L - 1):
        P = S[i + 1]
        codes = ((P[:, None] << np.uint64(3)) | np.arange(8, dtype=np.uint64)).ravel()
        idx = off[i] + np.arange(codes.size)
        code[idx] = codes
        real = np.isin(codes, S[i])
        leaf[idx] = 1 if i == 0 else (~real).astype(np.uint8)
        width[idx] = widths[i]
        if i >= 1:
            rank = np.searchsorted(S[i], codes[real])
            childs[idx[real]] = (off[i - 1] + 8 * rank)[:, None] + np.arange(8)
        # parent: node holding P at level i+1
        if i + 1 == L - 1:
            par = np.full(P.size, off[L - 1])
        else:
            grp = np.searchsorted(S[i + 2], P >> np.uint64(3))
            par = off[i + 1] + 8 * grp + (P & np.uint64(7)).astype(np.int64)
        parent[idx] = np.repeat(par, 8).astype(np.uint32)
    root = N - 1
    code[root] = S[L - 1][0]
    leaf[root] = 1 if L == 1 else 0
    width[root] = widths[L - 1]
    if L >= 2:
        childs[root] = off[L - 2] + np.arange(8)
    coords = _decode_vec(code).astype(np.float32)
    org = (np.asarray(origin, np.float32)[None, :] + width[:, None] * coords).astype(np.float32)
    return dict(depth=L, code=code, is_leaf=leaf, origin=org, width=width, parent=parent,
                childs=childs, off=off)


def _meshes():
    from svo_amd import meshgen
    P, T = meshgen.icosphere(3)
    yield "ico3", P, T, 32
    P, T = meshgen.torus(0.4, 0.1, 40, 20)
    yield "torus", P, T, 64
    P, T = meshgen.soup(300, seed=4, median=1 / 32)
    yield "soup", P, T, 128
    P, T = meshgen.icosphere(2)
    yield "ico2-noncubic", P, T, (20, 13, 7)
    yield "ico2-res100", P, T, 100


def _check_reference(name, P, T, res):
    r = po.build(P, T, res)
    assert r.status == 0, name
    cf = closed_form(r["voxels"], r.info.unit[0], r.info.bbox_min)
    assert r.depth == cf["depth"]
    assert np.array_equal(r["level_nodes"], np.diff(cf["off"]).astype(np.uint64))
    for k in ("code", "is_leaf", "parent", "childs"):
        assert np.array_equal(np.asarray(r[k]).astype(cf[k].dtype), cf[k]), (name, k)
    for k in ("origin", "width"):
        assert np.array_equal(r[k].view(np.uint32), cf[k].view(np.uint32)), (name, k)
    N = r.info.n_nodes
    nb = r["neighbors"]
    expect = np.full((N, 27), NONE, np.uint32)
    if r.depth >= 2:
        expect[N - 1, 13] = N - 1
    assert np.array_equal(nb, expect)
    # 8 / 12 records per node, owned by the node (no adjacent duplicates here)
    v, e = r["vertices"], r["edges"]
    assert v.size == 8 * N and e.size == 12 * N
    assert np.array_equal(v["owner"], np.repeat(np.arange(N, dtype=np.uint32), 8))
    bits = np.array([[j & 1, (j >> 1) & 1, (j >> 2) & 1] for j in range(8)], np.float32)
    want_p = (r["width"][:, None, None] * bits[None]).reshape(-1, 3).astype(np.float32)
    assert np.array_equal(v["p"].view(np.uint32), want_p.view(np.uint32))
    assert np.array_equal(e["owner"], np.repeat(np.arange(N, dtype=np.uint32), 12))
    # AoS records equal the SoA fields
    aos = r["nodes_aos"]
    assert np.array_equal(aos["code"], r["code"]) and np.array_equal(aos["parent"], r["parent"])
    assert np.array_equal(aos["neighbors"], nb) and np.all(aos["pad"] == 0)


@pytest.mark.parametrize("case", list(range(5)))
def test_reference_closed_form(case):
    name, P, T, res = list(_meshes())[case]
    _check_reference(name, P, T, res)


def test_reference_bunny(bunny):
    P, T = bunny
    _check_reference("bunny", P, T, 64)


def test_single_voxel_tree():
    P = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
    T = np.array([[0, 1, 2]], np.int32)
    r = po.build(P, T, 1)
    assert r.status == 0 and r.depth == 1 and r.info.n_nodes == 1
    assert r["is_leaf"][0] == 1 and r["parent"][0] == NONE
    assert np.all(r["neighbors"] == NONE)


def _brute_intended(r):
    N = r.info.n_nodes
    off = np.concatenate([[0], np.cumsum(r["level_nodes"])]).astype(np.int64)
    coords = _decode_vec(r["code"])
    level = np.searchsorted(off, np.arange(N), side="right") - 1
    maps = [dict() for _ in range(r.depth)]
    for t in range(N):
        maps[level[t]][tuple(coords[t])] = t
    nb = np.full((N, 27), NONE, np.uint32)
    for t in range(N):
        x, y, z = coords[t]
        m = maps[level[t]]
        for s in range(27):
            d = (s % 3 - 1, (s // 3) % 3 - 1, s // 9 - 1)
            nb[t, s] = m.get((x + d[0], y + d[1], z + d[2]), NONE)
    if r.depth >= 2:
        nb[N - 1] = NONE
        nb[N - 1, 13] = N - 1
    else:
        nb[:] = NONE

    def owner(t, cells):
        m = maps[level[t]]
        return min(m.get(c, NONE) for c in cells)

    nv = ne = nf = 0
    for t in range(N):
        x, y, z = coords[t]
        for j in range(8):
            b = (j & 1, (j >> 1) & 1, (j >> 2) & 1)
            cells = [(x + b[0] - 1 + i, y + b[1] - 1 + k, z + b[2] - 1 + q)
                     for q in (0, 1) for k in (0, 1) for i in (0, 1)]
            nv += owner(t, cells) == t
        for axis in range(3):  # edges along `axis` at the 4 corners of the other two axes
            o = [a for a in range(3) if a != axis]
            for c0 in (0, 1):
                for c1 in (0, 1):
                    cells = []
                    for d0 in (c0 - 1, c0):
                        for d1 in (c1 - 1, c1):
                            c = [x, y, z]
                            c[o[0]] += d0
                            c[o[1]] += d1
                            cells.append(tuple(c))
                    ne += owner(t, cells) == t
        for axis in range(3):
            for side in (-1, 1):
                c = [x, y, z]
                c[axis] += side
                nf += min(t, maps[level[t]].get(tuple(c), NONE)) == t
    return nb, nv, ne, nf


@pytest.mark.parametrize("case", [0, 1, 3])
def test_intended_bruteforce(case):
    name, P, T, res = list(_meshes())[case]
    r = po.build(P, T, res, semantics=po.SEM_INTENDED)
    assert r.status == 0
    nb, nv, ne, nf = _brute_intended(r)
    assert np.array_equal(r["neighbors"], nb), name
    assert (r.info.n_vertices, r.info.n_edges, r.info.n_faces) == (nv, ne, nf)
    # every kept record is owned by its writer and records of one node are adjacent
    for k in ("vertices", "edges", "faces"):
        own = r[k]["owner"]
        assert np.all(np.diff(own.astype(np.int64)) >= 0)
