"""Pins the oracle's surface voxelization (SVO.cu:37-188) with exact arithmetic.

The reference's predicate -- triangle box clamped to the grid, plane test
(n.p + d1)(n.p + d2) <= 0 with the critical corner c, and the three 2-D
projection tests with inward edge normals and critical-point offsets -- is
evaluated here in rational arithmetic (fractions.Fraction) on the float32
vertex values.  The float32 restatement may disagree only at near-ties: every
disagreeing cell must have one of its test values within a tiny relative
distance of zero.  A dropped term, a wrong sign or swapped axes produces
disagreements far from any tie and fails the test.
"""
from fractions import Fraction as Fr

import numpy as np
import pytest

import pyoracle as po

TIE = Fr(1, 10 ** 4)  # relative to the voxel size


def _f(x):
    return Fr(float(x))


def exact_cells(v, u, res):
    """Cells hit by one triangle under the reference's predicate, exactly.
    Returns (hits, near): near[cell] is the smallest |test value| / scale."""
    v = [[_f(c) for c in p] for p in v]
    u = [_f(c) for c in u]
    e = [[v[1][a] - v[0][a] for a in range(3)], [v[2][a] - v[1][a] for a in range(3)],
         [v[0][a] - v[2][a] for a in range(3)]]
    n = [e[0][1] * e[1][2] - e[0][2] * e[1][1], e[0][2] * e[1][0] - e[0][0] * e[1][2],
         e[0][0] * e[1][1] - e[0][1] * e[1][0]]
    nn = sum(abs(c) for c in n) or Fr(1)
    lo = [min(p[a] for p in v) for a in range(3)]
    hi = [max(p[a] for p in v) for a in range(3)]
    g0 = [max(0, min(res[a] - 1, int(lo[a] / u[a]))) for a in range(3)]
    g1 = [max(0, min(res[a] - 1, int(hi[a] / u[a]))) for a in range(3)]
    c = [u[a] if n[a] > 0 else Fr(0) for a in range(3)]
    d1 = sum(n[a] * (c[a] - v[0][a]) for a in range(3))
    d2 = sum(n[a] * ((u[a] - c[a]) - v[0][a]) for a in range(3))
    planes = []  # (axis a, axis b, flip sign)
    for (a, b, flip_axis) in ((0, 1, 2), (1, 2, 0), (2, 0, 1)):
        sgn = -1 if n[flip_axis] < 0 else 1
        tests = []
        for i in range(3):
            ne = (sgn * -e[i][b], sgn * e[i][a])
            d = -(ne[0] * v[i][a] + ne[1] * v[i][b]) + max(Fr(0), u[a] * ne[0]) + max(Fr(0), u[b] * ne[1])
            scale = (abs(ne[0]) + abs(ne[1])) or Fr(1)
            tests.append((ne, d, scale))
        planes.append((a, b, tests))
    hits, near = set(), {}
    for z in range(g0[2], g1[2] + 1):
        for y in range(g0[1], g1[1] + 1):
            for x in range(g0[0], g1[0] + 1):
                p = [x * u[0], y * u[1], z * u[2]]
                t = sum(n[a] * p[a] for a in range(3))
                vals = [(t + d1) / nn, (t + d2) / nn]
                ok = (t + d1) * (t + d2) <= 0
                for a, b, tests in planes:
                    for ne, d, scale in tests:
                        val = ne[0] * p[a] + ne[1] * p[b] + d
                        vals.append(val / scale)
                        ok = ok and val >= 0
                # box-boundary ties of the clamped cell range
                for k in range(3):
                    vals.append(lo[k] - g0[k] * u[k])
                    vals.append((g1[k] + 1) * u[k] - hi[k])
                cell = (x, y, z)
                near[cell] = min(abs(w) for w in vals) / max(u)
                if ok:
                    hits.add(cell)
    return hits, near


def _case(P, T, res, max_bad_frac=0.02):
    st, codes = po.voxelize(P, T, res)
    assert st == 0
    res3 = (res,) * 3 if np.isscalar(res) else res
    info = po.build(P, T, res, fields=["voxels"]).info
    origin = np.array(info.bbox_min, np.float32)
    u = np.array(info.unit, np.float32)
    got = {po.morton_decode(int(c)) for c in codes}
    want, near = set(), {}
    for t in T:
        v = (P[t].astype(np.float32) - origin).astype(np.float32)  # SVO.cu:58-63 in float32
        h, nr = exact_cells(v, u, res3)
        want |= h
        for k, w in nr.items():
            near[k] = min(near.get(k, w), w)
    diff = got ^ want
    far = [c for c in diff if near.get(c, Fr(1)) > TIE]
    assert not far, f"{len(far)} disagreements away from ties, e.g. {sorted(far)[:5]}"
    assert len(diff) <= max(2, max_bad_frac * len(want)), (len(diff), len(want))
    return len(want), len(diff)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_random_triangles_res8(seed):
    rng = np.random.default_rng(seed)
    P = rng.random((30, 3)).astype(np.float32)
    T = np.arange(30, dtype=np.int32).reshape(10, 3)
    n, d = _case(P, T, 8)
    assert n > 10


def test_small_triangles_res16():
    from svo_amd import meshgen
    P, T = meshgen.soup(40, seed=9, median=1.0 / 8)
    _case(P, T, 16)


def test_icosphere_res12_noncubic():
    from svo_amd import meshgen
    P, T = meshgen.icosphere(1)
    _case(P, T, (12, 9, 7))


def test_axis_aligned_triangles():
    # triangles in axis planes: the bbox padding keeps them off the cell faces
    P = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [1, 1, 1]], np.float32)
    T = np.array([[0, 1, 2], [3, 4, 5], [0, 3, 1]], np.int32)
    _case(P, T, 8, max_bad_frac=0.2)  # hypotenuses through cell corners: exact ties


def test_every_vertex_voxel_hit():
    # the cell containing each triangle vertex is a hit (up to ties)
    from svo_amd import meshgen
    P, T = meshgen.icosphere(3)
    res = 32
    info = po.build(P, T, res, fields=["voxels"]).info
    st, codes = po.voxelize(P, T, res)
    got = set(int(c) for c in codes)
    origin = np.array(info.bbox_min, np.float32)
    u = np.array(info.unit, np.float32)
    missing = 0
    for p in P:
        v = (p - origin).astype(np.float32)
        cell = np.clip((v / u).astype(np.int64), 0, res - 1)
        if po.morton_encode_u16(*map(int, cell)) not in got:
            frac = (v / u) - np.floor(v / u)
            if not np.any((frac < 1e-4) | (frac > 1 - 1e-4)):
                missing += 1
    assert missing == 0
