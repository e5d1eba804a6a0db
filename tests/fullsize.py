"""Checks of full-size trees (hundreds of millions of nodes) against their
voxel set, in chunks of nodes so that host memory stays bounded.

The expected node fields are the closed form of the reference's level
construction (SVO.cu:383-495, createNode_1/2 SVO.cu:237-309, createRemainNode
SVO.cu:311-339), the same closed form `test_oracle_levels.closed_form` checks
the oracle against, evaluated here for any range of node indices:

- level l+1 holds the distinct parents (code >> 3) of level l, sorted;
- level l's nodes are eight per level-(l+1) code p, codes 8p + k, in order;
  the root level is one node;
- a node at level l >= 2 whose code is in level l is not a leaf and links to
  the eight nodes of its group one level down; every node below the root
  links to the node holding its parent code;
- origin = gridOrigin + width * decode(code) per axis in float32 (multiply,
  then add), width = unit.x doubled per level.

`tests/test_fullsize_helpers.py` pins these functions against
`closed_form` and the oracle on small meshes.
"""
from __future__ import annotations

import numpy as np

NONE = 0xFFFFFFFF

# 9-bit Morton chunk -> three 3-bit coordinates
_LX = np.zeros(512, np.uint64)
_LY = np.zeros(512, np.uint64)
_LZ = np.zeros(512, np.uint64)
for _v in range(512):
    for _b in range(3):
        _LX[_v] |= np.uint64(((_v >> (3 * _b)) & 1) << _b)
        _LY[_v] |= np.uint64(((_v >> (3 * _b + 1)) & 1) << _b)
        _LZ[_v] |= np.uint64(((_v >> (3 * _b + 2)) & 1) << _b)


_POPC = np.array([bin(v).count("1") for v in range(256)], np.int64)


def decode(codes: np.ndarray) -> np.ndarray:
    """Morton codes (up to 63 bits) -> int64 [n, 3] coordinates."""
    c = np.asarray(codes, dtype=np.uint64)
    x = np.zeros(c.size, np.uint64)
    y = np.zeros(c.size, np.uint64)
    z = np.zeros(c.size, np.uint64)
    top = int(c.max()).bit_length() if c.size else 0
    for k in range((top + 8) // 9):
        idx = ((c >> np.uint64(9 * k)) & np.uint64(511)).astype(np.intp)
        sh = np.uint64(3 * k)
        x |= _LX[idx] << sh
        y |= _LY[idx] << sh
        z |= _LZ[idx] << sh
    return np.stack([x, y, z], axis=1).astype(np.int64)


def _unique_sorted(x: np.ndarray) -> np.ndarray:
    if x.size == 0:
        return x
    m = np.empty(x.size, bool)
    m[0] = True
    np.not_equal(x[1:], x[:-1], out=m[1:])
    return x[m]


class Levels:
    """Level sets, node offsets and widths of the tree over `voxels`."""

    def __init__(self, voxels: np.ndarray, unit_x: float, origin):
        v = np.asarray(voxels, dtype=np.uint64)
        assert v.size > 0 and np.all(v[1:] > v[:-1]), "voxels must be sorted and unique"
        S = [v]
        while S[-1].size > 1:
            S.append(_unique_sorted(S[-1] >> np.uint64(3)))
        self.S = S
        self.L = len(S)
        counts = [8 * S[i + 1].size for i in range(self.L - 1)] + [1]
        self.off = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
        self.N = int(self.off[-1])
        # per level i >= 1 and each of its groups (a distinct code >> 3, one
        # code of level i + 1): index of the group's first code in S[i], and
        # the mask of the child slots k present; per level i >= 1, the group
        # of each code (its rank in level i + 1)
        self.first = [None] * self.L
        self.mask = [None] * self.L
        self.group = [None] * self.L
        for i in range(1, self.L - 1):
            x = S[i] >> np.uint64(3)
            m = np.empty(x.size, bool)
            m[0] = True
            np.not_equal(x[1:], x[:-1], out=m[1:])
            self.first[i] = np.flatnonzero(m)
            bit = (np.uint8(1) << (S[i] & np.uint64(7)).astype(np.uint8)).astype(np.uint8)
            self.mask[i] = np.bitwise_or.reduceat(bit, self.first[i])
            self.group[i] = np.cumsum(m, dtype=np.int64) - 1
        w = np.float32(unit_x)
        self.widths = []
        for _ in range(self.L):
            self.widths.append(w)
            w = np.float32(w * np.float32(2.0))
        self.origin = np.asarray(origin, np.float32)

    @property
    def depth(self) -> int:
        return self.L

    def level_nodes(self) -> np.ndarray:
        return np.diff(self.off).astype(np.uint64)

    def rows(self, start: int, count: int, want_origin: bool = True) -> dict:
        """Expected fields of nodes [start, start + count)."""
        end = min(self.N, start + count)
        n = end - start
        code = np.zeros(n, np.uint64)
        leaf = np.ones(n, np.uint8)
        width = np.zeros(n, np.float32)
        parent = np.full(n, NONE, np.uint32)
        childs = np.full((n, 8), NONE, np.uint32)
        S, off, L = self.S, self.off, self.L
        for i in range(L):
            a, b = max(start, int(off[i])), min(end, int(off[i + 1]))
            if a >= b:
                continue
            sl = slice(a - start, b - start)
            width[sl] = self.widths[i]
            if i == L - 1:  # the root
                code[sl] = S[L - 1][0]
                leaf[sl] = 1 if L == 1 else 0
                if L >= 2:
                    childs[sl] = off[L - 2] + np.arange(8)
                continue
            j = np.arange(a, b, dtype=np.int64) - off[i]
            g = j >> 3
            k = (j & 7).astype(np.uint8)
            P = S[i + 1][g]
            code[sl] = (P << np.uint64(3)) | k.astype(np.uint64)
            if i >= 1:
                mk = self.mask[i][g]
                real = ((mk >> k) & 1).astype(bool)
                leaf[sl] = (~real).astype(np.uint8)
                below = _POPC[mk & ((np.uint8(1) << k) - np.uint8(1)).astype(np.uint8)]
                pos = self.first[i][g[real]] + below[real]
                rows = np.nonzero(real)[0] + (a - start)
                childs[rows] = (off[i - 1] + 8 * pos)[:, None] + np.arange(8)
            if i + 1 == L - 1:
                parent[sl] = np.uint32(off[L - 1])
            else:
                par = off[i + 1] + 8 * self.group[i + 1][g] + (P & np.uint64(7)).astype(np.int64)
                parent[sl] = par.astype(np.uint32)
        out = dict(code=code, is_leaf=leaf, width=width, parent=parent, childs=childs)
        if want_origin:
            coords = decode(code).astype(np.float32)
            out["origin"] = (self.origin[None, :] + width[:, None] * coords).astype(np.float32)
        return out


def edge_points(origin, width, ofs, bare):
    """Reference edge endpoints (SVO.cu:638-663): [n, 12, 2, 3] float32, each
    the node origin (bare term) or origin + offset * width."""
    o = np.asarray(origin, np.float32)
    w = np.asarray(width, np.float32)
    out = np.empty((o.shape[0], 12, 2, 3), np.float32)
    for k in range(12):
        for s in range(2):
            if bare[k, s]:
                out[:, k, s] = o
            else:
                out[:, k, s] = o + ofs[k, s].astype(np.float32)[None, :] * w[:, None]
    return out


def vertex_points(width):
    """Reference vertex positions (SVO.cu:586-598): width * corner bits, no origin."""
    bits = np.array([[j & 1, (j >> 1) & 1, (j >> 2) & 1] for j in range(8)], np.float32)
    w = np.asarray(width, np.float32)
    return (w[:, None, None] * bits[None]).astype(np.float32)
