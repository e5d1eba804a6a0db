"""The oracle's Morton encode / decode and lookup tables against the values
printed in the reference's headers (tests/golden/luts.json, written by
tests/golden/make_golden.py from MortonLUT.h and SVO.cu)."""
import json
import os

import numpy as np
import pytest

import pyoracle as po

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "luts.json")


@pytest.fixture(scope="module")
def golden():
    with open(GOLDEN) as f:
        return json.load(f)


@pytest.fixture(scope="module")
def luts():
    return po.luts()


def test_host_and_device_tables_agree(golden):
    # MortonLUT.h prints host (h_) and device (d_) copies of each table
    for k in ("enc_x", "enc_y", "enc_z", "dec_x", "dec_y", "dec_z"):
        assert golden[k] == golden["d_" + k], k


@pytest.mark.parametrize("k", ["enc_x", "enc_y", "enc_z", "dec_x", "dec_y", "dec_z"])
def test_morton_tables(golden, luts, k):
    assert list(luts[k]) == golden[k]


@pytest.mark.parametrize("k,shape", [("nb_parent", (8, 27)), ("nb_child", (8, 27)),
                                     ("vert_shared", (8, 8)), ("edge_shared", (12, 4))])
def test_neighbor_tables(golden, luts, k, shape):
    assert luts[k].shape == shape
    assert np.array_equal(luts[k].ravel(), np.array(golden[k]))


def _interleave(x, y, z, bits):
    r = 0
    for i in range(bits):
        r |= ((x >> i) & 1) << (3 * i) | ((y >> i) & 1) << (3 * i + 1) | ((z >> i) & 1) << (3 * i + 2)
    return r


def _reference_encode(golden, x, y, z):
    """mortonEncode_LUT(uint16, uint16, uint16), MortonLUT.h:482-501, on the printed tables."""
    ex, ey, ez = golden["enc_x"], golden["enc_y"], golden["enc_z"]
    ans = 0
    for shift in (8, 0):
        ans = ((ans << 24) | ez[(z >> shift) & 0xFF] | ey[(y >> shift) & 0xFF] |
               ex[(x >> shift) & 0xFF]) & 0xFFFFFFFF
    return ans


def _reference_decode(golden, m):
    """morton3D_32_decode, MortonLUT.h:503-529, on the printed tables (7 chunks)."""
    out = []
    for k in ("dec_x", "dec_y", "dec_z"):
        lut = golden[k]
        a = 0
        for i in range(7):
            a |= lut[(m >> (i * 9)) & 0x1FF] << (3 * i)
        out.append(a & 0xFFFF)
    return tuple(out)


def test_encode_matches_reference_tables_and_closed_form(golden):
    rng = np.random.default_rng(0)
    pts = rng.integers(0, 1024, size=(2000, 3))
    corners = [(0, 0, 0), (1023, 1023, 1023), (1, 0, 0), (0, 1, 0), (0, 0, 1), (512, 256, 128)]
    for x, y, z in list(map(tuple, pts)) + corners:
        x, y, z = int(x), int(y), int(z)
        want = _reference_encode(golden, x, y, z)
        assert po.morton_encode_u16(x, y, z) == want
        assert _interleave(x, y, z, 10) == want
        assert po.morton_encode_u64(x, y, z) == want


def test_u32_encode_truncates_like_reference(golden):
    # at coordinate 1024 the y bit reaches the flag bit 31 and z's is dropped
    for x, y, z in [(1024, 0, 0), (0, 1024, 0), (0, 0, 1024), (2047, 2047, 2047)]:
        assert po.morton_encode_u16(x, y, z) == _reference_encode(golden, x, y, z)
    assert po.morton_encode_u16(0, 1024, 0) == 1 << 31
    assert po.morton_encode_u16(0, 0, 1024) == 0


def test_u64_encode_closed_form():
    rng = np.random.default_rng(1)
    for x, y, z in rng.integers(0, 1 << 21, size=(500, 3)):
        x, y, z = int(x), int(y), int(z)
        assert po.morton_encode_u64(x, y, z) == _interleave(x, y, z, 21)


def test_decode_matches_reference_tables(golden):
    rng = np.random.default_rng(2)
    codes = list(rng.integers(0, 1 << 30, size=2000)) + [0, 1, 7, (1 << 30) - 1]
    for m in codes:
        m = int(m)
        assert po.morton_decode(m) == _reference_decode(golden, m)
        x, y, z = po.morton_decode(m)
        assert po.morton_encode_u16(x, y, z) == m


def test_neighbor_lut_semantics(luts):
    # child k's neighbour in direction d is child ((k+d) mod 2) of the parent-level
    # neighbour floor((k+d)/2): check by coordinates on a 4x4x4 lattice
    for k in range(8):
        kc = np.array([k & 1, (k >> 1) & 1, (k >> 2) & 1])
        for s in range(27):
            d = np.array([s % 3 - 1, (s // 3) % 3 - 1, s // 9 - 1])
            child = 2 + kc + d  # absolute child coordinate, parent of self at (1,1,1)
            pd = child // 2 - 1
            ps = (pd[0] + 1) + 3 * (pd[1] + 1) + 9 * (pd[2] + 1)
            ck = (child[0] & 1) | (child[1] & 1) << 1 | (child[2] & 1) << 2
            assert luts["nb_parent"][k, s] == ps
            assert luts["nb_child"][k, s] == ck


def test_sharing_luts_contain_self(luts):
    # every corner / edge row lists slot 13 (the cell itself): the reference's
    # owner rule therefore falls back to self when only root[13] is stored
    assert all(13 in row for row in luts["vert_shared"])
    assert all(13 in row for row in luts["edge_shared"])
