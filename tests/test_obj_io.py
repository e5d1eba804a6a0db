"""OBJ input rules (BaseModel.cpp:3-25) and the tree .obj format (IO.h:23-90).

The product parser (libsvo_amd.so, host code, no device needed) and the
oracle's parser must agree on every file, and both must follow the
reference's getline(buffer, 255) + sscanf rules.  The writer format is checked
against Python's printf-style "%g" formatting of the oracle's node arrays.
"""
import os

import numpy as np
import pytest

import pyoracle as po


def _both(path):
    import svo_amd
    st_o, P_o, T_o = po.load_obj(path)
    st_p = svo_amd.load_obj_status(path)
    assert st_o == st_p
    if st_o != 0:
        return st_o, None, None
    P, T = svo_amd.load_obj(path)
    assert np.array_equal(P.view(np.uint32), P_o.view(np.uint32))
    assert np.array_equal(T, T_o)
    return 0, P, T


def _write(tmp_path, name, text, binary=False):
    p = tmp_path / name
    p.write_bytes(text if binary else text.encode())
    return str(p)


def test_face_forms(tmp_path):
    text = ("# comment\n"
            "v 0 0 0\nv 1 0 0\nv 0 1 0\nv 0 0 1\n"
            "vn 0 0 1\nvt 0 0\n"
            "f 1 2 3\n"              # plain
            "f 1//1 2//1 4//1\n"     # v//vn
            "f 1/1 3/1 4/1\n"        # v/vt
            "f 1/1/1 2/1/1 3/1/1\n"  # v/vt/vn: matches no pattern, dropped
            "f 2 3 4 1\n"            # quad: first triangle
            "f\t1 2 3\n"             # tab after f: not a face line
            "  v 9 9 9\n"            # leading space: not a vertex line
            "v 1 2\n"                # two coordinates: dropped
            "v 1e-3 -2.5 +7\n")
    st, P, T = _both(_write(tmp_path, "forms.obj", text))
    assert st == 0
    assert P.shape == (5, 3)
    assert np.allclose(P[4], [1e-3, -2.5, 7])
    assert T.tolist() == [[0, 1, 2], [0, 1, 3], [0, 2, 3], [1, 2, 3]]


def test_last_line_without_newline_is_dropped(tmp_path):
    st, P, T = _both(_write(tmp_path, "nonl.obj", "v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 3"))
    assert st == 0 and P.shape == (3, 3) and T.shape == (0, 3)
    st, P, T = _both(_write(tmp_path, "nl.obj", "v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 3\n"))
    assert T.shape == (1, 3)


def test_crlf(tmp_path):
    st, P, T = _both(_write(tmp_path, "crlf.obj", "v 0 0 0\r\nv 1 0 0\r\nv 0 1 0\r\nf 1 2 3\r\n"))
    assert st == 0 and P.shape == (3, 3) and T.shape == (1, 3)


@pytest.mark.parametrize("n,ok", [(254, True), (255, False), (300, False)])
def test_long_lines(tmp_path, n, ok):
    # getline(buffer, 255): a line of 255+ characters sets failbit (the
    # reference then loops forever); 254 characters are fine
    line = "v 1 2 3" + " " * (n - 7)
    assert len(line) == n
    st, P, T = _both(_write(tmp_path, f"long{n}.obj", line + "\nv 0 0 0\n"))
    if ok:
        assert st == 0 and P.shape == (2, 3)
    else:
        assert st == -6  # SVO_E_PARSE


def test_missing_file(tmp_path):
    st, _, _ = _both(str(tmp_path / "nope.obj"))
    assert st == -5  # SVO_E_IO


def test_empty_file(tmp_path):
    st, P, T = _both(_write(tmp_path, "empty.obj", ""))
    assert st == 0 and P.shape == (0, 3) and T.shape == (0, 3)


def test_bunny_counts(bunny_path):
    st, P, T = _both(bunny_path)
    assert st == 0
    assert P.shape == (14290, 3) and T.shape == (28576, 3)
    assert np.allclose(P.min(0), [-0.500044, -0.493323, -0.387304], atol=1e-6)
    assert np.allclose(P.max(0), [0.499956, 0.493617, 0.385848], atol=1e-6)


def test_large_file_parallel_parse(tmp_path):
    # several MiB so the product parser splits the file across threads
    from svo_amd import meshgen
    P, T = meshgen.icosphere(6)
    lines = ["v %.9g %.9g %.9g" % tuple(p) for p in P.astype(np.float64)]
    lines += ["f %d %d %d" % tuple(t + 1) for t in T]
    st, P2, T2 = _both(_write(tmp_path, "big.obj", "\n".join(lines) + "\n"))
    assert st == 0
    assert np.array_equal(P2, P) and np.array_equal(T2, T)


def test_extreme_face_indices(tmp_path):
    # index - 1 wraps for INT_MIN (two's complement); out-of-range values are
    # kept as parsed and rejected later by validation
    text = ("v 0 0 0\nv 1 0 0\nv 0 1 0\n"
            "f -2147483648 2147483647 0\n"
            "f 1e39 nan -inf\n"          # not integers: %d reads 1, then fails
            "v 1e39 nan -inf\n")         # overflow to inf and the nan/inf words
    st, P, T = _both(_write(tmp_path, "extreme.obj", text))
    assert st == 0
    assert T.shape == (1, 3) and T[0].tolist() == [2147483647, 2147483646, -1]
    assert P.shape == (4, 3) and np.isinf(P[3, 0]) and np.isnan(P[3, 1])


def _mutate(rng, base, n_ops):
    s = bytearray(base)
    for _ in range(n_ops):
        pos = int(rng.integers(len(s) + 1))
        op = rng.random()
        if op < 0.3 and pos < len(s):
            s[pos] = int(rng.integers(256))
        elif op < 0.5:
            del s[pos:pos + int(rng.integers(1, 21))]
        elif op < 0.75:
            s[pos:pos] = rng.integers(0, 256, int(rng.integers(1, 300)), dtype=np.uint8).tobytes()
        elif op < 0.85:
            s[pos:pos] = b"\n" * int(rng.integers(1, 4))
        else:
            s[pos:pos] = b"f 2147483647 -2147483648 99999999999999999999\nv 1e39 nan -inf\n"
    return bytes(s)


def test_fuzzed_files_agree_with_oracle(tmp_path):
    # random byte edits of a valid file (NULs, CRs, overlong lines, binary
    # junk, no final newline): status, points and indices must match the
    # oracle's istream/sscanf reading on every file; two files of several
    # MiB take the product's multi-threaded split
    rng = np.random.default_rng(11)
    lines = ["v %.6g %.6g %.6g\n" % tuple(rng.random(3)) for _ in range(150)]
    forms = ["f %d %d %d\n", "f %d//1 %d//2 %d//3\n", "f %d/1 %d/2 %d/3\n", "f %d/1/1 %d/2/2 %d/3/3\n"]
    lines += [forms[int(rng.integers(4))] % tuple(rng.integers(-5, 200, 3)) for _ in range(250)]
    base = "".join(lines).encode()
    statuses = set()
    for k in range(120):
        data = _mutate(rng, base, int(rng.integers(0, 30)))
        if rng.random() < 0.3:
            data = data.rstrip(b"\n")
        st, _, _ = _both(_write(tmp_path, f"fz{k}.obj", data, binary=True))
        statuses.add(st)
    for k in range(2):
        data = _mutate(rng, base * 400, 40)
        if k:
            data += b"x" * 300  # overlong final line without a newline
        st, _, _ = _both(_write(tmp_path, f"fzbig{k}.obj", data, binary=True))
        statuses.add(st)
    assert statuses == {0, -6}  # both outcomes exercised


def expected_tree_text(origin, width):
    C = [(0, 1, 1), (0, 0, 1), (0, 1, 0), (0, 0, 0), (1, 0, 0), (1, 1, 0), (1, 0, 1), (1, 1, 1)]
    Lp = [(1, 2), (2, 7), (7, 8), (8, 1), (3, 4), (4, 5), (5, 6), (6, 3), (3, 1), (4, 2), (5, 7),
          (6, 8)]
    out = []
    for t in range(width.size):
        o = origin[t]
        w = width[t]
        for k, c in enumerate(C):
            v = [o[a] if k == 3 else np.float32(o[a] + (w if c[a] else np.float32(0)))
                 for a in range(3)]
            out.append("v %g %g %g\n" % tuple(float(x) for x in v))
        for a, b in Lp:
            out.append("l %d %d\n" % (8 * t + a, 8 * t + b))
    return "".join(out)


def test_oracle_writer_format(tmp_path):
    from svo_amd import meshgen
    P, T = meshgen.icosphere(2)
    path = str(tmp_path / "ico_tree.obj")
    st, depth = po.write_obj(P, T, 16, path)
    assert st == 0
    r = po.build(P, T, 16)
    assert open(path).read() == expected_tree_text(r["origin"], r["width"])


def test_oracle_voxel_writer_is_level1_cubes(tmp_path):
    # writeVoxel's cube for voxel c is bboxMin + unit.x * decode(c) with edge
    # unit.x, which is the level-1 node of code c: the file must equal the
    # tree writer's text for the level-1 nodes that are occupied voxels.
    from svo_amd import meshgen
    P, T = meshgen.torus(0.4, 0.12, 24, 12)
    for res in (16, (20, 12, 9)):
        path = str(tmp_path / "t_1_voxel.obj")
        st, _ = po.write_obj(P, T, res, path, voxels=True)
        assert st == 0
        r = po.build(P, T, res)
        n1 = int(r.info.level_nodes[0])
        code1 = r["code"][:n1]
        occ = np.isin(code1, r["voxels"])
        assert occ.sum() == r["voxels"].size
        assert np.array_equal(code1[occ], r["voxels"])
        exp = expected_tree_text(r["origin"][:n1][occ], r["width"][:n1][occ])
        assert open(path).read() == exp
