"""Field-by-field comparison of a HIP-built tree with the oracle (bit-exact)."""
import numpy as np

import pyoracle as po

NODE_FIELDS = ["code", "is_leaf", "origin", "width", "parent", "childs", "neighbors"]
RECORD_FIELDS = ["vertices", "edges", "faces"]


def bits(a: np.ndarray) -> np.ndarray:
    """Raw bytes view so that -0.0 / +0.0 and NaN payloads compare exactly."""
    a = np.ascontiguousarray(a)
    return a.view(np.uint8).reshape(a.shape[0], -1) if a.ndim else a.view(np.uint8)


def assert_same(name, got, want):
    assert got.shape == want.shape, f"{name}: shape {got.shape} != {want.shape}"
    if got.size == 0:
        return
    g, w = bits(got), bits(want)
    if not np.array_equal(g, w):
        rows = np.nonzero((g != w).reshape(g.shape[0], -1).any(axis=1))[0]
        i = rows[0]
        raise AssertionError(f"{name}: {rows.size} of {g.shape[0]} rows differ; first at {i}: "
                             f"got {got[i]!r} want {want[i]!r}")


def compare(tree, ref: "po.OracleTree", fields=None, check_aos=True):
    info = tree.info
    assert info.depth == ref.info.depth, (info.depth, ref.info.depth)
    assert info.n_voxels == ref.info.n_voxels
    assert info.n_nodes == ref.info.n_nodes
    for a in range(3):
        for k in ("bbox_min", "bbox_max", "bbox_width", "unit"):
            assert np.float32(getattr(info, k)[a]).tobytes() == \
                np.float32(getattr(ref.info, k)[a]).tobytes(), k
    names = fields or (["voxels"] + NODE_FIELDS + RECORD_FIELDS + ["level_nodes"])
    for name in names:
        if name not in ref.arrays:
            continue
        assert_same(name, tree.field(name), ref[name])
    if check_aos and "nodes_aos" in ref.arrays and (fields is None or "nodes_aos" in fields):
        assert_same("nodes_aos", tree.field("nodes_aos"), ref["nodes_aos"])
