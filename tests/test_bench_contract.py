"""bench.py's contract pieces that need no GPU: one GPU only, and the
algorithmic-byte accounting behind `roofline` and `dedup_gbs`."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


def _stats(**kw):
    s = {"key_bytes": 4, "n_hits": 1000, "n_voxels": 600, "n_nodes": 900, "n_tris": 50,
         "bitmap_bytes": 0, "n_groups1": 150, "sort_passes": 4, "level_elems": 800,
         "n_vertices": 7200, "n_edges": 10800, "n_faces": 0, "bitmap_fill_bytes": 0,
         "bitmap_reset_bytes": 0}
    s.update(kw)
    return s


@pytest.mark.parametrize("n", [2, 8])
def test_more_than_one_gpu_exits_with_a_message(n):
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", str(n)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""


def test_sort_path_bytes():
    s = _stats()
    assert bench.stage_bytes(s, "sort") == 2 * 4 * 1000 * 4
    # sorted hits in; unique keys, parent keys and first children out
    assert bench.stage_bytes(s, "unique_level1") == 4 * 1000 + 4 * 600 + 8 * 150


def test_bitmap_path_bytes_count_fill_and_reset():
    s = _stats(bitmap_bytes=4096, bitmap_fill_bytes=1024, bitmap_reset_bytes=256)
    # the fill stage moves what it filled, not the whole bitmap
    assert bench.stage_bytes(s, "bitmap_clear") == 1024
    assert bench.stage_bytes(s, "unique_level1") == 4096 + 4 * 600 + 8 * 150 + 256
    # voxelization: triangles in, one bitmap word update per hit
    assert bench.stage_bytes(s, "voxelize") == 36 * 50 + 4 * 1000


def test_dedup_rate_uses_the_stages_present():
    s = _stats(bitmap_bytes=4096, bitmap_reset_bytes=256)
    stages = {"voxelize": 1.0, "unique_level1": 0.002}
    want = bench.stage_bytes(s, "unique_level1") / 0.002e-3 / 1e9
    assert bench.dedup_gbs(s, stages) == round(want, 1)
    assert bench.dedup_gbs(s, {"voxelize": 1.0}) is None


def test_compulsory_bytes():
    # 36 B per triangle, 4 B per unique voxel, 165 B of node arrays per node
    # (u32 code), 16 B per vertex and 28 B per edge record
    s = _stats()
    assert bench.compulsory_bytes(s) == 36 * 50 + 4 * 600 + 165 * 900 + 16 * 7200 + 28 * 10800
    assert bench.compulsory_gbs(s, 2.0) == round(bench.compulsory_bytes(s) / 2e-3 / 1e9, 1)
    s8 = _stats(key_bytes=8, n_faces=300)
    assert bench.compulsory_bytes(s8) - bench.compulsory_bytes(s) == 4 * 600 + 4 * 900 + 28 * 300


def test_roofline_picks_the_longest_device_stage():
    s = _stats()
    stages = {"setup_b": 9.0, "voxelize": 0.5, "attr_write": 2.0, "nodes": 0.4}
    r = bench._roofline(s, stages)
    assert r["kernel"] == "attr_write"
    b = bench.stage_bytes(s, "attr_write")
    assert r["algorithmic_bytes"] == b
    assert r["achieved"] == round(b / 2.0e-3 / 1e9, 1)
    assert r["frac"] == round(r["achieved"] / 8000.0, 4)
    assert r["traffic"] == round(bench.TRAFFIC_PER_ALG_BYTE["attr_write"] * b)
