"""The `svo` command line (main.cpp:7-17) and the SparseVoxelOctree class
behind it: `svo [mesh.obj=bunny.obj] [res=64]` prints the depth line of
create() (SVO.cu:488) and the writeTree line (SVO.cu:744), and writes
<stem>_<depth>_tree.obj byte-identical to the oracle's writer; with
--write-voxels the voxel file (SVO.cu:755-777) comes first, as in the
reference's NDEBUG build."""
import os
import shutil
import subprocess

import pytest

import pyoracle as po

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "gpu-svo_amd", "svo")


def _run(args, cwd):
    if not os.path.exists(CLI):
        pytest.fail("the svo command line is not built (make -C gpu-svo_amd)")
    print("svo", " ".join(args), flush=True)
    return subprocess.run([CLI] + args, cwd=cwd, capture_output=True, text=True, timeout=120)


def _oracle_file(bunny, res, path, voxels=False):
    st, P, T = po.load_obj(bunny)
    assert st == 0
    st, depth = po.write_obj(P, T, res, path, voxels=voxels)
    assert st == 0
    return depth


def test_default_arguments(bunny_path, tmp_path):
    # no arguments: bunny.obj in the working directory at res 64
    shutil.copy(bunny_path, tmp_path / "bunny.obj")
    r = _run([], str(tmp_path))
    assert r.returncode == 0, r.stderr
    depth = _oracle_file(bunny_path, 64, str(tmp_path / "oracle.obj"))
    name = f"bunny_{depth}_tree.obj"
    assert r.stdout == (f"Sparse Voxel Octree depth: {depth}\n"
                        f"[I/O] Writing octree data in obj format to file {name} \n")
    with open(tmp_path / name, "rb") as a, open(tmp_path / "oracle.obj", "rb") as b:
        assert a.read() == b.read()


@pytest.mark.parametrize("res", [16, 100])
def test_mesh_res_and_voxels(bunny_path, tmp_path, res):
    out = tmp_path / "out"
    out.mkdir()
    r = _run([bunny_path, str(res), "--out-dir", str(out), "--write-voxels"], str(tmp_path))
    assert r.returncode == 0, r.stderr
    depth = _oracle_file(bunny_path, res, str(tmp_path / "oracle_tree.obj"))
    _oracle_file(bunny_path, res, str(tmp_path / "oracle_vox.obj"), voxels=True)
    tree = f"{out}/bunny_{depth}_tree.obj"
    vox = f"{out}/bunny_1_voxel.obj"
    assert r.stdout == (f"[I/O] Writing data in obj voxels format to file {vox} \n"
                        f"Sparse Voxel Octree depth: {depth}\n"
                        f"[I/O] Writing octree data in obj format to file {tree} \n")
    for mine, ref in ((tree, "oracle_tree.obj"), (vox, "oracle_vox.obj")):
        with open(mine, "rb") as a, open(tmp_path / ref, "rb") as b:
            assert a.read() == b.read(), mine


def test_missing_mesh_is_an_error(tmp_path):
    # the reference exits(1) after its loader's message; the CLI reports the
    # status and exits 1 without writing anything
    r = _run([str(tmp_path / "nope.obj"), "32"], str(tmp_path))
    assert r.returncode == 1
    assert "svo: " in r.stderr
    assert os.listdir(tmp_path) == []
