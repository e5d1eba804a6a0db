"""The device-memory bound of the served path (svo_estimate_bytes_mesh) on the
CPU: at least the output block the build allocates for the oracle's node and
voxel counts (every node array, the voxel codes and the vertex / edge / face
records, each on a 256-byte boundary), on meshes whose triangles are smaller
than, about as large as, and much larger than the voxels, at several
resolutions, in both semantics and both key widths.  The device test
(test_gpu_serve.py) checks it against what a build actually takes."""
import numpy as np
import pytest

import pyoracle as po


def _align(x):
    return (int(x) + 255) // 256 * 256


def _output_bytes(info, sem, kb, attrs=True):
    N, U = int(info.n_nodes), int(info.n_voxels)
    b = kb // 8
    arrays = [b * N, N, 12 * N, 4 * N, 4 * N, 32 * N, 108 * N, b * U]
    if attrs:
        arrays += [16 * 8 * N, 28 * 12 * N] + ([28 * 6 * N] if sem == 1 else [])
    return sum(_align(x) for x in arrays)


def _cases():
    from svo_amd import meshgen
    P, T = meshgen.icosphere(5)
    yield "ico5", P, T, (64, 256)
    P, T = meshgen.icosphere(2)
    yield "ico2", P, T, (64, 300)
    P, T = meshgen.torus(0.4, 0.1, 120, 40)
    yield "torus", P, T, (128, (96, 64, 40))
    P, T = meshgen.noisy_icosphere(4, seed=3)
    yield "ico4-noise", P, T, (512,)
    P, T = meshgen.soup(20000, seed=5)
    yield "soup", P, T, (256, 1024)
    P = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0.2, 0.2, 1e-7]], np.float32)
    T = np.array([[0, 1, 2], [0, 1, 3]], np.int32)
    yield "flat-and-sliver", P, T, (33, 700)


@pytest.mark.parametrize("case", range(6))
def test_bound_covers_output(case):
    import svo_amd
    name, P, T, resolutions = list(_cases())[case]
    for res in resolutions:
        for sem, kb in ((0, 32), (1, 64)):
            r = po.build(P, T, res, semantics=sem, key_bits=kb, fields=("voxels",))
            assert r.status == 0, (name, res)
            need = _output_bytes(r.info, sem, kb)
            bound = svo_amd.estimate_bytes_mesh(P, T, res, semantics=sem, key_bits=kb)
            assert bound >= need, (name, res, sem, bound, need)
            # records off: the bound shrinks, still covering the node arrays
            nb = svo_amd.estimate_bytes_mesh(P, T, res, semantics=sem, key_bits=kb,
                                             flags=svo_amd.BUILD_NO_VERTEX_EDGE)
            assert _output_bytes(r.info, sem, kb, attrs=False) <= nb < bound


def test_bound_rejected_inputs_fall_back():
    import svo_amd
    P = np.zeros((3, 3), np.float32)
    T = np.array([[0, 1, 2]], np.int32)
    # zero-extent box and out-of-range index: the build rejects these; the
    # bound falls back to the count-only estimate
    assert svo_amd.estimate_bytes_mesh(P, T, 64) == svo_amd.estimate_bytes(3, 1, 64)
    P2 = np.random.default_rng(0).random((3, 3)).astype(np.float32)
    T2 = np.array([[0, 1, 5]], np.int32)
    assert svo_amd.estimate_bytes_mesh(P2, T2, 64) == svo_amd.estimate_bytes(3, 1, 64)
