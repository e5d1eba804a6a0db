"""How many voxels change under the arithmetic the restatement cannot pin.

The voxel tests of SVO.cu:69-185 go through Eigen 3.4 (not available here)
and were compiled by nvcc, which contracts multiply-adds by default.  The
oracle fixes one reading (Eigen's a0 + (a1 + a2) split, no contraction); this
script counts how many voxels differ from it under the other readings:

    variant 1  3-term dots summed left to right
    variant 2  FMA contraction in dots and the cross product
    variant 3  both

Usage: python oracle/arith_sensitivity.py   (CPU only, a few minutes)
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
sys.path.insert(0, os.path.join(HERE, "..", "gpu-svo_amd"))

import pyoracle as po  # noqa: E402
from svo_amd import meshgen  # noqa: E402


def cases():
    import gzip
    import tempfile
    gz = os.path.join(HERE, "..", "tests", "golden", "bunny.obj.gz")
    with tempfile.NamedTemporaryFile(suffix=".obj") as f:
        f.write(gzip.open(gz).read())
        f.flush()
        _, P, T = po.load_obj(f.name)
    yield "bunny.obj", P, T, 64
    yield "bunny.obj", P, T, 256
    P, T = meshgen.icosphere(4)
    yield "icosphere-4", P, T, 64
    P, T = meshgen.noisy_icosphere(6, seed=1)
    yield "icosphere-6 + noise", P, T, 512
    P, T = meshgen.soup(200_000, seed=2)
    yield "soup 200K triangles", P, T, 1024


def main():
    print("| mesh | res | voxels | variant 1 | variant 2 | variant 3 |")
    print("|---|---:|---:|---:|---:|---:|")
    for name, P, T, res in cases():
        po.set_arith_variant(0)
        st, base = po.voxelize(P, T, res)
        assert st == 0
        row = [name, str(res), str(base.size)]
        for v in (1, 2, 3):
            po.set_arith_variant(v)
            st, other = po.voxelize(P, T, res)
            assert st == 0
            diff = np.setxor1d(base, other).size
            row.append(f"{diff} ({diff / base.size:.2e})")
        po.set_arith_variant(0)
        print("| " + " | ".join(row) + " |", flush=True)


if __name__ == "__main__":
    main()
