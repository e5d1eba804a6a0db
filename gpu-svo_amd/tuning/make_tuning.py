"""Writes a tuning file for this device: the library's tuning search
(svo_ctx_autotune: global one-lane box limit and concurrent passes, then the
one-lane box limit per size class) on the 256-mesh mixed batch (icospheres
and tori, 320-81,920 triangles, res 64-256, seed 3), device-resident.

    python gpu-svo_amd/tuning/make_tuning.py [out.tuning] [--reps 5]

The file holds "name value" lines (svo_ctx_load_tuning reads them); bench.py
loads gpu-svo_amd/tuning/mi355x.tuning by default.
"""
import argparse
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("out", nargs="?", default=os.path.join(HERE, "mi355x.tuning"))
    ap.add_argument("--reps", type=int, default=5)
    args = ap.parse_args()
    import svo_amd
    from svo_amd import meshgen
    meshes = meshgen.mixed_batch(256, seed=3)
    ctx = svo_amd.Context(0)
    dms = [ctx.upload_mesh(P, T) for P, T, _ in meshes]
    descs = [d.desc(r) for d, (_, _, r) in zip(dms, meshes)]
    print("tuning search on the 256-mesh batch", file=sys.stderr, flush=True)
    r = ctx.autotune(descs, on_device=True, reps=args.reps)
    ctx.save_tuning(args.out)
    print(json.dumps(r), flush=True)
    for d in dms:
        d.free()
    ctx.close()


if __name__ == "__main__":
    main()
