"""Writes a tuning file for this device: the library's tuning search
(svo_ctx_autotune: global one-lane box limit and concurrent passes, then the
one-lane box limit per size class) on the 256-mesh mixed batch (icospheres
and tori, 320-81,920 triangles, res 64-256, seed 3), device-resident; then
the one-lane box limit of the size classes of two single large meshes the
batch does not hold, by timing their builds: the 2,000,000-triangle soup at
res 1024 and icosphere-7 with value noise at res 512.

    python gpu-svo_amd/tuning/make_tuning.py [out.tuning] [--reps 5]

The file holds "name value" lines (svo_ctx_load_tuning reads them); bench.py
loads gpu-svo_amd/tuning/mi355x.tuning by default.
"""
import argparse
import json
import os
import statistics
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

SINGLE_CANDIDATES = (512, 1024, 2048, 4096)


def size_class(res, n_tris):
    """The library's size class of a mesh (size_class_of)."""
    band = 0 if res <= 64 else 1 if res <= 128 else 2 if res <= 256 else 3 if res <= 512 else 4
    q = res * res / max(n_tris, 1)
    return 3 * band + (0 if q <= 4.0 else 1 if q <= 64.0 else 2)


def tune_single(ctx, name, P, T, res, reps):
    """Times builds of one mesh for each one-lane box limit of its size class
    and keeps the fastest if it beats the current limit by 3%."""
    cls = size_class(res, T.shape[0])
    opt = f"class_small_max_{cls}"
    base = ctx.get_option(opt) or ctx.get_option("small_max_cells")
    dm = ctx.upload_mesh(P, T)

    def med_ms(v):
        ctx.set_option(opt, v)
        ctx.build_device(*dm.desc(res)).free()  # warm-up
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            ctx.build_device(*dm.desc(res)).free()
            ts.append((time.perf_counter() - t0) * 1e3)
        return statistics.median(ts)

    ref = med_ms(base)
    times = {v: (ref if v == base else med_ms(v)) for v in SINGLE_CANDIDATES}
    best = min(times, key=times.get)
    keep = best if times[best] < 0.97 * ref else base
    ctx.set_option(opt, keep if keep != ctx.get_option("small_max_cells") else 0)
    dm.free()
    return {"mesh": name, "class": cls, "ms": {str(k): round(v, 3) for k, v in times.items()},
            "small_max_cells": ctx.get_option(opt)}


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
    for d in dms:
        d.free()
    singles = []
    P, T = meshgen.soup(2_000_000, seed=2)
    print("one-lane limit of the soup's size class", file=sys.stderr, flush=True)
    singles.append(tune_single(ctx, "soup 2M res 1024", P, T, 1024, args.reps))
    P, T = meshgen.noisy_icosphere(7, seed=1)
    print("one-lane limit of the icosphere's size class", file=sys.stderr, flush=True)
    singles.append(tune_single(ctx, "icosphere-7 + noise res 512", P, T, 512, args.reps))
    r["single_meshes"] = singles
    ctx.save_tuning(args.out)
    print(json.dumps(r), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
