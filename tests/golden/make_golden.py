"""Writes the golden fixtures under tests/golden/ from the reference checkout.

Run in a container that has the reference at /root/reference:

    python tests/golden/make_golden.py

Outputs (data only, no source text):
  luts.json      numeric values of the lookup tables printed in the reference
                 headers: MortonLUT.h:13-233 (encode), 236-449 (decode),
                 531-597 (neighbor parent/child), SVO.cu:567-584 (corner
                 sharing) and SVO.cu:610-627 (edge sharing); and the 12 edge
                 endpoint pairs of determineNodeEdge, SVO.cu:638-663, as
                 offsets in units of the node width plus whether the endpoint
                 is the bare `origin` term (no addition).
  bunny.obj.gz   the reference's sample input mesh (bunny.obj), gzip-compressed.
"""
from __future__ import annotations

import gzip
import hashlib
import json
import os
import re
import sys

REF = os.environ.get("SVO_REFERENCE", "/root/reference")
HERE = os.path.dirname(os.path.abspath(__file__))


def _table(text: str, name: str) -> list[int]:
    m = re.search(re.escape(name) + r"\s*(\[[^\]]*\])+\s*=\s*\{", text)
    if not m:
        raise SystemExit(f"table {name} not found")
    depth, i = 0, m.end() - 1
    start = i
    while True:
        c = text[i]
        if c == "{":
            depth += 1
        elif c == "}":
            depth -= 1
            if depth == 0:
                break
        i += 1
    body = text[start:i + 1]
    body = re.sub(r"//[^\n]*", "", body)
    return [int(t, 0) for t in re.findall(r"0x[0-9A-Fa-f]+|\d+", body)]


def _split_top(s: str) -> list[str]:
    """Splits s on commas outside parentheses."""
    out, depth, cur = [], 0, ""
    for c in s:
        if c == "(":
            depth += 1
        elif c == ")":
            depth -= 1
        if c == "," and depth == 0:
            out.append(cur)
            cur = ""
        else:
            cur += c
    out.append(cur)
    return out


def _endpoint(term: str) -> tuple[list[int], bool]:
    """`origin` -> ([0,0,0], bare); `origin + Eigen::Vector3f(a, b, c)` with
    a, b, c each 0 or width -> (offsets in units of width, not bare)."""
    t = re.sub(r"\s+", "", term)
    if t == "origin":
        return [0, 0, 0], True
    m = re.fullmatch(r"origin\+Eigen::Vector3f\(([^()]*)\)", t)
    if not m:
        raise SystemExit(f"unexpected edge endpoint {term!r}")
    comps = m.group(1).split(",")
    if len(comps) != 3 or any(c not in ("0", "width") for c in comps):
        raise SystemExit(f"unexpected edge endpoint {term!r}")
    return [1 if c == "width" else 0 for c in comps], False


def _edge_table(text: str) -> tuple[list, list]:
    """The initializer of `thrust_edge edges[12]` in determineNodeEdge."""
    m = re.search(r"thrust_edge\s+edges\s*\[\s*12\s*\]\s*=\s*\{", text)
    if not m:
        raise SystemExit("edge table not found")
    i, depth = m.end() - 1, 0
    start = i
    while True:
        c = text[i]
        if c == "{":
            depth += 1
        elif c == "}":
            depth -= 1
            if depth == 0:
                break
        i += 1
    body = text[start + 1:i]
    pairs = re.findall(r"thrust::make_pair\(", body)
    if len(pairs) != 12:
        raise SystemExit(f"edge table: {len(pairs)} pairs, expected 12")
    offsets, bare = [], []
    pos = 0
    for _ in range(12):
        k = body.index("thrust::make_pair(", pos) + len("thrust::make_pair(")
        d, j = 1, k
        while d:
            if body[j] == "(":
                d += 1
            elif body[j] == ")":
                d -= 1
            j += 1
        args = _split_top(body[k:j - 1])
        if len(args) != 2:
            raise SystemExit(f"edge pair with {len(args)} arguments")
        (oa, ba), (ob, bb) = _endpoint(args[0]), _endpoint(args[1])
        offsets.append([oa, ob])
        bare.append([ba, bb])
        pos = j
    return offsets, bare


def main() -> int:
    lut_h = open(os.path.join(REF, "src", "MortonLUT.h"), encoding="utf-8-sig").read()
    svo_cu = open(os.path.join(REF, "src", "SVO.cu"), encoding="utf-8-sig").read()
    out = {
        "enc_x": _table(lut_h, "h_morton256_x"),
        "enc_y": _table(lut_h, "h_morton256_y"),
        "enc_z": _table(lut_h, "h_morton256_z"),
        "d_enc_x": _table(lut_h, "d_morton256_x"),
        "d_enc_y": _table(lut_h, "d_morton256_y"),
        "d_enc_z": _table(lut_h, "d_morton256_z"),
        "dec_x": _table(lut_h, "h_Morton3D_decode_x_512"),
        "dec_y": _table(lut_h, "h_Morton3D_decode_y_512"),
        "dec_z": _table(lut_h, "h_Morton3D_decode_z_512"),
        "d_dec_x": _table(lut_h, "d_Morton3D_decode_x_512"),
        "d_dec_y": _table(lut_h, "d_Morton3D_decode_y_512"),
        "d_dec_z": _table(lut_h, "d_Morton3D_decode_z_512"),
        "nb_parent": _table(lut_h, "neighbor_LUTparent"),
        "nb_child": _table(lut_h, "neighbor_LUTchild"),
        "vert_shared": _table(svo_cu, "d_vertSharedLUT"),
        "edge_shared": _table(svo_cu, "d_edgeSharedLUT"),
    }
    edge_ofs, edge_bare = _edge_table(svo_cu)
    sizes = {"enc": 256, "dec": 512, "nb_parent": 216, "nb_child": 216, "vert_shared": 64,
             "edge_shared": 48}
    for k, v in out.items():
        want = sizes.get(k) or sizes["enc" if "enc" in k else "dec"]
        if len(v) != want:
            raise SystemExit(f"{k}: {len(v)} values, expected {want}")
    out["edge_endpoints"] = edge_ofs
    out["edge_endpoint_bare"] = edge_bare
    with open(os.path.join(HERE, "luts.json"), "w") as f:
        json.dump(out, f, separators=(",", ":"))
    src = open(os.path.join(REF, "bunny.obj"), "rb").read()
    with gzip.GzipFile(os.path.join(HERE, "bunny.obj.gz"), "wb", mtime=0) as f:
        f.write(src)
    meta = {"bunny.obj": {"sha256": hashlib.sha256(src).hexdigest(), "bytes": len(src)}}
    with open(os.path.join(HERE, "fixtures_meta.json"), "w") as f:
        json.dump(meta, f, indent=1)
    print("wrote luts.json, bunny.obj.gz, fixtures_meta.json")
    return 0


if __name__ == "__main__":
    sys.exit(main())
