"""Deterministic synthetic meshes for the SVO workloads.

All randomness comes from a splitmix64 stream implemented here with numpy
uint64 arithmetic, so the meshes do not depend on numpy's distribution code.
Meshes are returned as (points float32[P,3], tris int32[T,3]).

Workloads (sizes as named by the benchmark configurations):
  icosphere(4)                         5,120 triangles, 2,562 vertices
  noisy_icosphere(7, seed=1)           327,680 triangles, radial value noise 0.1
  soup(2_000_000, seed=2)              independent equilateral triangles,
                                       lognormal edge (median 1/256, sigma 0.6)
  mixed_batch(256, seed=3)             icospheres (subdiv 2-6) and tori,
                                       320-81,920 triangles log-uniform,
                                       resolution from {64, 128, 256}
  request_stream(...)                  Poisson arrivals for the served path
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def splitmix64(state: int, n: int) -> np.ndarray:
    """n consecutive splitmix64 outputs starting from `state` (uint64 array)."""
    with np.errstate(over="ignore"):
        idx = np.arange(1, n + 1, dtype=np.uint64)
        z = np.uint64(state & 0xFFFFFFFFFFFFFFFF) + idx * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


class Rng:
    """Counter-based generator over splitmix64."""

    def __init__(self, seed: int):
        self.state = (int(seed) * 0x9E3779B97F4A7C15 + 0x632BE59BD9B4E019) & 0xFFFFFFFFFFFFFFFF
        self.count = 0

    def u64(self, n: int) -> np.ndarray:
        out = splitmix64((self.state + self.count * 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF, n)
        self.count += n
        return out

    def uniform(self, n: int, lo: float = 0.0, hi: float = 1.0) -> np.ndarray:
        u = (self.u64(n) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
        return lo + (hi - lo) * u

    def normal(self, n: int) -> np.ndarray:
        """Box-Muller."""
        m = (n + 1) // 2
        u1 = self.uniform(m)
        u2 = self.uniform(m)
        r = np.sqrt(-2.0 * np.log1p(-u1))  # 1-u1 in (0, 1]
        z = np.concatenate([r * np.cos(2 * np.pi * u2), r * np.sin(2 * np.pi * u2)])
        return z[:n]

    def integers(self, n: int, lo: int, hi: int) -> np.ndarray:
        """Uniform integers in [lo, hi]."""
        span = hi - lo + 1
        return lo + (self.u64(n) % np.uint64(span)).astype(np.int64)


# --------------------------------------------------------------------------- icosphere
def icosphere(subdiv: int) -> tuple[np.ndarray, np.ndarray]:
    """Unit-sphere icosphere: 20*4^s triangles, 10*4^s + 2 vertices."""
    t = (1.0 + math.sqrt(5.0)) / 2.0
    v = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t],
                  [0, -1, -t], [0, 1, -t], [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]],
                 dtype=np.float64)
    f = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9],
                  [5, 11, 4], [11, 10, 2], [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2],
                  [3, 2, 6], [3, 6, 8], [3, 8, 9], [4, 9, 5], [2, 4, 11], [6, 2, 10],
                  [8, 6, 7], [9, 8, 1]], dtype=np.int64)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    for _ in range(subdiv):
        nv = v.shape[0]
        e = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]])
        e.sort(axis=1)
        key = e[:, 0] * (nv + 1) + e[:, 1]
        uniq, inv = np.unique(key, return_inverse=True)
        a = uniq // (nv + 1)
        b = uniq % (nv + 1)
        mid = v[a] + v[b]
        mid /= np.linalg.norm(mid, axis=1, keepdims=True)
        m = nv + inv.reshape(3, -1)  # midpoints of edges 01, 12, 20
        v = np.concatenate([v, mid])
        i0, i1, i2 = f[:, 0], f[:, 1], f[:, 2]
        m01, m12, m20 = m[0], m[1], m[2]
        f = np.concatenate([np.stack([i0, m01, m20], 1), np.stack([i1, m12, m01], 1),
                            np.stack([i2, m20, m12], 1), np.stack([m01, m12, m20], 1)])
    return v.astype(np.float32), f.astype(np.int32)


def _hash3(ix, iy, iz, seed: int) -> np.ndarray:
    """Lattice hash -> value in [-1, 1]."""
    with np.errstate(over="ignore"):
        k = (ix.astype(np.uint64) * np.uint64(0x8DA6B343) ^ iy.astype(np.uint64) * np.uint64(0xD8163841)
             ^ iz.astype(np.uint64) * np.uint64(0xCB1AB31F) ^ np.uint64(seed * 0x9E3779B9 & 0xFFFFFFFF))
        z = k * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) * (2.0 / 9007199254740992.0) - 1.0


def value_noise(p: np.ndarray, freq: float, seed: int) -> np.ndarray:
    """Trilinear value noise with smoothstep weights, values in [-1, 1]."""
    q = p * freq
    i = np.floor(q)
    fr = q - i
    w = fr * fr * (3.0 - 2.0 * fr)
    i = i.astype(np.int64)
    acc = np.zeros(p.shape[0])
    for dz in (0, 1):
        for dy in (0, 1):
            for dx in (0, 1):
                h = _hash3(i[:, 0] + dx, i[:, 1] + dy, i[:, 2] + dz, seed)
                wx = w[:, 0] if dx else 1 - w[:, 0]
                wy = w[:, 1] if dy else 1 - w[:, 1]
                wz = w[:, 2] if dz else 1 - w[:, 2]
                acc += h * wx * wy * wz
    return acc


def noisy_icosphere(subdiv: int, seed: int = 1, amplitude: float = 0.1, base_freq: float = 4.0):
    """Icosphere with each vertex displaced radially by 3-octave value noise."""
    v, f = icosphere(subdiv)
    p = v.astype(np.float64)
    n = (value_noise(p, base_freq, seed) + 0.5 * value_noise(p, 2 * base_freq, seed + 1) +
         0.25 * value_noise(p, 4 * base_freq, seed + 2)) / 1.75
    r = 1.0 + amplitude * n
    return (p * r[:, None]).astype(np.float32), f


# --------------------------------------------------------------------------- torus
def torus(R: float, r: float, a: int, b: int) -> tuple[np.ndarray, np.ndarray]:
    """Torus with a segments around the ring and b around the tube: 2ab triangles."""
    i = np.arange(a)
    j = np.arange(b)
    th = 2 * np.pi * i / a
    ph = 2 * np.pi * j / b
    TH, PH = np.meshgrid(th, ph, indexing="ij")
    x = (R + r * np.cos(PH)) * np.cos(TH)
    y = (R + r * np.cos(PH)) * np.sin(TH)
    z = r * np.sin(PH)
    v = np.stack([x, y, z], -1).reshape(-1, 3)
    I, J = np.meshgrid(i, j, indexing="ij")
    v00 = I * b + J
    v10 = ((I + 1) % a) * b + J
    v01 = I * b + (J + 1) % b
    v11 = ((I + 1) % a) * b + (J + 1) % b
    f = np.concatenate([np.stack([v00, v10, v11], -1).reshape(-1, 3),
                        np.stack([v00, v11, v01], -1).reshape(-1, 3)])
    return v.astype(np.float32), f.astype(np.int32)


# --------------------------------------------------------------------------- soup
def soup(n: int, seed: int = 2, median: float = 1.0 / 256.0, sigma: float = 0.6):
    """n independent equilateral triangles, centroids uniform in [0,1)^3,
    side length lognormal(ln median, sigma), uniformly random orientation."""
    rng = Rng(seed)
    c = rng.uniform(3 * n).reshape(n, 3)
    L = np.exp(math.log(median) + sigma * rng.normal(n))
    # random unit normal (uniform on the sphere) and in-plane angle
    zc = rng.uniform(n, -1.0, 1.0)
    az = rng.uniform(n, 0.0, 2 * np.pi)
    s = np.sqrt(np.maximum(0.0, 1 - zc * zc))
    nrm = np.stack([s * np.cos(az), s * np.sin(az), zc], 1)
    helper = np.where(np.abs(nrm[:, :1]) < 0.9, np.array([[1.0, 0, 0]]), np.array([[0, 1.0, 0]]))
    e1 = np.cross(nrm, helper)
    e1 /= np.linalg.norm(e1, axis=1, keepdims=True)
    e2 = np.cross(nrm, e1)
    th0 = rng.uniform(n, 0.0, 2 * np.pi)
    rad = L / math.sqrt(3.0)
    pts = []
    for k in range(3):
        th = th0 + 2 * np.pi * k / 3
        pts.append(c + rad[:, None] * (np.cos(th)[:, None] * e1 + np.sin(th)[:, None] * e2))
    P = np.stack(pts, 1).reshape(-1, 3).astype(np.float32)
    T = np.arange(3 * n, dtype=np.int32).reshape(n, 3)
    return P, T


# --------------------------------------------------------------------------- batches
@dataclass
class MeshSpec:
    kind: str
    params: tuple
    res: int

    def build(self):
        if self.kind == "ico":
            return icosphere(*self.params)
        if self.kind == "torus":
            return torus(*self.params)
        raise ValueError(self.kind)

    @property
    def n_tris(self) -> int:
        if self.kind == "ico":
            return 20 * 4 ** self.params[0]
        return 2 * self.params[2] * self.params[3]


def mixed_specs(n: int = 256, seed: int = 3, res_choices=(64, 128, 256)) -> list[MeshSpec]:
    """Half icospheres (subdivision 2-6), half tori (R ~ U(0.3,0.45),
    r ~ U(0.05,0.2), 16-512 segments per ring), triangle counts log-uniform
    in [320, 81920]; resolution uniform over res_choices."""
    rng = Rng(seed)
    kinds = rng.integers(n, 0, 1)
    logt = rng.uniform(n, math.log(320), math.log(81920))
    res = rng.integers(n, 0, len(res_choices) - 1)
    Rr = rng.uniform(n, 0.3, 0.45)
    rr = rng.uniform(n, 0.05, 0.2)
    ratio = rng.uniform(n, 1.0, 4.0)
    out = []
    for k in range(n):
        target = math.exp(logt[k])
        rs = int(res_choices[res[k]])
        if kinds[k] == 0:
            s = int(round(math.log(target / 20.0, 4)))
            out.append(MeshSpec("ico", (min(6, max(2, s)),), rs))
        else:
            half = target / 2.0
            b = int(round(math.sqrt(half / ratio[k])))
            b = min(512, max(16, b))
            a = int(round(half / b))
            a = min(512, max(16, a))
            out.append(MeshSpec("torus", (float(Rr[k]), float(rr[k]), a, b), rs))
    return out


def mixed_batch(n: int = 256, seed: int = 3, res_choices=(64, 128, 256)):
    """List of (points, tris, res) for the mixed batch workload."""
    cache: dict = {}
    out = []
    for s in mixed_specs(n, seed, res_choices):
        key = (s.kind, s.params)
        if key not in cache:
            cache[key] = s.build()
        P, T = cache[key]
        out.append((P, T, s.res))
    return out


def request_stream(n: int, rate: float = 2000.0, seed: int = 4, small_frac: float = 0.05,
                   res_choices=(64, 128, 256, 512)):
    """Poisson arrivals: list of (arrival_time_s, MeshSpec)."""
    rng = Rng(seed)
    gaps = -np.log1p(-rng.uniform(n)) / rate
    t = np.cumsum(gaps)
    base = mixed_specs(n, seed + 100, res_choices)
    small = rng.uniform(n)
    res = rng.integers(n, 0, len(res_choices) - 1)
    out = []
    for k in range(n):
        s = base[k]
        if small[k] < small_frac:
            s = MeshSpec("ico", (4,), int(res_choices[res[k]]))
        out.append((float(t[k]), s))
    return out
