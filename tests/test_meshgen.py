"""Synthetic workload generators: sizes of the benchmark configurations and
determinism (own splitmix64 stream, no numpy distribution code)."""
import numpy as np

from svo_amd import meshgen


def test_icosphere_counts():
    for s in range(6):
        P, T = meshgen.icosphere(s)
        assert T.shape == (20 * 4 ** s, 3) and P.shape == (10 * 4 ** s + 2, 3)
        assert np.allclose(np.linalg.norm(P, axis=1), 1, atol=1e-6)
        # closed 2-manifold: every edge used twice
        e = np.sort(np.concatenate([T[:, [0, 1]], T[:, [1, 2]], T[:, [2, 0]]]), axis=1)
        _, c = np.unique(e, axis=0, return_counts=True)
        assert np.all(c == 2)


def test_plumbing_config_sizes():
    P, T = meshgen.icosphere(4)
    assert (P.shape[0], T.shape[0]) == (2562, 5120)


def test_noisy_icosphere_deterministic():
    a = meshgen.noisy_icosphere(3, seed=1)
    b = meshgen.noisy_icosphere(3, seed=1)
    assert np.array_equal(a[0], b[0])
    r = np.linalg.norm(a[0], axis=1)
    assert r.min() >= 0.9 - 1e-6 and r.max() <= 1.1 + 1e-6
    assert not np.array_equal(a[0], meshgen.noisy_icosphere(3, seed=2)[0])


def test_torus_counts():
    P, T = meshgen.torus(0.4, 0.1, 32, 16)
    assert P.shape == (512, 3) and T.shape == (1024, 3)


def test_soup_distribution():
    P, T = meshgen.soup(20000, seed=2)
    tri = P[T]
    side = np.linalg.norm(tri[:, 0] - tri[:, 1], axis=1)
    assert abs(np.median(side) * 256 - 1) < 0.05
    assert abs(np.std(np.log(side)) - 0.6) < 0.02
    c = tri.mean(1)
    assert c.min() > -0.01 and c.max() < 1.01
    assert np.array_equal(P, meshgen.soup(20000, seed=2)[0])


def test_mixed_specs():
    specs = meshgen.mixed_specs(256, seed=3)
    t = np.array([s.n_tris for s in specs])
    assert len(specs) == 256 and t.min() >= 320 and t.max() <= 81920
    assert set(s.res for s in specs) == {64, 128, 256}
    assert 100 < sum(s.kind == "ico" for s in specs) < 156
    assert [s.params for s in specs] == [s.params for s in meshgen.mixed_specs(256, seed=3)]


def test_request_stream():
    rs = meshgen.request_stream(4000, rate=2000.0, seed=4)
    t = np.array([a for a, _ in rs])
    assert np.all(np.diff(t) >= 0)
    assert abs(t[-1] - 2.0) < 0.2
    # 5% of the requests are icosphere-4 (5,120 triangles), besides the ones
    # the mixed distribution draws itself
    small = sum(1 for _, s in rs if s.kind == "ico" and s.params == (4,))
    assert small >= 0.05 * len(rs)
    assert {s.res for _, s in rs} == {64, 128, 256, 512}
