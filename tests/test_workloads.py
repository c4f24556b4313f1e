"""Benchmark workload generators (host only)."""
import numpy as np

from oxpt import workloads


def test_serving_stream_refines_follow_their_render():
    reqs = workloads.serving_stream(duration=30.0, rate=20.0, pool=64, refine_frac=0.25, seed=1)
    seen = {}
    n_ref = 0
    for i, r in enumerate(reqs):
        assert r.arrival <= 30.0 + 1e3
        if r.kind == "render":
            assert 0 <= r.job < 64
            seen[r.seq] = i
        else:
            n_ref += 1
            assert r.job in seen, "a refine arrives before the render it continues"
            assert r.spp == 4
    arr = np.array([r.arrival for r in reqs])
    assert np.all(np.diff(arr) >= 0)
    n_render = len(seen)
    assert abs(n_render - 600) < 4 * np.sqrt(600)          # Poisson count at 20 req/s for 30 s
    assert abs(n_ref / n_render - 0.25) < 0.06


def test_uneven_draws_cover_the_stated_ranges():
    rng = np.random.default_rng(0)
    d = [workloads.draw_uneven(rng) for _ in range(4000)]
    tris = np.array([x[0] for x in d])
    side = np.array([x[1] for x in d])
    spp = np.array([x[2] for x in d])
    assert tris.min() >= 1000 and tris.max() <= 2_000_000
    assert abs(np.median(np.log(tris)) - 0.5 * (np.log(1e3) + np.log(2e6))) < 0.2   # log-uniform
    assert side.min() >= 128 and side.max() <= 1024
    assert set(np.unique(spp)) == {4, 16, 64, 256}
