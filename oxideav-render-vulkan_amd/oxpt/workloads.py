"""Seeded benchmark workloads.

* single_large(): one terrain scene of ~1.3 M triangles with a sky, 1920x1080,
  64 spp, depth 12.
* uneven_batch(): 64 scenes of the same generator type, triangle counts
  log-uniform over 1 k..2 M, square resolutions with side uniform over
  128..1024, spp uniform over {4, 16, 64, 256}, depth 12, per-scene seeds.
* serving_stream(): Poisson arrivals of render requests drawn from the uneven
  batch distribution plus progressive refine(4) requests on resident scenes.

Scenes are generated and packed on the host (in parallel threads: the C++
packer releases the GIL) before any timing starts.
"""
from __future__ import annotations

import math
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass
from typing import List, Optional

import numpy as np

from . import EnvironmentMap, Packed, RenderOptions, Scene


def terrain_of_size(n_tris: int, seed: int) -> Scene:
    """The single-large-scene generator scaled to about n_tris triangles:
    80 % heightfield (gx x gx/2 quads), 20 % subdivision-3 icospheres, 64
    emissive triangles."""
    gx = max(4, int(round(math.sqrt(0.8 * n_tris))))
    gz = max(2, gx // 2)
    spheres = int(round(0.2 * n_tris / 1280))
    return Scene.terrain(seed=seed, gx=gx, gz=gz, n_spheres=spheres, subdiv=3, n_emissive=64)


def single_large():
    scene = Scene.terrain(seed=1, gx=1024, gz=512, n_spheres=200, subdiv=3, n_emissive=64)
    env = EnvironmentMap.sky(1024, 512, 0)
    opts = RenderOptions(width=1920, height=1080, samples_per_pixel=64, max_bounces=12, rr_start=3)
    return scene, env, opts


@dataclass
class Job:
    scene: Scene
    env: EnvironmentMap
    opts: RenderOptions
    packed: Optional[Packed]
    n_tris: int

    @property
    def paths(self) -> int:
        return self.opts.width * self.opts.height * self.opts.samples_per_pixel


def draw_uneven(rng: np.random.Generator):
    """One draw of the uneven-batch distribution: (n_tris, side, spp)."""
    n_tris = int(round(math.exp(rng.uniform(math.log(1e3), math.log(2e6)))))
    side = int(rng.integers(128, 1025))
    spp = int(rng.choice([4, 16, 64, 256]))
    return n_tris, side, spp


def make_job(n_tris: int, side: int, spp: int, seed: int, pack: bool = True) -> Job:
    scene = terrain_of_size(n_tris, seed)
    env = EnvironmentMap.sky(1024, 512, seed)
    opts = RenderOptions(width=side, height=side, samples_per_pixel=spp, max_bounces=12, rr_start=3, seed=seed)
    pk = Packed(scene, opts, env) if pack else None
    return Job(scene, env, opts, pk, scene.triangle_count())


def uneven_batch(n: int = 64, seed: int = 3, threads: int = 16, scale: float = 1.0, pack: bool = True) -> List[Job]:
    """`scale` < 1 shrinks triangle counts and sides (for tests); `pack`
    False leaves the packing to the consumer (the server packs on add)."""
    rng = np.random.default_rng(seed)
    draws = [draw_uneven(rng) for _ in range(n)]
    if scale != 1.0:
        draws = [(max(200, int(t * scale)), max(16, int(s * math.sqrt(scale))), p) for t, s, p in draws]
    with ThreadPoolExecutor(max_workers=threads) as ex:
        return list(ex.map(lambda a: make_job(a[1][0], a[1][1], a[1][2], seed * 1000 + a[0], pack), enumerate(draws)))


@dataclass
class Request:
    arrival: float      # seconds from the start
    kind: str           # "render" or "refine"
    job: int            # pool scene (render) or the sequence number of the render a refine continues
    spp: int
    seq: int = 0        # sequence number of a render (renders only)


def serving_stream(duration: float, rate: float, pool: int, refine_frac: float = 0.25, seed: int = 9):
    """Poisson arrivals at `rate` req/s for `duration` s: renders of a pool
    scene; after each render a refine(4) follows with probability
    `refine_frac` (on the then-resident scene of that render)."""
    rng = np.random.default_rng(seed)
    t, out, n_render = 0.0, [], 0
    while True:
        t += rng.exponential(1.0 / rate)
        if t > duration:
            break
        out.append(Request(t, "render", int(rng.integers(0, pool)), 0, n_render))
        if rng.random() < refine_frac:
            out.append(Request(t + rng.exponential(1.0 / rate), "refine", n_render, 4))
        n_render += 1
    out.sort(key=lambda r: r.arrival)
    return out
