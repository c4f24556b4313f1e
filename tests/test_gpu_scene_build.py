"""Device scene preparation (oxpt_scene_build): the BVH, the leaf-ordered
triangle / attribute words, the 4-wide nodes and the f16 texture atlas
built on the GPU are the host packer's words bit for bit, and so the
images are the same.

The reference prepares its scene on the CPU (GpuPathTracer::upload,
src/pathtrace.rs:970-1239, with oxideav-mesh3d's Bvh); the host packer
restates that step, and its tree is pinned by the brute-force closest-hit
tests in test_oracle.py.
"""
import numpy as np
import pytest

import oxpt
import ptscenes

pytestmark = pytest.mark.gpu

OPTS = oxpt.RenderOptions(width=48, height=32, samples_per_pixel=4, max_bounces=6)


def coincident(n):
    """n identical triangles: every centroid coincides, so every split is an
    object median (at sizes above and below the one-wave subtree limit)."""
    s = oxpt.Scene()
    m = s.add_material(oxpt.material(base_color=(0.6, 0.5, 0.4, 1.0)))
    pos = np.tile(np.array([0, 0, 0, 1, 0, 0, 0, 1, 0], np.float32), (n, 1))
    s.add_triangles(pos, np.full(n, m, np.uint32))
    s.set_camera(eye=(0.3, 0.3, 3.0), target=(0.3, 0.3, 0.0), yfov=0.8)
    return s


def signed_zeros():
    """Coordinates of both signs of zero and a NaN vertex: box unions must
    not depend on the order triangles are visited in."""
    rng = np.random.default_rng(7)
    n = 3000
    pos = rng.uniform(-1, 1, (n, 9)).astype(np.float32)
    z = rng.random((n, 9)) < 0.3
    pos[z] = 0.0
    pos[z & (rng.random((n, 9)) < 0.5)] = -0.0
    pos[17, 4] = np.nan
    s = oxpt.Scene()
    m = s.add_material(oxpt.material(base_color=(0.5, 0.5, 0.5, 1.0)))
    s.add_triangles(pos, np.full(n, m, np.uint32))
    return s


def flat_grid(n):
    """Triangles on the plane y = 0: one centroid axis has no extent."""
    xs = np.arange(n, dtype=np.float32)
    pos = np.zeros((n, 9), np.float32)
    pos[:, 0] = xs
    pos[:, 3] = xs + 1
    pos[:, 8] = 1.0
    pos[:, 6] = xs
    s = oxpt.Scene()
    m = s.add_material(oxpt.material())
    s.add_triangles(pos, np.full(n, m, np.uint32))
    return s


def textures_odd():
    """Odd-sized textures bound in both colour spaces (sRGB base colour,
    linear metallic-roughness / normal), and one wider than the largest
    atlas layer, whose level 0 is dropped."""
    rng = np.random.default_rng(11)
    s = oxpt.Scene()
    tex = [s.add_texture(rng.integers(0, 256, (h, w, 4), dtype=np.uint8), smp)
           for (w, h), smp in (((37, 19), (2, 2, 1, 5)), ((5, 300), (0, 1, 0, 2)), ((5000, 4), (2, 2, 1, 4)))]
    mats = []
    for i, t in enumerate(tex):
        m = oxpt.material(base_color=(0.8, 0.7, 0.6, 1.0), metallic=0.1, roughness=0.6)
        m.tex[0].texture = t          # base colour: sRGB
        m.tex[2].texture = t          # metallic-roughness: linear
        m.tex[1].texture = tex[(i + 1) % 3]  # normal map: linear
        mats.append(s.add_material(m))
    pos, uv = ptscenes.quad((-1, 0, -1), (1, 0, -1), (1, 0, 1), (-1, 0, 1))
    for i, m in enumerate(mats):
        s.add_triangles(pos + np.float32(0.1 * i), m, uv0=uv)
    return s


def emissive_many():
    """Hundreds of emissive triangles of uneven area, some of zero area,
    an unlit emitter, a material with a negative emission channel and a
    black one: the lights, their order and their CDF are found on the
    device."""
    rng = np.random.default_rng(3)
    s = oxpt.Scene()
    mats = [s.add_material(oxpt.material(base_color=(0.5, 0.5, 0.5, 1.0), emissive=tuple(e), unlit=u))
            for e, u in (((4.0, 3.0, 2.0), 0), ((0.0, 0.0, 0.0), 0), ((2.0, 2.0, 2.0), 1), ((3.0, -1.0, 0.5), 0),
                         ((0.0, 0.0, 7.0), 0))]
    n = 900
    pos = rng.uniform(-2, 2, (n, 9)).astype(np.float32)
    pos[::37, 6:9] = pos[::37, 3:6]  # degenerate: zero area
    s.add_triangles(pos, rng.integers(0, len(mats), n).astype(np.uint32))
    return s


SCENES = {
    "emissive_many": emissive_many,
    "textures_odd": textures_odd,
    "checker": ptscenes.checker_floor,
    "alpha": ptscenes.alpha_planes,
    "room": lambda: oxpt.Scene.room(0),
    "single": lambda: coincident(1),
    "coincident_small": lambda: coincident(300),
    "coincident_large": lambda: coincident(5000),
    "signed_zeros_nan": signed_zeros,
    "flat": lambda: flat_grid(2000),
    "materials": ptscenes.materials_scene,
    "terrain": lambda: oxpt.Scene.terrain(seed=3, gx=256, gz=128, n_spheres=24, subdiv=3, n_emissive=16),
    "layered": lambda: oxpt.Scene.layered(seed=2, n_tris=60000, n_textures=2, tex_size=64, n_lights=3),
}


def device_tracer(scene, opts, env=None):
    pt = oxpt.GpuPathTracer(0)
    pt.set_device_build(True)
    if env is not None:
        pt.set_environment(env)
    pt.sync(scene, opts)
    return pt


@pytest.mark.parametrize("name", sorted(SCENES))
def test_device_build_writes_the_host_words(name):
    scene = SCENES[name]()
    print(f"device build: {name}, {scene.triangle_count()} triangles", flush=True)
    pk = oxpt.Packed(scene, OPTS)
    pt = device_tracer(scene, OPTS)
    info = pt.build_info()
    P = pt.params()
    assert bytes(P) == bytes(pk.params), f"{name}: Params differ"
    assert info.node_count == pk.params.env[3] and info.tri_base == pk.params.offs2[2]
    assert info.wide_base == pk.params.offs2[3]
    assert info.bvh_depth == pk.bvh_depth
    for which, host in ((0, pk.geom), (1, pk.attrs), (2, pk.tables)):
        dev = pt.scene_words(which)
        assert dev.size == host.size, f"{name}: word count {which}: {dev.size} vs {host.size}"
        hb, db = host.view(np.uint32), dev.view(np.uint32)
        bad = np.flatnonzero(hb != db)
        assert bad.size == 0, f"{name}: buffer {which} differs at {bad[:8]} ({bad.size} words)"
    dev = pt.scene_words(3)
    assert dev.size == pk.atlas.size
    bad = np.flatnonzero(dev != pk.atlas)
    assert bad.size == 0, f"{name}: atlas differs at {bad[:8]} ({bad.size} texel words)"
    assert (info.atlas_ms > 0.0) == (pk.atlas.size > 4)  # a scene without textures keeps one blank texel
    # The host copies the tracer hands the oracle are the device words.
    assert np.array_equal(pt.words(0).view(np.uint32), pk.geom.view(np.uint32))
    assert np.array_equal(pt.words(3), pk.atlas)


def test_device_built_scene_renders_the_same_image():
    scene = oxpt.Scene.terrain(seed=5, gx=128, gz=64, n_spheres=12, subdiv=2, n_emissive=8)
    env = oxpt.EnvironmentMap.sky(64, 32, 1)
    opts = oxpt.RenderOptions(width=96, height=64, samples_per_pixel=8, max_bounces=8)
    host = oxpt.GpuPathTracer(0)
    host.set_environment(env)
    host.sync(scene, opts)
    host.refine(8)
    dev = device_tracer(scene, opts, env)
    dev.refine(8)
    assert np.array_equal(host.hdr().view(np.uint32), dev.hdr().view(np.uint32))
    assert np.array_equal(host.image(), dev.image())


def test_device_build_switch_and_empty_scene():
    """Switching the mode re-prepares the scene; an empty scene packs on the
    host (nothing to build) and renders the background."""
    scene = oxpt.Scene.room(1)
    pt = oxpt.GpuPathTracer(0)
    pt.sync(scene, OPTS)
    assert pt.build_info().node_count == 0
    pt.set_device_build(True)
    assert pt.sync(scene, OPTS)
    assert pt.build_info().node_count > 0 and pt.build_info().build_ms > 0.0
    empty = oxpt.Scene()
    pt.invalidate_scene()
    pt.sync(empty, OPTS)
    pt.refine(1)
    img = pt.image()
    assert (img[..., :3] == np.array(OPTS.background[:3], np.uint8)).all()


def env_maps():
    rng = np.random.default_rng(9)
    noisy = rng.uniform(-0.5, 3.0, (37, 64, 3)).astype(np.float32)
    noisy[3, 5] = (np.nan, 1.0, 1.0)
    noisy[7, 9, 1] = np.inf
    noisy[11] = 0.0  # a black row
    return {
        "sky": oxpt.EnvironmentMap.sky(256, 128, 2),
        "noisy": oxpt.EnvironmentMap(noisy, intensity=1.7),
        "black": oxpt.EnvironmentMap(np.zeros((8, 16, 3), np.float32)),
    }


@pytest.mark.parametrize("env_name", ["sky", "noisy", "black"])
def test_device_environment_tables_equal_the_host(env_name):
    env = env_maps()[env_name]
    scene = oxpt.Scene.room(0)
    pk = oxpt.Packed(scene, OPTS, env)
    pt = device_tracer(scene, OPTS, env)
    assert bytes(pt.params()) == bytes(pk.params)
    dev = pt.scene_words(2)
    assert dev.size == pk.tables.size
    bad = np.flatnonzero(dev != pk.tables)
    assert bad.size == 0, f"{env_name}: tables differ at {bad[:8]} ({bad.size} words)"
    assert np.array_equal(pt.words(2), pk.tables)
