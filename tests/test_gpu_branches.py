"""Estimator and resolve options against the CPU oracle on the GPU.

Each option branch of trace_sample and fs_resolve_pt the reference supports
is rendered on the device and by the oracle at equal spp, and compared with
the reference's parity method (tests/pathtrace.rs:88-142: PSNR floor plus
image and 4 x 4 region means within 4 standard errors + 1 %):

* firefly clamp (pathtrace.wgsl:1672-1679),
* constant ambient sky on a miss (pathtrace.wgsl:1780-1798),
* the orthographic camera (pathtrace.wgsl:1715-1764, Params.eye.w),
* an environment-only scene, P_env = 1 (src/pathtrace.rs:873-877),
* light-only and BSDF-only strategies at pixel level,
* Reinhard / ACES tone maps with exposure != 1 (resolve.wgsl:53-75), through
  image() and through the device resolve of the same accumulator,
* 2048^2 textures with full mip chains spread over several atlas layers.
"""
import numpy as np
import pytest

import oracle
import oxpt
import ptscenes
from ptutil import oracle_frame, parity, resolve_params

pytestmark = pytest.mark.gpu


def terrain(n_emissive=64):
    return oxpt.Scene.terrain(seed=3, gx=128, gz=64, n_spheres=24, subdiv=2, n_emissive=n_emissive)


def _gpu_vs_oracle(gpu, name, scene, opts, env=None, min_psnr=60.0):
    gpu.set_environment(env)
    gh = gpu.render_hdr(scene, opts)
    gi = gpu.image()
    _, ci, ch = oracle_frame(scene, opts, env)
    p = parity(name, gh, gi, ch, ci, min_psnr)
    print(f"{name}: PSNR {p:.2f} dB")
    return gh, ch


def test_firefly_clamp(gpu):
    """clamp > 0 caps each bounce's contribution (k = 0 emission unclamped):
    parity with the oracle, and the clamp visibly lowers the mean."""
    env = oxpt.EnvironmentMap.sky(256, 128, 5)
    opts = oxpt.RenderOptions(width=128, height=80, samples_per_pixel=16, max_bounces=8, clamp=0.6)
    gh, _ = _gpu_vs_oracle(gpu, "clamp", terrain(), opts, env)
    gpu.set_environment(env)
    free = gpu.render_hdr(terrain(), oxpt.RenderOptions(width=128, height=80, samples_per_pixel=16, max_bounces=8))
    assert gh[..., :3].mean() < free[..., :3].mean()


def test_ambient_sky_on_miss(gpu):
    """No environment, ambient > 0: misses return the constant sky."""
    opts = oxpt.RenderOptions(width=128, height=96, samples_per_pixel=16, max_bounces=8, ambient=0.35)
    gh, _ = _gpu_vs_oracle(gpu, "ambient", ptscenes.materials_scene(), opts)
    gpu.set_environment(None)
    dark = gpu.render_hdr(ptscenes.materials_scene(),
                          oxpt.RenderOptions(width=128, height=96, samples_per_pixel=16, max_bounces=8))
    assert gh[..., :3].mean() > dark[..., :3].mean() + 0.05


def test_orthographic_camera(gpu):
    scene = ptscenes.sphere_grid()
    scene.set_camera((0.0, 0.6, 5.0), (0.0, 0.4, 0.0), orthographic=True, ymag=1.6)
    opts = oxpt.RenderOptions(width=128, height=96, samples_per_pixel=16, max_bounces=8)
    gpu.set_environment(None)
    gpu.sync(scene, opts)
    assert gpu.params().eye[3] != 0.0, "Params.eye.w must flag the orthographic camera"
    _gpu_vs_oracle(gpu, "ortho", scene, opts)


def test_environment_only_scene(gpu):
    """No emissive triangles, no punctual lights, an environment: every NEE
    sample goes to the environment (P_env = 1)."""
    scene = terrain(n_emissive=0)
    env = oxpt.EnvironmentMap.sky(256, 128, 7)
    opts = oxpt.RenderOptions(width=128, height=80, samples_per_pixel=16, max_bounces=8)
    pk = oxpt.Packed(scene, opts, env)
    assert pk.params.cfg[3] == 0 and pk.params.cfg[2] == 0
    assert pk.params.fparams[2] == 1.0
    _gpu_vs_oracle(gpu, "env-only", scene, opts, env)


@pytest.mark.parametrize("strategy", ["light", "bsdf"])
def test_single_strategy_pixels(gpu, strategy):
    """Light-only and BSDF-only estimators on every lobe and light kind."""
    opts = oxpt.RenderOptions(width=128, height=96, samples_per_pixel=32, max_bounces=8, strategy=strategy)
    _gpu_vs_oracle(gpu, f"materials-{strategy}", ptscenes.materials_scene(), opts)


@pytest.mark.parametrize("tone_map,exposure", [("reinhard", 1.7), ("aces", 0.6), ("aces", 2.5),
                                               ("clamp", 0.45)])
def test_tone_maps_and_exposure(gpu, tone_map, exposure):
    """image() with a tone map and exposure vs the oracle's render + resolve,
    and the device resolve of one accumulator vs the oracle's resolve of the
    same accumulator (bytes within 1, almost all equal; HDR to f32 rounding)."""
    env = oxpt.EnvironmentMap.sky(256, 128, 5)
    opts = oxpt.RenderOptions(width=128, height=80, samples_per_pixel=16, max_bounces=8, tone_map=tone_map,
                              exposure=exposure, background=(40, 60, 90, 200))
    scene = terrain()
    gpu.set_environment(env)
    gi = gpu.render(scene, opts)
    gh = gpu.hdr()
    _, ci, ch = oracle_frame(scene, opts, env)
    p = parity(f"{tone_map}x{exposure}", gh, gi, ch, ci, 60.0)
    print(f"{tone_map} x {exposure}: PSNR {p:.2f} dB")
    # Same accumulator, both resolves (partial coverage from a sparse refine).
    dev = oxpt.Device(0)
    pk = oxpt.Packed(scene, opts, env)
    ds = dev.upload(pk)
    acc = dev.accum(opts.width, opts.height)
    dev.refine(ds, acc, ds.params, 0, 3)
    dev.finish()
    a = acc.read()
    for hdr in (False, True):
        rp = resolve_params(opts, 3, hdr)
        got = acc.resolve(rp)
        want = oracle.resolve(a, rp)
        if hdr:
            assert np.allclose(got, want, rtol=1e-6, atol=1e-7)
        else:
            d = np.abs(got.astype(np.int32) - want.astype(np.int32))
            assert d.max() <= 1 and (d > 0).mean() < 1e-3, (d.max(), (d > 0).mean())


def test_large_textures_over_several_atlas_layers(gpu):
    """2048^2 value-noise textures with full mip chains: the atlas needs
    several 2048^2 layers; sampling, LOD selection and layer addressing match
    the oracle at a small frame size."""
    scene = oxpt.Scene.layered(seed=8, n_tris=20000, n_textures=2, tex_size=2048, n_lights=4)
    opts = oxpt.RenderOptions(width=96, height=64, samples_per_pixel=8, max_bounces=8)
    pk = oxpt.Packed(scene, opts)
    assert pk.atlas_side == 2048 and pk.atlas_layers >= 2, (pk.atlas_side, pk.atlas_layers)
    gpu.set_environment(None)
    gpu.sync(scene, opts)
    assert gpu.atlas_shape() == (pk.atlas_side, pk.atlas_layers)
    _gpu_vs_oracle(gpu, "layered-2048", scene, opts)
