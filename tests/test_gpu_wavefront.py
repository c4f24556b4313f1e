"""Wavefront kernels (tuning kernel 7): bit-identical accumulators to the megakernel."""
import numpy as np
import pytest

import oxpt
import ptscenes
from oxpt import workloads

pytestmark = pytest.mark.gpu


def _scenes():
    yield "room", oxpt.Scene.room(0), None, oxpt.RenderOptions(width=96, height=80, samples_per_pixel=8)
    yield ("terrain+env", oxpt.Scene.terrain(seed=3, gx=96, gz=48, n_spheres=12, subdiv=2, n_emissive=64),
           oxpt.EnvironmentMap.sky(128, 64, 5), oxpt.RenderOptions(width=88, height=56, samples_per_pixel=6,
                                                                   max_bounces=12))
    yield "materials", ptscenes.materials_scene(), None, oxpt.RenderOptions(width=64, height=48, samples_per_pixel=8)
    yield "alpha", ptscenes.alpha_planes(), None, oxpt.RenderOptions(width=64, height=48, samples_per_pixel=8)
    yield ("layered", oxpt.Scene.layered(seed=2, n_tris=20000, n_textures=3, tex_size=128, n_lights=16), None,
           oxpt.RenderOptions(width=72, height=40, samples_per_pixel=4, max_bounces=16))


# wf_variant: traversal / shade occupancy forms; bvh=2: binary tree walk.
@pytest.mark.parametrize("case,variant,bvh", [(c, 0x00, 0) for c in range(5)] +
                         [(1, 0x02, 0), (1, 0x03, 0), (1, 0x11, 0), (1, 0x00, 2), (3, 0x00, 2)])
def test_wavefront_equals_megakernel(gpu, case, variant, bvh):
    name, scene, env, opts = list(_scenes())[case]
    gpu.set_environment(env)
    gpu.set_tuning(kernel=0)
    a = gpu.render_hdr(scene, opts)
    gpu.set_tuning(kernel=7, wf_variant=variant, bvh=bvh)
    b = gpu.render_hdr(scene, opts)
    # Progressive split through the wavefront too.
    gpu.reset()
    gpu.refine(1)
    gpu.refine(opts.samples_per_pixel - 1)
    c = gpu.hdr()
    gpu.set_tuning(kernel=0)
    gpu.set_environment(None)
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), name
    assert np.array_equal(a.view(np.uint32), c.view(np.uint32)), name + " (split)"


def test_wavefront_batch_equals_alone():
    dev = oxpt.Device(0)
    jobs = workloads.uneven_batch(n=8, seed=5, threads=8, scale=0.01)
    ds = [dev.upload(j.packed) for j in jobs]
    acc = [dev.accum(j.opts.width, j.opts.height) for j in jobs]
    alone = []
    for j, s, a in zip(jobs, ds, acc):
        dev.refine(s, a, s.params, 0, j.opts.samples_per_pixel)
        alone.append(a.read())
        a.clear()
    for variant in (0x00, 0x11):
        for a in acc:
            a.clear()
        dev.set_tuning(kernel=7, wf_variant=variant)
        dev.refine_batch([(s, a, s.params, 0, j.opts.samples_per_pixel) for j, s, a in zip(jobs, ds, acc)])
        for i, a in enumerate(acc):
            assert np.array_equal(a.read().view(np.uint32), alone[i].view(np.uint32)), (i, variant)
    # Passes: a small path cap splits the samples over several passes.
    for a in acc:
        a.clear()
    dev.set_tuning(kernel=7, max_paths=50000)
    dev.refine_batch([(s, a, s.params, 0, j.opts.samples_per_pixel) for j, s, a in zip(jobs, ds, acc)])
    for i, a in enumerate(acc):
        assert np.array_equal(a.read().view(np.uint32), alone[i].view(np.uint32)), f"{i} (passes)"
