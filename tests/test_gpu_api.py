"""Progressive API semantics on the GPU (tests/pathtrace.rs:420-565)."""
import numpy as np
import pytest

import oxpt
from ptutil import oracle_frame

pytestmark = pytest.mark.gpu


def opts(w, h, spp, bounces):
    return oxpt.RenderOptions(width=w, height=h, samples_per_pixel=spp, max_bounces=bounces,
                              background=(16, 16, 20, 255))


def test_deterministic_for_a_seed(gpu):
    scene = oxpt.Scene.room(0)
    o = opts(64, 64, 8, 8)
    a = gpu.render(scene, o)
    b = gpu.render(scene, o)
    assert np.array_equal(a, b)
    o2 = opts(64, 64, 8, 8)
    o2.seed = 7
    c = gpu.render(scene, o2)
    assert not np.array_equal(a, c)


def test_progressive_refines_add_up(gpu):
    scene = oxpt.Scene.room(0)
    o = opts(64, 48, 0, 8)
    gpu.invalidate_scene()
    gpu.sync(scene, o)
    gpu.refine(3)
    gpu.refine(5)
    assert gpu.samples() == 8
    split = gpu.hdr()
    gpu.reset()
    assert gpu.samples() == 0
    gpu.refine(8)
    whole = gpu.hdr()
    assert np.array_equal(split.view(np.uint32), whole.view(np.uint32))


def test_reset_rules_match_the_cpu_tracer(gpu):
    scene = oxpt.Scene.room(0)
    o = opts(32, 32, 16, 4)
    gpu.set_environment(None)
    gpu.invalidate_scene()
    assert gpu.sync(scene, o), "first sync prepares"
    gpu.refine(2)
    assert not gpu.sync(scene, o), "no change, no reset"
    assert gpu.samples() == 2
    d = opts(32, 32, 999, 4)
    d.background = (200, 0, 0, 255)
    d.tone_map = "aces"
    d.exposure = 2.0
    assert not gpu.sync(scene, d)
    assert gpu.samples() == 2
    assert not gpu.is_converged()
    changes = [("seed", 3), ("max_bounces", 2), ("width", 40), ("ambient", 0.5)]
    for field, value in changes:
        gpu.refine(1)
        c = opts(32, 32, 999, 4)
        c.background, c.tone_map, c.exposure = d.background, d.tone_map, d.exposure
        setattr(c, field, value)
        assert gpu.sync(scene, c), field
        assert gpu.samples() == 0, field
        gpu.sync(scene, d)
    gpu.refine(1)
    gpu.invalidate_scene()
    assert gpu.sync(scene, d)
    assert gpu.samples() == 0
    gpu.refine(1)
    gpu.set_environment(oxpt.EnvironmentMap.sky(64, 32, 1))
    assert gpu.samples() == 0
    gpu.refine(1)
    assert gpu.samples() == 1
    gpu.set_environment(None)
    fresh = oxpt.GpuPathTracer(0)
    fresh.refine(4)
    assert fresh.samples() == 0


def test_uncovered_pixels_keep_background_bytes(gpu):
    # Camera looking past the room's side: many camera rays miss everything.
    scene = oxpt.Scene.room(0)
    scene.set_camera((3.0, 1.0, 3.0), (0.5, 1.0, 0.0), yfov=1.2)
    o = opts(64, 48, 4, 2)
    o.background = (37, 99, 201, 180)
    gi = gpu.render(scene, o)
    _, ci, _ = oracle_frame(scene, o)
    bgmask = (ci == np.array([37, 99, 201, 180], dtype=np.uint8)).all(axis=2)
    assert bgmask.sum() > 100, f"{bgmask.sum()} background pixels"
    assert np.array_equal(gi[bgmask], ci[bgmask])


def test_empty_scene_is_all_background(gpu):
    o = opts(16, 12, 4, 4)
    o.background = (1, 2, 3, 4)
    img = gpu.render(oxpt.Scene(), o)
    assert (img == np.array([1, 2, 3, 4], dtype=np.uint8)).all()


def test_registry_and_render_agree(gpu):
    scene = oxpt.Scene.room(0)
    o = opts(40, 30, 2, 2)
    gpu.invalidate_scene()
    gpu.sync(scene, o)
    gpu.refine(2)
    assert gpu.size() == (40, 30)
    a = gpu.render(scene, o)
    reg = oxpt.RenderRegistry()
    oxpt.register_into(reg)
    viareg = reg.make(oxpt.PATHTRACE_BACKEND_NAME)
    assert np.array_equal(viareg.render(scene, o), a)
    with pytest.raises(oxpt.InvalidOptions):
        reg.make("gpu")


def test_oversize_dimensions_are_invalid_options(gpu):
    scene = oxpt.Scene.room(0)
    with pytest.raises(oxpt.InvalidOptions):
        gpu.sync(scene, opts(40000, 8, 1, 1))


def test_hdr_budget_exceeded_is_backend_error():
    import ctypes as C

    pt = oxpt.GpuPathTracer(0)
    assert oxpt.lib.oxpt_ctx_set_budget(pt.ctx, C.c_uint64(1 << 20)) == 0
    with pytest.raises(oxpt.BackendError):
        pt.sync(oxpt.Scene.terrain(seed=1, gx=64, gz=64, n_spheres=4, subdiv=1, n_emissive=4), opts(64, 64, 1, 1))


def test_draw_into_device_memory_equals_the_readback():
    """draw_texture (src/pathtrace.rs:687-691): the resolve written into
    device memory without a readback holds the bytes image() / hdr() return;
    host memory and short buffers are refused."""
    scene = oxpt.Scene.room(0)
    opts = oxpt.RenderOptions(width=80, height=48, samples_per_pixel=4, tone_map="reinhard", exposure=1.3)
    pt = oxpt.GpuPathTracer(0)
    pt.sync(scene, opts)
    pt.refine(4)
    ldr = oxpt.DeviceBuffer(48 * 80 * 4)
    hdr = oxpt.DeviceBuffer(48 * 80 * 16)
    pt.draw(ldr.ptr, ldr.nbytes, hdr=False)
    pt.draw(hdr.ptr, hdr.nbytes, hdr=True)
    pt.finish()
    assert np.array_equal(ldr.read(np.uint8).reshape(48, 80, 4), pt.image())
    assert np.array_equal(hdr.read(np.uint32).reshape(48, 80, 4), pt.hdr().view(np.uint32))
    host = np.zeros(48 * 80 * 4, np.uint8)
    with pytest.raises(oxpt.Error):
        pt.draw(host.ctypes.data, host.size)
    with pytest.raises(oxpt.Error):
        pt.draw(ldr.ptr, ldr.nbytes - 1)
