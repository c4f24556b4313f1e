"""set_texture_resolver (src/pathtrace.rs:474-478, 1261-1263): textures named
by uri get their pixels from the tracer's resolver at prepare time. A scene
whose texture is resolved renders bit for bit like the same scene with the
pixels given directly; an unresolved texture is absent (as an unbound slot);
installing a resolver re-prepares and resets on the next sync."""
import numpy as np
import pytest

import oxpt
import ptscenes

pytestmark = pytest.mark.gpu

OPTS = oxpt.RenderOptions(width=96, height=64, samples_per_pixel=4, max_bounces=4)


def floor_scene(bind):
    """The textured checker floor; bind(scene) -> texture index or -1."""
    s = oxpt.Scene()
    m = oxpt.material(base_color=(1, 1, 1, 1), metallic=0.0, roughness=0.8)
    m.tex[0].texture = bind(s)
    m.tex[0].xform[0] = 8.0
    m.tex[0].xform[4] = 8.0
    mid = s.add_material(m)
    pos, uv = ptscenes.quad([-20, 0, 20], [20, 0, 20], [20, 0, -20], [-20, 0, -20])
    s.add_triangles(pos, mid, uv0=uv)
    s.add_light("directional", direction=(-0.4, -1.0, -0.3), color=(1, 1, 1), intensity=2.0)
    s.set_camera((0.0, 1.0, 6.0), (0.0, 0.0, -10.0), yfov=0.9)
    return s


def render(scene, resolver=None, set_resolver=False):
    pt = oxpt.GpuPathTracer(0)
    if set_resolver:
        pt.set_texture_resolver(resolver)
    return pt.render_hdr(scene, OPTS).view(np.uint32).copy(), pt


def test_resolved_texture_renders_like_direct_pixels():
    tex = ptscenes.checker()
    direct, _ = render(floor_scene(lambda s: s.add_texture(tex, (2, 2, 1, 5))))
    calls = []

    def resolver(uri):
        calls.append(uri)
        return tex if uri == "mem://checker" else None

    named, pt = render(floor_scene(lambda s: s.add_texture_ref("mem://checker", (2, 2, 1, 5))), resolver, True)
    assert np.array_equal(named, direct)
    assert calls == ["mem://checker"]
    # Decoded once: a re-prepare of the same scene does not call it again.
    pt.invalidate_scene()
    pt.render_hdr(floor_scene(lambda s: s.add_texture_ref("mem://checker", (2, 2, 1, 5))), OPTS)
    assert calls == ["mem://checker"]


def test_unresolved_texture_is_absent():
    plain, _ = render(floor_scene(lambda s: -1))
    nothing, _ = render(floor_scene(lambda s: s.add_texture_ref("mem://checker")))
    refused, _ = render(floor_scene(lambda s: s.add_texture_ref("mem://other")), lambda uri: None, True)
    failing, _ = render(floor_scene(lambda s: s.add_texture_ref("mem://x")), lambda uri: 1 / 0, True)
    assert np.array_equal(nothing, plain)
    assert np.array_equal(refused, plain)
    assert np.array_equal(failing, plain)


def test_installing_a_resolver_reprepares_and_resets():
    tex = ptscenes.checker()
    scene = floor_scene(lambda s: s.add_texture_ref("mem://checker"))
    pt = oxpt.GpuPathTracer(0)
    assert pt.sync(scene, OPTS)
    pt.refine(2)
    without = pt.hdr().view(np.uint32).copy()
    assert not pt.sync(scene, OPTS)  # nothing changed: no reset
    assert pt.samples() == 2
    pt.set_texture_resolver(lambda uri: tex)
    assert pt.sync(scene, OPTS)  # re-prepared: reset
    assert pt.samples() == 0
    pt.refine(2)
    with_tex = pt.hdr().view(np.uint32)
    assert not np.array_equal(with_tex, without)
    pt.set_texture_resolver(None)
    assert pt.sync(scene, OPTS)
    pt.refine(2)
    assert np.array_equal(pt.hdr().view(np.uint32), without)
