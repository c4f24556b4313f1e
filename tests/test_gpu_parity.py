"""GPU kernels vs the CPU oracle: bit-exact integers, closest hits, and images.

The integer sampler words and light-CDF indices must match bit for bit
(they are closed-form integer code, pathtrace.wgsl:158-228,1440-1484); the
images must reach PSNR >= 60 dB at equal spp, with the reference's
statistical region checks (tests/pathtrace.rs:88-142) on top.
"""
import numpy as np
import pytest

import oracle
import oxpt
from ptutil import oracle_frame, parity, psnr
import ptscenes

pytestmark = pytest.mark.gpu


def small_terrain():
    return oxpt.Scene.terrain(seed=3, gx=128, gz=64, n_spheres=24, subdiv=2, n_emissive=64)


def test_pattern_words_bit_exact(gpu):
    scene = oxpt.Scene.room(0)
    opts = oxpt.RenderOptions(width=64, height=64, samples_per_pixel=1)
    gpu.invalidate_scene()
    gpu.sync(scene, opts)
    gpu.refine(1)
    rng = np.random.default_rng(7)
    n = 1 << 20
    x = rng.integers(0, 4096, n, dtype=np.uint32)
    y = rng.integers(0, 4096, n, dtype=np.uint32)
    idx = rng.integers(0, 1 << 32, n, dtype=np.uint64).astype(np.uint32)
    idx[: n // 4] = np.arange(n // 4, dtype=np.uint32)  # low sample indices, the common case
    p = rng.integers(0, 40, n, dtype=np.uint32)
    for seed in (0, 7, 0xDEADBEEF):
        got = gpu.debug_pattern_words(seed, x, y, idx, p)
        want = oracle.pattern_words(oracle.sobol_directions(), seed, x, y, idx, p)
        bad = np.nonzero((got != want).any(axis=1))[0]
        assert bad.size == 0, f"seed {seed}: {bad.size} mismatches, first at tuple {bad[:4]}"


def test_light_cdf_indices_bit_exact(gpu):
    scene = small_terrain()
    env = oxpt.EnvironmentMap.sky(256, 128, 5)
    opts = oxpt.RenderOptions(width=32, height=32, samples_per_pixel=1, max_bounces=4)
    gpu.set_environment(env)
    gpu.invalidate_scene()  # a different Scene3D must be signalled (README: sync rules)
    gpu.sync(scene, opts)
    gpu.refine(1)
    pk = oxpt.Packed(scene, opts, env)
    osc = oracle.Scene.from_packed(pk)
    P = gpu.params()
    assert P.cfg[3] == 64 and P.env[2] == 1
    rng = np.random.default_rng(11)
    u = rng.random(1 << 20, dtype=np.float32)
    # Edges: 0, values just below 1, and every CDF value exactly (ties go right).
    t = pk.tables
    cdf = t[P.offs1[0]: P.offs1[0] + 4 * P.cfg[3]: 4].view(np.float32)
    marg = t[P.offs1[2]: P.offs1[2] + P.env[1] + 1].view(np.float32)
    edges = np.concatenate([[0.0, 0.999999, np.nextafter(np.float32(1), np.float32(0))], cdf, marg,
                            np.nextafter(cdf, np.float32(0)), np.nextafter(marg, np.float32(0))]).astype(np.float32)
    u = np.concatenate([edges, u]).astype(np.float32)
    got = gpu.debug_light_picks(u)
    want = osc.light_picks(u)
    bad = np.nonzero((got != want).any(axis=1))[0]
    assert bad.size == 0, f"{bad.size} mismatches; first u={u[bad[:4]]} got {got[bad[:4]]} want {want[bad[:4]]}"
    gpu.set_environment(None)


def test_closest_hit_matches_oracle(gpu):
    scene = small_terrain()
    opts = oxpt.RenderOptions(width=32, height=32, samples_per_pixel=1)
    gpu.set_environment(None)
    gpu.invalidate_scene()
    gpu.sync(scene, opts)
    gpu.refine(1)
    pk = oxpt.Packed(scene, opts)
    osc = oracle.Scene.from_packed(pk)
    rng = np.random.default_rng(3)
    n = 4096
    o = np.stack([rng.uniform(-6, 6, n), rng.uniform(2, 8, n), rng.uniform(-3, 3, n)], axis=1).astype(np.float32)
    d = rng.normal(size=(n, 3)).astype(np.float32)
    d[:, 1] = -np.abs(d[:, 1])
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    got = gpu.debug_closest_hit(o, d)
    want = np.array([osc.closest_hit(o[i], d[i]) for i in range(n)], dtype=object)
    ws = np.array([w[0] for w in want], dtype=np.uint32)
    wt = np.array([w[1] for w in want], dtype=np.float32).view(np.uint32)
    wt[ws == 0xFFFFFFFF] = 0
    assert (got[:, 0] == ws).all(), f"slot mismatches: {(got[:, 0] != ws).sum()}"
    assert (got[:, 1] == wt).all(), f"t mismatches: {(got[:, 1] != wt).sum()}"
    assert (ws != 0xFFFFFFFF).sum() > n // 4


def test_wide_and_binary_walks_agree(gpu):
    """The 4-wide walk reaches the same leaves through the same boxes as the
    binary walk: same closest hits (slot and t bits) on random rays, and the
    same image."""
    scene = small_terrain()
    opts = oxpt.RenderOptions(width=96, height=64, samples_per_pixel=4, max_bounces=8)
    gpu.set_environment(oxpt.EnvironmentMap.sky(256, 128, 5))
    gpu.invalidate_scene()
    gpu.sync(scene, opts)
    gpu.refine(1)
    assert gpu.params().offs2[3] > 0, "scene carries no 4-wide tree"
    rng = np.random.default_rng(5)
    n = 1 << 16
    o = np.stack([rng.uniform(-6, 6, n), rng.uniform(0.5, 8, n), rng.uniform(-3, 3, n)], axis=1).astype(np.float32)
    d = rng.normal(size=(n, 3)).astype(np.float32)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    hits = {}
    imgs = {}
    for bvh in (0, 2):
        gpu.set_tuning(bvh=bvh)
        hits[bvh] = gpu.debug_closest_hit(o, d)
        gpu.reset()
        gpu.refine(4)
        imgs[bvh] = gpu.hdr().view(np.uint32).copy()
    gpu.set_tuning(bvh=0)
    gpu.set_environment(None)
    assert (hits[0][:, 0] != 0xFFFFFFFF).sum() > n // 8
    assert np.array_equal(hits[0], hits[2]), f"{(hits[0] != hits[2]).any(axis=1).sum()} rays differ"
    assert np.array_equal(imgs[0], imgs[2])


def test_room_full_frame_psnr(gpu):
    """Plain small workload: the room at 256x256, 16 spp, depth 8 — full frame vs the oracle."""
    scene = oxpt.Scene.room(0)
    opts = oxpt.RenderOptions(width=256, height=256, samples_per_pixel=16, max_bounces=8)
    gh = gpu.render_hdr(scene, opts)
    gi = gpu.image()
    _, ci, ch = oracle_frame(scene, opts)
    p = parity("room", gh, gi, ch, ci, 60.0)
    # Against the f64 light sampling of the reference's CPU tracer.
    _, ci64, ch64 = oracle_frame(scene, opts, flags=oracle.F64_LIGHTS)
    p64 = parity("room-f64", gh, gi, ch64, ci64, 60.0)
    print(f"room PSNR {p:.2f} dB (f32 oracle), {p64:.2f} dB (f64-light oracle)")


def test_white_furnace(gpu):
    """tests/pathtrace.rs:295-317: L = Le / (1 - rho) = 2 for every strategy."""
    scene = oxpt.Scene.furnace(0.5, 1.0)
    for strategy in ("mis", "light", "bsdf"):
        opts = oxpt.RenderOptions(width=8, height=8, samples_per_pixel=64, max_bounces=64, strategy=strategy,
                                  ambient=0.0)
        img = gpu.render_hdr(scene, opts)
        m = img[..., :3].mean(axis=(0, 1))
        assert np.all(np.abs(m - 2.0) < 0.04), f"{strategy}: {m}"


def test_terrain_env_psnr(gpu):
    """MR materials + 64 emissive triangles + sky CDF, full frame vs the oracle."""
    scene = small_terrain()
    env = oxpt.EnvironmentMap.sky(256, 128, 5)
    opts = oxpt.RenderOptions(width=160, height=96, samples_per_pixel=16, max_bounces=12)
    gpu.set_environment(env)
    gh = gpu.render_hdr(scene, opts)
    gi = gpu.image()
    gpu.set_environment(None)
    _, ci, ch = oracle_frame(scene, opts, env)
    p = psnr(gi, ci)
    print(f"terrain+env PSNR {p:.2f} dB")
    parity("terrain", gh, gi, ch, ci, 60.0)


def test_plain_kernel_matches_persistent(gpu):
    """Both launch forms add each pixel's samples in order: identical
    accumulators. The persistent form is forced (kernel 3: 4 waves per SIMD),
    since a default refine of a frame smaller than the GPU's resident lanes
    runs as a one-job batch; the kernel that ran is asserted."""
    scene = oxpt.Scene.room(0)
    opts = oxpt.RenderOptions(width=72, height=40, samples_per_pixel=6, max_bounces=8)
    gpu.set_tuning(kernel=1)
    a = gpu.render_hdr(scene, opts)
    assert gpu.last_kernel() == "trace_tiles", gpu.last_kernel()
    gpu.set_tuning(kernel=3)
    b = gpu.render_hdr(scene, opts)
    assert gpu.last_kernel().startswith("trace_persistent"), gpu.last_kernel()
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    gpu.set_tuning(kernel=0)
    c = gpu.render_hdr(scene, opts)
    assert gpu.last_kernel().startswith("trace_batch"), gpu.last_kernel()
    assert np.array_equal(a.view(np.uint32), c.view(np.uint32))


def test_default_form_of_a_large_frame_is_persistent(gpu):
    """A frame of at least the GPU's resident lanes (512 x 544 px > 256 CUs x
    4 SIMDs x 4 waves x 64 lanes) takes the default persistent megakernel with
    its path tail — the form the single-large-scene bench measures — and
    equals the plain per-pixel kernel bit for bit."""
    scene = small_terrain()
    env = oxpt.EnvironmentMap.sky(128, 64, 2)
    opts = oxpt.RenderOptions(width=512, height=544, samples_per_pixel=3, max_bounces=8)
    gpu.set_environment(env)
    gpu.set_tuning(kernel=1)
    ref = gpu.render_hdr(scene, opts).view(np.uint32).copy()
    gpu.set_tuning(kernel=0)
    got = gpu.render_hdr(scene, opts).view(np.uint32)
    assert gpu.last_kernel().startswith("trace_persistent"), gpu.last_kernel()
    assert np.array_equal(got, ref)


def test_punctual_lights_forms_bitwise(gpu):
    """With punctual lights the persistent forms trace a hit's shadow rays
    before claiming new work; images stay bitwise those of the plain form."""
    scene = ptscenes.materials_scene()
    opts = oxpt.RenderOptions(width=64, height=48, samples_per_pixel=6, max_bounces=8)
    gpu.set_environment(None)
    gpu.set_tuning(kernel=1)
    ref = gpu.render_hdr(scene, opts).view(np.uint32).copy()
    for kernel in (0, 3, 7):
        gpu.set_tuning(kernel=kernel)
        got = gpu.render_hdr(scene, opts).view(np.uint32)
        assert np.array_equal(got, ref), f"kernel {kernel}"
    gpu.set_tuning(kernel=0)


def test_punctual_queue_forms_bitwise(gpu):
    """With 1..31 punctual lights the default forms evaluate every NEE job at
    the hit, keep the pending contributions (and the area job's ray) in a
    per-lane queue in device memory and make the BSDF continuation at once:
    identical accumulators to the general forms — layered materials, 16
    punctual lights + environment + emissive NEE, single refine and batched
    (jobs with 16, 3 and no punctual lights in one launch). With more than
    31 punctual lights the general form runs."""
    scene = oxpt.Scene.layered(seed=5, n_tris=30000, n_textures=2, tex_size=64, n_lights=16)
    env = oxpt.EnvironmentMap.sky(64, 32, 3)
    opts = oxpt.RenderOptions(width=80, height=56, samples_per_pixel=5, max_bounces=10)
    gpu.set_environment(env)
    gpu.set_tuning(kernel=0, early=oxpt.EARLY_OFF)
    ref = gpu.render_hdr(scene, opts).view(np.uint32).copy()
    assert "(early" not in gpu.last_kernel()
    for kernel in (0, 2, 3, 4, 5):
        gpu.set_tuning(kernel=kernel, early=0)
        got = gpu.render_hdr(scene, opts).view(np.uint32)
        assert gpu.last_kernel().endswith("(early, queue)"), gpu.last_kernel()
        assert np.array_equal(got, ref), f"kernel {kernel}"
    gpu.set_tuning(kernel=0)
    many = oxpt.Scene.layered(seed=6, n_tris=8000, n_textures=1, tex_size=32, n_lights=40)
    gpu.render_hdr(many, opts)
    assert "(early" not in gpu.last_kernel(), gpu.last_kernel()
    gpu.set_environment(None)
    dev = oxpt.Device(0)
    packs = [oxpt.Packed(scene, opts, env),
             oxpt.Packed(oxpt.Scene.layered(seed=7, n_tris=12000, n_textures=1, tex_size=32, n_lights=3), opts),
             oxpt.Packed(small_terrain(), opts)]
    ds = [dev.upload(pk) for pk in packs]
    out = []
    # General forms, queue forms, and the queue forms replayed from a
    # captured graph (twice: the second replay reuses the queue).
    for early, graph in ((oxpt.EARLY_OFF, False), (0, False), (0, True), (0, True)):
        dev.set_tuning(early=early)
        dev.set_graph(graph)
        acc = [dev.accum(opts.width, opts.height) for _ in ds]
        dev.refine_batch([(s, a, s.params, 0, 5) for s, a in zip(ds, acc)])
        dev.finish()
        out.append([a.read().view(np.uint32).copy() for a in acc])
    dev.set_graph(False)
    dev.set_tuning(early=0)
    for run in out[1:]:
        for a, b in zip(run, out[0]):
            assert np.array_equal(a, b)


def test_punctual_queue_limit(gpu):
    """31 punctual lights fill the NEE queue's job mask (bit 31 is the area /
    environment job): still the queue form, bitwise the general form; 32
    lights run the general form."""
    env = oxpt.EnvironmentMap.sky(64, 32, 3)
    opts = oxpt.RenderOptions(width=64, height=48, samples_per_pixel=4, max_bounces=6)
    gpu.set_environment(env)
    for n_lights, queue in ((31, True), (32, False)):
        scene = oxpt.Scene.layered(seed=11, n_tris=8000, n_textures=1, tex_size=32, n_lights=n_lights)
        gpu.set_tuning(kernel=2, early=oxpt.EARLY_OFF)
        ref = gpu.render_hdr(scene, opts).view(np.uint32).copy()
        gpu.set_tuning(kernel=2, early=0)
        got = gpu.render_hdr(scene, opts).view(np.uint32)
        assert gpu.last_kernel().endswith("(early, queue)") == queue, (n_lights, gpu.last_kernel())
        assert np.array_equal(got, ref), n_lights
    gpu.set_tuning(kernel=0)
    gpu.set_environment(None)


def test_kernel_forms_and_stack_split_bitwise(gpu):
    """Every occupancy form and every LDS/private split of the traversal stack
    (including a 1-entry LDS share, so deep pushes land in private memory)
    walks the same nodes in the same order: identical accumulators."""
    scene = small_terrain()
    opts = oxpt.RenderOptions(width=96, height=64, samples_per_pixel=4, max_bounces=6)
    gpu.set_environment(None)
    gpu.set_tuning(kernel=8, stack_lds=oxpt.STACK_PRIVATE)
    ref = gpu.render_hdr(scene, opts).view(np.uint32).copy()
    for kernel, lds in ((0, 0), (0, 1), (0, 5), (2, 24), (3, oxpt.STACK_PRIVATE), (4, 0), (5, 3), (6, 0), (8, 16)):
        gpu.set_tuning(kernel=kernel, stack_lds=lds)
        got = gpu.render_hdr(scene, opts).view(np.uint32)
        assert np.array_equal(got, ref), f"kernel {kernel} stack_lds {lds}"
    gpu.set_tuning(kernel=0)


def test_tail_shares_bitwise(gpu):
    """The persistent kernel's path-granular tail (the last tail/16 of the
    frame's tiles traced one path per claim, reduced in sample order) makes
    the same additions as whole-pixel runs: identical accumulators for every
    tail share, on frames with ragged edge tiles, across launches split by
    max_paths, and with punctual lights (shadow rays traced before claims)."""
    gpu.set_environment(None)
    for scene, (w, h), spp in ((small_terrain(), (100, 70), 6), (ptscenes.materials_scene(), (72, 40), 5)):
        opts = oxpt.RenderOptions(width=w, height=h, samples_per_pixel=spp, max_bounces=8)
        gpu.set_tuning(kernel=3, tail=oxpt.TAIL_NONE, max_paths=1 << 30)
        ref = gpu.render_hdr(scene, opts).view(np.uint32).copy()
        # max_paths 0 keeps the previous cap: the split launches come last.
        for tail, max_paths in ((1, 0), (8, 0), (16, 0), (0, 0), (8, w * h * 2), (16, w * h * 3)):
            gpu.set_tuning(kernel=3, tail=tail, max_paths=max_paths)
            got = gpu.render_hdr(scene, opts).view(np.uint32)
            assert np.array_equal(got, ref), f"tail {tail} max_paths {max_paths}"
    gpu.set_tuning(kernel=0, max_paths=1 << 30)


def test_tuning_rejects_bad_stack_split(gpu):
    with pytest.raises(oxpt.Error):
        gpu.set_tuning(tail=17)
    with pytest.raises(oxpt.Error):
        gpu.set_tuning(stack_lds=65)
    with pytest.raises(oxpt.Error):
        gpu.set_tuning(kernel=9)
    with pytest.raises(oxpt.Error):
        gpu.set_tuning(bvh=3)
    gpu.set_tuning(kernel=0)


# ---------------------------------------------------------------------------
# The reference's parity scenes (tests/pathtrace.rs:319-418): every lobe,
# punctual lights, alpha modes, textures with mips, direct-only.
# ---------------------------------------------------------------------------



@pytest.mark.parametrize("name,bounces,spp", [("materials", 8, 32), ("alpha", 8, 32), ("checker", 4, 16),
                                              ("spheres", 8, 16), ("materials-direct", 1, 32)])
def test_reference_scene_parity(gpu, name, bounces, spp):
    scene = {"materials": ptscenes.materials_scene, "materials-direct": ptscenes.materials_scene,
             "alpha": ptscenes.alpha_planes, "checker": ptscenes.checker_floor,
             "spheres": ptscenes.sphere_grid}[name]()
    opts = oxpt.RenderOptions(width=128, height=96, samples_per_pixel=spp, max_bounces=bounces)
    gpu.set_environment(None)
    gh = gpu.render_hdr(scene, opts)
    gi = gpu.image()
    _, ci, ch = oracle_frame(scene, opts)
    p = parity(name, gh, gi, ch, ci, 60.0)
    print(f"{name}: PSNR {p:.2f} dB")


def test_layered_texture_scene_parity(gpu):
    """Layered BSDF + textures + 16 punctual lights (the large layered workload, scaled down)."""
    scene = oxpt.Scene.layered(seed=2, n_tris=40000, n_textures=4, tex_size=256, n_lights=16)
    opts = oxpt.RenderOptions(width=128, height=72, samples_per_pixel=8, max_bounces=16)
    gpu.set_environment(None)
    gh = gpu.render_hdr(scene, opts)
    gi = gpu.image()
    _, ci, ch = oracle_frame(scene, opts)
    p = parity("layered", gh, gi, ch, ci, 60.0)
    print(f"layered: PSNR {p:.2f} dB")


def test_autotune_keeps_the_image(gpu):
    """The autotuner picks one of the kernel forms; the image does not change."""
    scene = small_terrain()
    opts = oxpt.RenderOptions(width=96, height=64, samples_per_pixel=4, max_bounces=6)
    gpu.set_environment(None)
    gpu.set_tuning(kernel=0)
    before = gpu.render_hdr(scene, opts).view(np.uint32).copy()
    t, ms = gpu.autotune(2)
    assert ms > 0 and t.kernel in (0, 2, 3, 4, 5, 7)
    gpu.reset()
    gpu.refine(opts.samples_per_pixel)
    after = gpu.hdr().view(np.uint32)
    gpu.set_tuning(kernel=0)
    assert np.array_equal(before, after), f"tuning {t.kernel}/{t.stack_lds}"


def test_early_continuation_forms_bitwise(gpu):
    """Without punctual lights the default megakernel forms make the NEE
    shadow ray and the BSDF continuation together at a hit (the BSDF is not
    kept across traversals): the same random numbers and sums as the general
    forms, so identical accumulators — layered materials (volumes,
    transmission, clearcoat, sheen, blending), environment + emissive NEE,
    single refine and batched."""
    scene = oxpt.Scene.layered(seed=4, n_tris=30000, n_textures=2, tex_size=64, n_lights=0)
    env = oxpt.EnvironmentMap.sky(64, 32, 3)
    opts = oxpt.RenderOptions(width=80, height=56, samples_per_pixel=5, max_bounces=10)
    gpu.set_environment(env)
    gpu.set_tuning(kernel=0, early=oxpt.EARLY_OFF)
    ref = gpu.render_hdr(scene, opts).view(np.uint32).copy()
    assert "(early)" not in gpu.last_kernel()
    for kernel in (0, 2, 4):
        gpu.set_tuning(kernel=kernel, early=0)
        got = gpu.render_hdr(scene, opts).view(np.uint32)
        assert gpu.last_kernel().endswith("(early)"), gpu.last_kernel()
        assert np.array_equal(got, ref), f"kernel {kernel}"
    gpu.set_tuning(kernel=0)
    gpu.set_environment(None)
    # Batched: the same two scenes in one launch, early and general.
    dev = oxpt.Device(0)
    pk = oxpt.Packed(scene, opts, env)
    pk2 = oxpt.Packed(small_terrain(), opts)
    ds = [dev.upload(pk), dev.upload(pk2)]
    out = []
    for early in (0, oxpt.EARLY_OFF):
        dev.set_tuning(early=early)
        acc = [dev.accum(opts.width, opts.height) for _ in ds]
        dev.refine_batch([(s, a, s.params, 0, 5) for s, a in zip(ds, acc)])
        dev.finish()
        out.append([a.read().view(np.uint32).copy() for a in acc])
    for a, b in zip(out[0], out[1]):
        assert np.array_equal(a, b)


def test_opaque_walk_bitwise(gpu):
    """Scenes whose materials all have alpha mode OPAQUE run the traversal
    built without the alpha test: identical accumulators to the general
    traversal (OPAQUE_OFF) for the persistent forms, the small-frame batch,
    the NEE-queue form (opaque terrain + punctual lights) and a batch; a
    scene with one MASK or BLEND material keeps the general traversal."""
    env = oxpt.EnvironmentMap.sky(64, 32, 3)
    gpu.set_environment(env)
    opts = oxpt.RenderOptions(width=96, height=64, samples_per_pixel=4, max_bounces=8)
    lit = small_terrain()
    lit.add_light(kind="point", position=(0.0, 6.0, 0.0), color=(1.0, 0.9, 0.8), intensity=40.0)
    lit.add_light(kind="directional", direction=(0.3, -1.0, 0.2), color=(0.9, 0.9, 1.0), intensity=2.0)
    for scene, tag in ((small_terrain(), "(early"), (lit, "(early, queue")):
        for kernel in (0, 2, 3, 4):
            gpu.set_tuning(kernel=kernel, opaque=oxpt.OPAQUE_OFF)
            ref = gpu.render_hdr(scene, opts).view(np.uint32).copy()
            assert tag in gpu.last_kernel() and "opaque" not in gpu.last_kernel(), gpu.last_kernel()
            gpu.set_tuning(kernel=kernel)
            got = gpu.render_hdr(scene, opts).view(np.uint32)
            assert tag in gpu.last_kernel() and gpu.last_kernel().endswith("opaque)"), gpu.last_kernel()
            assert np.array_equal(got, ref), (tag, kernel)
    gpu.set_tuning(kernel=0)
    gpu.render_hdr(ptscenes.alpha_planes(), opts)
    assert "opaque" not in gpu.last_kernel(), gpu.last_kernel()
    gpu.set_environment(None)
    dev = oxpt.Device(0)
    ds = [dev.upload(oxpt.Packed(small_terrain(), opts, env)), dev.upload(oxpt.Packed(lit, opts, env)),
          dev.upload(oxpt.Packed(oxpt.Scene.room(1), opts))]
    out = []
    for opaque in (oxpt.OPAQUE_OFF, 0):
        dev.set_tuning(opaque=opaque)
        acc = [dev.accum(opts.width, opts.height) for _ in ds]
        dev.refine_batch([(s, a, s.params, 0, 4) for s, a in zip(ds, acc)])
        dev.finish()
        assert ("opaque" in dev.last_kernel()) == (opaque == 0), dev.last_kernel()
        out.append([a.read().view(np.uint32).copy() for a in acc])
    dev.set_tuning()
    for a, b in zip(out[0], out[1]):
        assert np.array_equal(a, b)


def _back_face_scene():
    """A lit back wall facing the camera and, in front of it, a single-sided
    red quad turned away from the camera: camera rays pass through the quad
    (acceptance culls back faces of camera rays on single-sided materials,
    wgsl:625), while bounce and shadow rays hit it (its shadow and its red
    light on the wall)."""
    s = oxpt.Scene()
    wall = s.add_material(ptscenes.lambert((0.8, 0.8, 0.8)))
    pos, uv = ptscenes.quad([-3, -2, -2], [3, -2, -2], [3, 2, -2], [-3, 2, -2])  # normal +z
    s.add_triangles(pos, wall, uv0=uv)
    red = s.add_material(ptscenes.lambert((0.9, 0.1, 0.1), double_sided=0))
    pos, uv = ptscenes.quad([-1, -1, 0], [-1, 1, 0], [1, 1, 0], [1, -1, 0])  # normal -z
    s.add_triangles(pos, red, uv0=uv)
    light = s.add_material(ptscenes.lambert((0, 0, 0), emissive=(8.0, 8.0, 8.0)))
    pos, uv = ptscenes.quad([-0.5, 2.5, -0.5], [0.5, 2.5, -0.5], [0.5, 2.5, 0.5], [-0.5, 2.5, 0.5])
    s.add_triangles(pos, light)  # normal -y: facing down
    s.set_camera((0.3, 0.2, 3.0), (0.0, 0.0, -2.0), yfov=0.9)
    return s


def test_camera_rays_cull_single_sided_back_faces(gpu):
    """The batched form tests a candidate's front face only for camera rays
    and sets it once for the hit a walk returns; the single-frame forms test
    it per candidate. Both equal the plain kernel bit for bit on a scene
    where camera rays meet a single-sided back face and bounce rays hit it,
    and match the oracle."""
    scene = _back_face_scene()
    opts = oxpt.RenderOptions(width=96, height=64, samples_per_pixel=8, max_bounces=6)
    gpu.set_environment(None)
    gpu.set_tuning(kernel=1)
    ref = gpu.render_hdr(scene, opts).view(np.uint32).copy()
    assert gpu.last_kernel() == "trace_tiles", gpu.last_kernel()
    for kernel, form in ((0, "trace_batch"), (4, "trace_persistent"), (5, "trace_persistent")):
        gpu.set_tuning(kernel=kernel)
        got = gpu.render_hdr(scene, opts).view(np.uint32)
        assert gpu.last_kernel().startswith(form), (kernel, gpu.last_kernel())
        assert np.array_equal(got, ref), f"kernel {kernel}"
    gpu.set_tuning(kernel=0)
    gh = gpu.render_hdr(scene, opts)
    gi = gpu.image()
    # The quad is seen through: the centre of the frame is the lit wall, not red.
    h, w = gh.shape[:2]
    c = gh[h // 2 - 4:h // 2 + 4, w // 2 - 4:w // 2 + 4, :3].mean(axis=(0, 1))
    assert c[0] < 3.0 * c[1] + 0.05, c
    _, ci, ch = oracle_frame(scene, opts)
    parity("back faces", gh, gi, ch, ci, 60.0)
