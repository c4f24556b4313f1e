"""Pins of the CPU oracle and the host packer (no GPU needed).

Each check is chosen so that a plausible mistake fails it: the reference's
own known-answer tests where they exist (f16 KAT, src/scene.rs:496-511;
white furnace, tests/pathtrace.rs:295-317), otherwise closed forms,
invariants and brute force.
"""
import math

import numpy as np
import pytest

import oracle
import oxpt


# ---------------------------------------------------------------------------
# f32 -> f16 (src/scene.rs:244-272, KAT :496-511).
# ---------------------------------------------------------------------------

KAT = [(0.0, 0x0000), (-0.0, 0x8000), (1.0, 0x3C00), (0.5, 0x3800), (-2.0, 0xC000), (65504.0, 0x7BFF),
       (1.0e6, 0x7C00), (float("inf"), 0x7C00), (5.9604645e-8, 0x0001), (1.0 / 3.0, 0x3555)]


@pytest.mark.parametrize("conv", [oracle.f32_to_f16, oxpt.f32_to_f16], ids=["oracle", "product"])
def test_f16_reference_kat(conv):
    for v, want in KAT:
        assert conv(v) == want, (v, hex(conv(v)), hex(want))
    assert conv(float("nan")) & 0x7E00 == 0x7E00


@pytest.mark.parametrize("conv", [oracle.f32_to_f16, oxpt.f32_to_f16], ids=["oracle", "product"])
def test_f16_matches_numpy_rne(conv):
    rng = np.random.default_rng(0)
    bits = np.concatenate([rng.integers(0, 1 << 32, 20000, dtype=np.uint64).astype(np.uint32),
                           np.arange(0x33000000, 0x33800000, 997, dtype=np.uint32),   # subnormal boundary
                           np.arange(0x477FE000, 0x47800000, 17, dtype=np.uint32)])    # overflow boundary
    vals = bits.view(np.float32)
    finite = np.isfinite(vals)
    want = vals[finite].astype(np.float16).view(np.uint16)
    got = np.array([conv(float(v)) for v in vals[finite]], dtype=np.uint16)
    assert np.array_equal(got, want)


# ---------------------------------------------------------------------------
# Sobol' matrices and the Owen-scrambled sampler (pathtrace.wgsl:158-228).
# ---------------------------------------------------------------------------

def test_sobol_directions_oracle_equals_product():
    assert np.array_equal(oracle.sobol_directions(), oxpt.sobol_directions())


def test_sobol_direction_closed_forms():
    v = oracle.sobol_directions().reshape(4, 32)
    assert all(int(v[0, k]) == 1 << (31 - k) for k in range(32))
    # x + 1: v_k = v_{k-1} ^ (v_{k-1} >> 1) from 0x80000000 — Pascal's triangle mod 2.
    for k in range(32):
        word = 0
        for j in range(k + 1):
            if math.comb(k, j) % 2:
                word |= 1 << (31 - j)
        assert int(v[1, k]) == word
    assert [hex(int(x)) for x in v[2, :4]] == ["0x80000000", "0xc0000000", "0x60000000", "0x90000000"]
    assert [hex(int(x)) for x in v[3, :3]] == ["0x80000000", "0xc0000000", "0x20000000"]


def _is_net(xs, ys, m):
    """(0, m, 2)-net: every elementary box of area 2^-m holds exactly one point."""
    for a in range(m + 1):
        bx = (xs >> np.uint64(32 - a)) if a else np.zeros_like(xs)
        by = (ys >> np.uint64(32 - (m - a))) if m - a else np.zeros_like(ys)
        cells = bx * np.uint64(1 << (m - a)) + by
        if len(np.unique(cells)) != len(xs):
            return False
    return True


@pytest.mark.parametrize("seed,x,y,p", [(0, 0, 0, 0), (7, 13, 200, 1), (0xDEADBEEF, 1919, 1079, 5)])
def test_scrambled_sobol_is_a_0m2_net(seed, x, y, p):
    """The first 2^m samples of a pixel: dims (0,1) and (2,3) form (0,m,2)-nets."""
    sob = oracle.sobol_directions()
    for m in range(1, 11):
        n = 1 << m
        w = oracle.pattern_words(sob, seed, np.full(n, x), np.full(n, y), np.arange(n), np.full(n, p))
        w = w.astype(np.uint64)
        assert _is_net(w[:, 0], w[:, 1], m), (m, "dims 0,1")
        for d in (2, 3):  # each remaining dimension alone is stratified
            assert len(np.unique(w[:, d] >> np.uint64(32 - m))) == n, (m, d)


def test_owen_scramble_is_a_bijection_preserving_prefixes():
    """Nested uniform scrambling: bit k depends only on bits above it (pathtrace.wgsl:185-196)."""
    rng = np.random.default_rng(1)
    for seed in rng.integers(0, 1 << 32, 8, dtype=np.uint64):
        xs = rng.integers(0, 1 << 32, 64, dtype=np.uint64)
        for x in xs:
            x = int(x)
            y = oracle.lib.oxo_nested_uniform_scramble(x, int(seed))
            for k in (4, 12, 20, 28):
                x2 = (x & ~((1 << (32 - k)) - 1)) | (int(rng.integers(0, 1 << (32 - k))))
                y2 = oracle.lib.oxo_nested_uniform_scramble(x2, int(seed))
                assert (y >> (32 - k)) == (y2 >> (32 - k))
    seen = {oracle.lib.oxo_nested_uniform_scramble(i, 12345) & 0xFFFF for i in range(1 << 16)}
    assert len(seen) == 1 << 16


def test_pcg_hash_is_injective_on_a_block():
    vals = {oracle.lib.oxo_pcg_hash(i) for i in range(1 << 18)}
    assert len(vals) == 1 << 18


# ---------------------------------------------------------------------------
# Sheen LUT (pathtrace.wgsl:997-1013): product table vs an independent quadrature.
# ---------------------------------------------------------------------------

def test_sheen_lut_matches_independent_quadrature():
    a = oxpt.sheen_table()
    b = oracle.sheen_table(192)
    assert np.all(np.isfinite(a)) and np.all(a >= 0)
    # The smooth end (roughness 0: alpha 1e-3, a very peaked lobe) is the
    # hardest for either quadrature: 1 % relative there, 2e-3 elsewhere.
    tol = 2e-3 + 1e-2 * np.abs(b)
    assert np.all(np.abs(a - b) <= tol), np.max(np.abs(a - b) / tol)


# ---------------------------------------------------------------------------
# Estimator: the reference's white furnace (tests/pathtrace.rs:295-317).
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("strategy", ["mis", "light", "bsdf"])
@pytest.mark.parametrize("flags", [0, oracle.F64_LIGHTS], ids=["f32", "f64"])
def test_white_furnace_oracle(strategy, flags):
    scene = oxpt.Scene.furnace(0.5, 1.0)
    o = oxpt.RenderOptions(width=8, height=8, samples_per_pixel=64, max_bounces=64, strategy=strategy, ambient=0.0)
    osc = oracle.Scene.from_packed(oxpt.Packed(scene, o))
    acc = osc.render(0, 64, flags=flags)
    m = acc[..., :3].sum(axis=(0, 1)) / acc[..., 3].sum()
    assert np.all(np.abs(m - 2.0) < 0.04), m


# ---------------------------------------------------------------------------
# BVH + traversal vs brute force (pathtrace.wgsl:651-735).
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("which", ["room", "terrain"])
def test_bvh_closest_hit_equals_brute_force(which):
    scene = oxpt.Scene.room(0) if which == "room" else oxpt.Scene.terrain(seed=2, gx=48, gz=32, n_spheres=10,
                                                                           subdiv=2, n_emissive=8)
    o = oxpt.RenderOptions(width=8, height=8)
    osc = oracle.Scene.from_packed(oxpt.Packed(scene, o))
    rng = np.random.default_rng(5)
    hits = 0
    for _ in range(1500):
        if which == "room":  # inside the room: every ray hits a wall
            org = rng.uniform([-0.9, 0.1, -0.9], [0.9, 1.9, 0.9])
        else:
            org = rng.uniform(-3, 3, 3) + np.array([0, 1.5, 0])
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        a = osc.closest_hit(org, d)
        b = osc.closest_hit(org, d, brute=True)
        assert a[0] == b[0] or a[1] == b[1], (org, d, a, b)
        assert a[1] == b[1]
        hits += a[0] != 0xFFFFFFFF
    assert hits > 300


def test_bvh_leaf_slots_cover_every_triangle_once():
    scene = oxpt.Scene.terrain(seed=2, gx=48, gz=32, n_spheres=10, subdiv=2, n_emissive=8)
    pk = oxpt.Packed(scene, oxpt.RenderOptions(width=8, height=8))
    n_nodes = pk.params.env[3]
    nodes = pk.geom[: 8 * n_nodes].reshape(-1, 8)
    first = nodes[:, 3].view(np.uint32)
    count = nodes[:, 7].view(np.uint32)
    leaves = count > 0
    covered = np.zeros(scene.triangle_count(), dtype=np.int32)
    for f, c in zip(first[leaves], count[leaves]):
        covered[f: f + c] += 1
    assert (covered == 1).all()
    # Walking from the root: children boxes lie inside their parent's box,
    # child pairs start at even indices (64-byte aligned), every leaf is reached.
    stack, reached = [0], 0
    while stack:
        i = stack.pop()
        if count[i] > 0:
            reached += 1
            continue
        assert first[i] % 2 == 0
        for ch in (first[i], first[i] + 1):
            assert np.all(nodes[ch, :3] >= nodes[i, :3]) and np.all(nodes[ch, 4:7] <= nodes[i, 4:7])
            stack.append(ch)
    assert reached == int(leaves.sum())
    # Global ids in the triangle words are a permutation.
    tris = pk.geom[8 * n_nodes: 8 * n_nodes + 12 * scene.triangle_count()].reshape(-1, 12)
    gids = tris[:, 3].view(np.uint32)
    assert np.array_equal(np.sort(gids), np.arange(scene.triangle_count(), dtype=np.uint32))
    assert pk.bvh_depth < 64
    # Stable partitions: every leaf lists its triangles in ascending id order
    # (the canonical tree the device builder reproduces), and every node's
    # box is exactly the union of its triangles' boxes.
    tri_lo = np.minimum(np.minimum(tris[:, 0:3], tris[:, 4:7]), tris[:, 8:11]) + np.float32(0.0)
    tri_hi = np.maximum(np.maximum(tris[:, 0:3], tris[:, 4:7]), tris[:, 8:11]) + np.float32(0.0)

    def span(i):
        if count[i] > 0:
            return first[i], first[i] + count[i]
        return span(first[i])[0], span(first[i] + 1)[1]

    for i in range(n_nodes):
        if i == 1:
            continue
        b, e = span(i)
        if count[i] > 0:
            assert np.all(np.diff(gids[b:e].astype(np.int64)) > 0), f"leaf {i} ids not ascending"
        assert np.array_equal(nodes[i, :3], tri_lo[b:e].min(axis=0)), f"node {i} lo"
        assert np.array_equal(nodes[i, 4:7], tri_hi[b:e].max(axis=0)), f"node {i} hi"


@pytest.mark.parametrize("gen", ["terrain", "room", "coincident"])
def test_wide_bvh_covers_every_slot_with_conservative_boxes(gen):
    """The 4-wide nodes the kernels walk: every leaf slot reached exactly once,
    every child box holds every triangle below it, child boxes taken from the
    binary tree unchanged, the stack bound 3 * depth <= 96 holds."""
    if gen == "terrain":
        scene = oxpt.Scene.terrain(seed=2, gx=48, gz=32, n_spheres=10, subdiv=2, n_emissive=8)
    elif gen == "room":
        scene = oxpt.Scene.room(0)
    else:
        # 200 coincident triangles: SAH finds no split, the object-median
        # fallback still bounds every leaf.
        scene = oxpt.Scene()
        m = scene.add_material(oxpt.material(base_color=(0.5, 0.5, 0.5, 1.0)))
        scene.add_triangles(np.tile(np.array([[0, 0, 0, 1, 0, 0, 0, 1, 0]], np.float32), (200, 1)), m)
    pk = oxpt.Packed(scene, oxpt.RenderOptions(width=8, height=8))
    P = pk.params
    ntri = scene.triangle_count()
    base = P.offs2[3]
    assert base > 0 and base % 8 == 0
    tri = pk.geom[8 * P.env[3]: 8 * P.env[3] + 12 * ntri].reshape(-1, 12)
    pos = tri.reshape(-1, 3, 4)[:, :, :3]
    wide = pk.geom[4 * base:].reshape(-1, 32)
    bin_boxes = {tuple(np.concatenate([n[:3], n[4:7]]).view(np.uint32)) for n in pk.geom[: 8 * P.env[3]].reshape(-1, 8)}
    covered = np.zeros(ntri, np.int32)
    stack, max_depth = [(0, 1)], 0
    while stack:
        node, depth = stack.pop()
        max_depth = max(max_depth, depth)
        w = wide[node]
        refs, cnts = w[24:28].view(np.uint32), w[28:32].view(np.uint32)
        for i in range(4):
            if cnts[i] == 0xFFFFFFFF:
                continue
            lo = np.array([w[i], w[4 + i], w[8 + i]], np.float32)
            hi = np.array([w[12 + i], w[16 + i], w[20 + i]], np.float32)
            if cnts[i] == 0:
                cw = wide[refs[i]]
                for j in range(4):
                    if cw[28:32].view(np.uint32)[j] != 0xFFFFFFFF:
                        clo = np.array([cw[j], cw[4 + j], cw[8 + j]])
                        chi = np.array([cw[12 + j], cw[16 + j], cw[20 + j]])
                        assert np.all(clo >= lo) and np.all(chi <= hi)
                stack.append((refs[i], depth + 1))
            else:
                assert 1 <= cnts[i] <= 30
                sl = slice(int(refs[i]), int(refs[i] + cnts[i]))
                covered[sl] += 1
                assert np.all(pos[sl] >= lo) and np.all(pos[sl] <= hi)
            assert tuple(np.concatenate([lo, hi]).view(np.uint32)) in bin_boxes
    assert (covered == 1).all()
    assert 3 * max_depth <= 96


# ---------------------------------------------------------------------------
# Light sampling (pathtrace.wgsl:1486-1654).
# ---------------------------------------------------------------------------

def _rand_dirs(rng, n):
    d = rng.normal(size=(n, 3))
    return d / np.linalg.norm(d, axis=1, keepdims=True)


@pytest.mark.parametrize("f64", [False, True])
def test_solid_angle_matches_monte_carlo(f64):
    rng = np.random.default_rng(2)
    p = np.zeros(3)
    a, b, c = np.array([1.0, -0.3, 2.0]), np.array([-0.5, 0.8, 1.7]), np.array([0.4, 0.9, 2.6])
    om = oracle.solid_angle(p, a, b, c, f64)
    # Fraction of uniform directions that hit the triangle.
    d = _rand_dirs(rng, 400000)
    n = np.cross(b - a, c - a)
    t = (a @ n) / (d @ n)
    q = d * t[:, None]
    e0, e1, e2 = np.cross(b - a, q - a), np.cross(c - b, q - b), np.cross(a - c, q - c)
    inside = (t > 0) & ((e0 @ n) >= 0) & ((e1 @ n) >= 0) & ((e2 @ n) >= 0)
    mc = 4 * np.pi * inside.mean()
    assert abs(om - mc) < 4 * 4 * np.pi * np.sqrt(inside.mean() / len(d)) + 1e-4, (om, mc)


@pytest.mark.parametrize("f64", [False, True])
def test_arvo_samples_are_uniform_over_the_spherical_triangle(f64):
    rng = np.random.default_rng(4)
    p = np.zeros(3)
    a, b, c = np.array([1.0, -0.3, 2.0]), np.array([-0.5, 0.8, 1.7]), np.array([0.4, 0.9, 2.6])
    n = np.cross(b - a, c - a)
    s = np.array([oracle.sample_spherical_triangle(p, a, b, c, u0, u1, f64) for u0, u1 in rng.random((20000, 2))])
    # Every sample points into the triangle.
    t = (a @ n) / (s @ n)
    q = s * t[:, None]
    e = [np.cross(b - a, q - a) @ n, np.cross(c - b, q - b) @ n, np.cross(a - c, q - c) @ n]
    assert np.all(t > 0) and all(np.all(x >= -1e-4) for x in e)
    # Uniform in solid angle: mean direction equals that of rejection samples.
    d = _rand_dirs(rng, 3000000)
    tt = (a @ n) / (d @ n)
    qq = d * tt[:, None]
    ins = (tt > 0) & ((np.cross(b - a, qq - a) @ n) >= 0) & ((np.cross(c - b, qq - b) @ n) >= 0) & (
        (np.cross(a - c, qq - c) @ n) >= 0)
    ref = d[ins].mean(axis=0)
    assert np.allclose(s.mean(axis=0), ref, atol=3e-3), (s.mean(axis=0), ref)


def test_emissive_cdf_invariants():
    scene = oxpt.Scene.terrain(seed=2, gx=48, gz=32, n_spheres=10, subdiv=2, n_emissive=64)
    pk = oxpt.Packed(scene, oxpt.RenderOptions(width=8, height=8))
    P = pk.params
    n = P.cfg[3]
    assert n == 64
    rec = pk.tables[P.offs1[0]: P.offs1[0] + 4 * n].reshape(n, 4)
    cdf, pmf = rec[:, 0].view(np.float32), rec[:, 1].view(np.float32)
    assert np.all(np.diff(cdf) >= 0) and cdf[-1] == 1.0
    assert np.allclose(np.diff(np.concatenate([[0.0], cdf])), pmf, atol=1e-7)
    # The slot of each record holds its global triangle id.
    tris = pk.geom[8 * P.env[3]: 8 * P.env[3] + 12 * scene.triangle_count()].reshape(-1, 12)
    for g, slot in rec[:, 2:4]:
        assert tris[slot, 3].view(np.uint32) == g
        assert tris[slot, 7].view(np.uint32) < n


def test_environment_tables_integrate_to_one():
    scene = oxpt.Scene.room(0)
    # A smooth map (no sun) so a uniform-direction Monte Carlo integral converges.
    yy, xx = np.mgrid[0:64, 0:128]
    rgb = np.stack([1 + np.sin(xx / 9.0), 1.5 + np.cos(yy / 7.0), 0.2 + (xx * yy) / 8192.0], axis=2)
    env = oxpt.EnvironmentMap(rgb.astype(np.float32), 1.0)
    pk = oxpt.Packed(scene, oxpt.RenderOptions(width=8, height=8), env)
    P = pk.params
    w, h = P.env[0], P.env[1]
    t = pk.tables
    marg = t[P.offs1[2]: P.offs1[2] + h + 1].view(np.float32)
    cond = t[P.offs1[3]: P.offs1[3] + h * (w + 1)].view(np.float32).reshape(h, w + 1)
    mass = t[P.offs2[0]: P.offs2[0] + w * h].view(np.float32).reshape(h, w)
    assert marg[0] == 0 and marg[-1] == 1 and np.all(np.diff(marg) >= 0)
    assert np.all(cond[:, 0] == 0) and np.all(cond[:, -1] == 1) and np.all(np.diff(cond, axis=1) >= 0)
    assert abs(mass.sum() - 1.0) < 1e-5
    # pdf(d) = mass w h / (2 pi^2 sin theta) integrates to 1 over the sphere.
    rng = np.random.default_rng(8)
    d = _rand_dirs(rng, 2000000)
    u = 0.5 + np.arctan2(d[:, 0], -d[:, 2]) / (2 * np.pi)
    th = np.arccos(np.clip(d[:, 1], -1, 1))
    x = np.minimum((u * w).astype(np.int64), w - 1)
    y = np.minimum((th / np.pi * h).astype(np.int64), h - 1)
    pdf = mass[y, x] * w * h / (2 * np.pi ** 2 * np.maximum(np.sin(th), 1e-6))
    assert abs(pdf.mean() * 4 * np.pi - 1.0) < 0.02


# ---------------------------------------------------------------------------
# BSDF (pathtrace.wgsl:1087-1380): sample pdf == eval pdf; energy.
# ---------------------------------------------------------------------------

def _mat(**kw):
    pk = oxpt.Packed(oxpt.Scene.room(0), oxpt.RenderOptions(width=8, height=8))  # a table to get the LUT
    lut = pk.tables[pk.params.offs0[1]: pk.params.offs0[1] + 512].view(np.float32)
    s = oxpt.Scene()
    s.add_material(oxpt.material(**kw))
    s.add_triangles(np.array([[0, 0, 0, 1, 0, 0, 0, 1, 0]], dtype=np.float32), 0)
    pk2 = oxpt.Packed(s, oxpt.RenderOptions(width=8, height=8))
    off = pk2.params.offs0[2]
    return pk2.tables[off: off + 136].copy(), lut


MATERIALS = {
    "lambert": dict(metallic=0.0, roughness=1.0, has_specular=1, specular=0.0, base_color=(0.7, 0.5, 0.3, 1)),
    "plastic": dict(metallic=0.0, roughness=0.3, base_color=(0.2, 0.6, 0.3, 1)),
    "metal": dict(metallic=1.0, roughness=0.25, base_color=(0.9, 0.7, 0.4, 1)),
    "glass": dict(metallic=0.0, roughness=0.1, has_transmission=1, transmission=1.0, has_volume=1, thickness=1.0,
                  base_color=(0.95, 0.95, 1.0, 1)),
    "thin": dict(metallic=0.0, roughness=0.3, has_transmission=1, transmission=0.8, base_color=(0.9, 0.7, 0.5, 1)),
    "clearcoat": dict(metallic=0.0, roughness=0.6, has_clearcoat=1, clearcoat=1.0, clearcoat_roughness=0.15,
                      base_color=(0.6, 0.1, 0.1, 1)),
    "sheen": dict(metallic=0.0, roughness=0.9, has_sheen=1, sheen_color=(0.9, 0.8, 0.6), sheen_roughness=0.5,
                  base_color=(0.1, 0.1, 0.3, 1)),
}


@pytest.mark.parametrize("name", sorted(MATERIALS))
def test_bsdf_sample_pdf_matches_eval_pdf(name):
    mw, lut = _mat(**MATERIALS[name])
    rng = np.random.default_rng(9)
    n = np.array([0.0, 0.0, 1.0])
    ok = 0
    for _ in range(400):
        v = rng.normal(size=3)
        v[2] = abs(v[2]) + 0.05
        v /= np.linalg.norm(v)
        r = oracle.bsdf_probe(mw, lut, n, v, True, rng.random(3), v)
        if r[0] > 0:
            ok += 1
            assert r[7] > 0 and abs(r[7] - r[8]) <= 1e-4 * max(1.0, r[7]), r
    assert ok > 200


@pytest.mark.parametrize("name", ["lambert", "plastic", "metal", "clearcoat"])
def test_bsdf_albedo_at_most_one(name):
    """E[f |n.l| / pdf] over BSDF samples is the albedo: <= 1 (+ MC noise)."""
    mw, lut = _mat(**MATERIALS[name])
    rng = np.random.default_rng(10)
    n = np.array([0.0, 0.0, 1.0])
    v = np.array([0.3, 0.0, 1.0])
    v /= np.linalg.norm(v)
    tot = np.zeros(3)
    N = 6000
    for u in rng.random((N, 3)):
        r = oracle.bsdf_probe(mw, lut, n, v, True, u, v)
        if r[0] > 0:
            tot += r[4:7] / r[7]
    alb = tot / N
    assert np.all(alb <= 1.03), alb
    if name == "lambert":
        assert np.allclose(alb, [0.7, 0.5, 0.3], atol=0.02), alb


# ---------------------------------------------------------------------------
# Resolve (resolve.wgsl:38-75,151-188).
# ---------------------------------------------------------------------------

def test_resolve_background_and_tone_maps():
    acc = np.zeros((2, 3, 4), dtype=np.float32)
    acc[0, 1] = [0.5, 0.25, 2.0, 4.0]     # covered 4/4
    acc[1, 2] = [0.4, 0.4, 0.4, 2.0]      # covered 2/4 -> composite
    bg = (37, 99, 201, 180)
    for tm, f in ((0, lambda c: np.minimum(c, 1.0)),
                  (1, lambda c: np.minimum(c / (1 + (0.2126 * c[0] + 0.7152 * c[1] + 0.0722 * c[2])), 1.0)),
                  (2, lambda c: np.clip(((c * .6) * (2.51 * c * .6 + .03)) / ((c * .6) * (2.43 * c * .6 + .59) + .14),
                                        0, 1))):
        rp = oxpt.ResolveParams((1, 3, tm, 0), (1.0, 4.0, 3.0, 0.0), tuple(c / 255.0 for c in bg))
        img = oracle.resolve(acc, rp)
        assert tuple(img[0, 0]) == bg and tuple(img[1, 0]) == bg
        m = np.array([0.5, 0.25, 2.0]) / 4.0
        want = f(m)
        srgb = np.where(want <= 0.0031308, 12.92 * want, 1.055 * want ** (1 / 2.4) - 0.055)
        assert np.all(np.abs(img[0, 1, :3].astype(float) - np.floor(srgb * 255 + 0.5)) <= 1)
        assert img[0, 1, 3] == 255
        assert 180 < img[1, 2, 3] < 255
    rp = oxpt.ResolveParams((1, 3, 0, 1), (1.0, 4.0, 3.0, 0.0), tuple(c / 255.0 for c in bg))
    hdr = oracle.resolve(acc, rp)
    assert np.allclose(hdr[0, 1, :3], np.array([0.5, 0.25, 2.0]) / 4.0)


# ---------------------------------------------------------------------------
# The reference's parity scenes render finitely on the oracle (sanity of the
# scene builders the GPU parity tests use).
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("name", ["materials", "alpha", "checker", "spheres"])
def test_reference_scenes_render_on_the_oracle(name):
    import ptscenes

    scene = {"materials": ptscenes.materials_scene, "alpha": ptscenes.alpha_planes,
             "checker": ptscenes.checker_floor, "spheres": ptscenes.sphere_grid}[name]()
    opts = oxpt.RenderOptions(width=48, height=32, samples_per_pixel=4, max_bounces=6)
    osc = oracle.Scene.from_packed(oxpt.Packed(scene, opts))
    acc, cnt = osc.render(0, 4, counts=True)
    assert np.isfinite(acc).all()
    assert (acc[..., 3] > 0).mean() > 0.5
    m = acc[..., :3].sum() / max(acc[..., 3].sum(), 1)
    assert 0.005 < m < 50, m
    if name == "checker":
        assert cnt.tex_taps > 0
