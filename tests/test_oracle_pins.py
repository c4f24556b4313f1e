"""Pins of oracle and packer functions that invariants alone leave open.

Each test compares against something independent of the code under test: a
third-party implementation (scipy's Joe-Kuo Sobol' generator, scipy
quadrature), a published listing (PCG's RXS-M-XS constants, Burley 2020's
Laine-Karras permutation), a closed form (microfacet normalisation, area x
luminance light weights, luminance x sin(theta) environment weights), or the
reference's conversion rules (WGSL's saturating i32(), wrapping i32 adds).
No GPU needed.
"""
import math

import numpy as np
import pytest

import oracle
import oxpt

LUM = np.array([0.2126, 0.7152, 0.0722])


# ---------------------------------------------------------------------------
# Sobol' direction numbers (the table sobol() reads, pathtrace.wgsl:168-183).
# ---------------------------------------------------------------------------

def test_sobol_words_equal_scipy_joe_kuo():
    """scipy.stats.qmc.Sobol (Joe & Kuo 2008 direction numbers, unscrambled,
    32 bits) emits its points in Gray-code order: point n is the natural-order
    point of index n ^ (n >> 1). Point 2^(k+1) - 1 is therefore index 2^k,
    whose coordinates are direction word k of each dimension."""
    from scipy.stats import qmc

    v = oracle.sobol_directions().reshape(4, 32).astype(np.uint64)
    # Words 0..26 from scipy (stepping to point 2^28 - 1 takes ~1 s; the
    # last words would take minutes) ...
    for k in range(27):
        eng = qmc.Sobol(4, scramble=False, bits=32)
        eng.fast_forward((1 << (k + 1)) - 1)
        pt = np.round(eng.random(1)[0] * 2.0 ** 32).astype(np.uint64)
        assert np.array_equal(pt, v[:, k]), (k, pt, v[:, k])
    # ... and every word by the Bratley-Fox recurrence of the Joe-Kuo
    # primitive polynomials (degree s, inner coefficients a):
    # v_k = v_{k-s} ^ (v_{k-s} >> s) ^ XOR_{j<s} a_j v_{k-j}.
    for dim, (s_deg, a) in zip((1, 2, 3), ((1, 0), (2, 1), (3, 1))):
        for k in range(s_deg, 32):
            x = int(v[dim, k - s_deg]) ^ (int(v[dim, k - s_deg]) >> s_deg)
            for j in range(1, s_deg):
                if (a >> (s_deg - 1 - j)) & 1:
                    x ^= int(v[dim, k - j])
            assert x == int(v[dim, k]), (dim, k)
    # The first 2^10 points, in Gray-code order, as a whole.
    eng = qmc.Sobol(4, scramble=False, bits=32)
    pts = np.round(eng.random(1 << 10) * 2.0 ** 32).astype(np.uint64)
    for n in range(1 << 10):
        g = n ^ (n >> 1)
        want = np.zeros(4, dtype=np.uint64)
        for k in range(10):
            if (g >> k) & 1:
                want ^= v[:, k]
        assert np.array_equal(pts[n], want), n
    assert np.array_equal(oracle.sobol_directions(), oxpt.sobol_directions())


def _t_value(xs, ys, m):
    """Smallest t such that the 2^m points form a (t, m, 2)-net in base 2."""
    for t in range(m + 1):
        ok = True
        k = m - t
        for a in range(k + 1):
            bx = (xs >> np.uint64(32 - a)) if a else np.zeros_like(xs)
            by = (ys >> np.uint64(32 - (k - a))) if k - a else np.zeros_like(ys)
            cells = bx * np.uint64(1 << (k - a)) + by
            if np.bincount(cells.astype(np.int64), minlength=1 << k).max() != 1 << t:
                ok = False
                break
        if ok:
            return t
    return m


def test_sobol_two_dimensional_projections():
    """t-values of every 2-D projection of the first 2^m unscrambled points,
    m <= 12, computed by brute force from the oracle's words and from scipy's:
    equal, and (0, 1) is a (0, 2)-sequence."""
    from scipy.stats import qmc

    v = oracle.sobol_directions().reshape(4, 32).astype(np.uint64)
    n = 1 << 12
    idx = np.arange(n, dtype=np.uint64)
    ours = np.zeros((n, 4), dtype=np.uint64)
    for k in range(12):
        bit = (idx >> np.uint64(k)) & np.uint64(1)
        ours ^= bit[:, None] * v[None, :, k]
    ref = np.round(qmc.Sobol(4, scramble=False, bits=32).random(n) * 2.0 ** 32).astype(np.uint64)
    for m in (4, 8, 12):
        # The first 2^m points are the same set in either order.
        for a in range(4):
            for b in range(a + 1, 4):
                ta = _t_value(ours[: 1 << m, a], ours[: 1 << m, b], m)
                tb = _t_value(ref[: 1 << m, a], ref[: 1 << m, b], m)
                assert ta == tb, (m, a, b, ta, tb)
                if (a, b) == (0, 1):
                    assert ta == 0


# ---------------------------------------------------------------------------
# Hashes: PCG RXS-M-XS (pathtrace.wgsl:158-166) and Laine-Karras
# (pathtrace.wgsl:185-196).
# ---------------------------------------------------------------------------

# pcg-c (O'Neill 2014, pcg_variants.h): PCG_DEFAULT_MULTIPLIER_32,
# PCG_DEFAULT_INCREMENT_32, and the 32-bit RXS M XS output multiplier.
PCG_MULT, PCG_INC, RXS_M_XS_MULT = 747796405, 2891336453, 277803737


def _pcg_published(x):
    """One LCG step of pcg-c's 32-bit generator from state x, then the
    pcg_output_rxs_m_xs_32_32 permutation of the new state."""
    state = (x * PCG_MULT + PCG_INC) & 0xFFFFFFFF
    word = (((state >> ((state >> 28) + 4)) ^ state) * RXS_M_XS_MULT) & 0xFFFFFFFF
    return ((word >> 22) ^ word) & 0xFFFFFFFF


def test_pcg_hash_known_answers():
    rng = np.random.default_rng(4)
    xs = [0, 1, 2, 0xFFFFFFFF, 0x80000000, 12345] + [int(v) for v in rng.integers(0, 1 << 32, 4000, dtype=np.uint64)]
    for x in xs:
        assert oracle.lib.oxo_pcg_hash(x) == _pcg_published(x), hex(x)
        b = (x * 2654435761) & 0xFFFFFFFF
        assert oracle.lib.oxo_hash_mix(x, b) == _pcg_published(x ^ _pcg_published(b))


def _rev32(x):
    return int(f"{x:032b}"[::-1], 2)


def _lk_burley(x, seed):
    """Burley 2020, "Practical Hash-based Owen Scrambling", listing 2
    (laine_karras_permutation) and nested_uniform_scramble."""
    x = (x + seed) & 0xFFFFFFFF
    for c in (0x6C50B47C, 0xB82F1E52, 0xC7AFE638, 0x8D22F6E6):
        x ^= (x * c) & 0xFFFFFFFF
    return x


def test_nested_uniform_scramble_known_answers():
    rng = np.random.default_rng(5)
    for _ in range(4000):
        x, seed = (int(v) for v in rng.integers(0, 1 << 32, 2, dtype=np.uint64))
        want = _rev32(_lk_burley(_rev32(x), seed))
        assert oracle.lib.oxo_nested_uniform_scramble(x, seed) == want, (hex(x), hex(seed))


# ---------------------------------------------------------------------------
# Microfacet distributions and the sheen albedo table.
# ---------------------------------------------------------------------------

@pytest.mark.filterwarnings("ignore::scipy.integrate.IntegrationWarning")
@pytest.mark.parametrize("kind,alphas", [(0, (0.05, 0.2, 0.5, 0.9, 1.0)), (1, (0.1, 0.3, 0.6, 1.0))],
                         ids=["ggx", "charlie"])
def test_ndf_is_normalised(kind, alphas):
    """Projected-area normalisation: 2 pi * int_0^{pi/2} D(cos t) cos t sin t dt = 1.
    A dropped pi, a missing (2 + 1/a) or a / vs a^2 slip fails it."""
    from scipy.integrate import quad

    for a in alphas:
        f = lambda t: oracle.ndf(kind, math.cos(t), a) * math.cos(t) * math.sin(t)  # noqa: E731
        pts = [math.atan(a) * s for s in (0.25, 1.0, 4.0) if math.atan(a) * s < math.pi / 2]
        val, _ = quad(f, 0.0, math.pi / 2, points=pts, limit=400, epsabs=1e-9, epsrel=1e-7)
        assert abs(2 * math.pi * val - 1.0) < 2e-4, (kind, a, 2 * math.pi * val)


def test_sheen_table_matches_scipy_quadrature():
    """E(mu, r) = int_{hemisphere} D_charlie(n.h; max(r^2, 1e-3)) V_neubelt(n.l, n.v) (n.l) dw_l
    at the table's grid (mu = (i + 0.5)/32, r = j/15), by scipy's adaptive
    quadrature over (theta_l, phi_l) — independent of the table builder's
    midpoint rule. Rough lobes (r >= 0.4) are smooth enough for dblquad."""
    from scipy.integrate import dblquad

    table = oxpt.sheen_table().reshape(16, 32)
    for j in (6, 9, 12, 15):
        r = j / 15.0
        a = max(r * r, 1e-3)
        inv = 1.0 / a
        for i in (0, 7, 16, 31):
            mu = (i + 0.5) / 32.0
            vs = math.sqrt(max(1.0 - mu * mu, 0.0))

            def integrand(ph, th):
                cl, sl = math.cos(th), math.sin(th)
                hx, hy, hz = vs + sl * math.cos(ph), sl * math.sin(ph), mu + cl
                nh = hz / math.sqrt(hx * hx + hy * hy + hz * hz)
                s2 = max(1.0 - nh * nh, 0.0)
                d = (2.0 + inv) * s2 ** (0.5 * inv) / (2.0 * math.pi)
                v = 1.0 / max(4.0 * (cl + mu - cl * mu), 1e-6)
                return d * v * cl * sl

            want, _ = dblquad(integrand, 0.0, math.pi / 2, 0.0, 2.0 * math.pi, epsabs=1e-7, epsrel=1e-6)
            got = float(table[j, i])
            assert abs(got - want) <= 1e-3 + 5e-3 * abs(want), (i, j, got, want)


# ---------------------------------------------------------------------------
# Light-selection weights (EmissiveLights::build, EnvironmentMap; consumed at
# src/pathtrace.rs:1201-1210,1428-1461).
# ---------------------------------------------------------------------------

def test_emissive_pmf_is_area_times_luminance():
    """Three emitters of unequal areas and radiances: pmf_i = A_i lum_i /
    sum_j A_j lum_j. Area-only or luminance-only weights give other values."""
    s = oxpt.Scene()
    wall = s.add_material(oxpt.material(base_color=(0.5, 0.5, 0.5, 1), metallic=0.0, roughness=1.0))
    s.add_triangles(np.array([[-5, 0, -5, 5, 0, -5, 0, 0, 5]], np.float32), wall)
    lights = [((0.0, 2.0, 0.0), 1.0, (4.0, 1.0, 0.5)),
              ((2.0, 2.0, 0.0), 2.5, (0.2, 3.0, 0.1)),
              ((-2.0, 2.0, 0.0), 0.4, (9.0, 9.0, 20.0))]
    want = []
    for (cx, cy, cz), side, le in lights:
        m = s.add_material(oxpt.material(base_color=(0, 0, 0, 1), metallic=0.0, roughness=1.0, emissive=le))
        tri = np.array([[cx, cy, cz, cx + side, cy, cz, cx, cy, cz + side]], np.float32)
        s.add_triangles(tri, m)
        want.append(0.5 * side * side * float(LUM @ np.array(le)))
    want = np.array(want) / sum(want)
    pk = oxpt.Packed(s, oxpt.RenderOptions(width=8, height=8))
    P = pk.params
    assert P.cfg[3] == 3
    rec = pk.tables[P.offs1[0]: P.offs1[0] + 12].reshape(3, 4)
    pmf = rec[:, 1].view(np.float32).astype(np.float64)
    cdf = rec[:, 0].view(np.float32).astype(np.float64)
    assert np.allclose(pmf, want, rtol=2e-6, atol=1e-7), (pmf, want)
    assert np.allclose(cdf, np.cumsum(want), rtol=2e-6, atol=1e-7)
    assert list(rec[:, 2]) == [1, 2, 3]  # global triangle ids, in input order


def test_environment_mass_is_luminance_times_sin_theta():
    """A black map with bright texels in chosen rows: texel mass =
    lum(L) sin(pi (y + 0.5) / h) / sum, so a missing sin(theta) (equal
    masses for equal texels in different rows) fails."""
    w, h = 16, 8
    rgb = np.zeros((h, w, 3), np.float32)
    spots = [(0, 3, (1.0, 1.0, 1.0)), (3, 5, (1.0, 1.0, 1.0)), (6, 12, (0.5, 2.0, 0.25)), (7, 0, (3.0, 0.0, 0.0))]
    for y, x, c in spots:
        rgb[y, x] = c
    env = oxpt.EnvironmentMap(rgb, 2.0)
    pk = oxpt.Packed(oxpt.Scene.room(0), oxpt.RenderOptions(width=8, height=8), env)
    P = pk.params
    mass = pk.tables[P.offs2[0]: P.offs2[0] + w * h].view(np.float32).reshape(h, w).astype(np.float64)
    wt = np.zeros((h, w))
    for y, x, c in spots:
        wt[y, x] = 2.0 * float(LUM @ np.array(c)) * math.sin(math.pi * (y + 0.5) / h)
    want = wt / wt.sum()
    assert np.allclose(mass, want, rtol=2e-6, atol=1e-9), (mass[mass > 0], want[want > 0])
    rad = pk.tables[P.offs2[1]: P.offs2[1] + 3 * w * h].view(np.float32).reshape(h, w, 3)
    assert np.array_equal(rad, rgb * np.float32(2.0))


# ---------------------------------------------------------------------------
# Integer helpers: the texel index conversions (filter_level,
# pathtrace.wgsl:335-380) and the batched claim step's exact divisions.
# ---------------------------------------------------------------------------

def _texel_ref(u, size, mode):
    """The reference's index steps in exact arithmetic: f32 products, WGSL's
    saturating i32(), wrapping i32 adds, and wrap_index with mathematical
    remainders."""
    u = np.float32(u)
    if not np.isfinite(u):
        u = np.float32(0.0)
    u = np.float32(min(max(float(u), -1.0e6), 1.0e6))
    x = np.float32(u * np.float32(size))

    def i32(f):
        f = float(f)
        return 2147483647 if f >= 2147483648.0 else (-2147483648 if f < -2147483648.0 else int(f))

    def wrap_add1(i):
        return ((i + 1 + 2 ** 31) % 2 ** 32) - 2 ** 31

    def wrap(i, n):
        if mode == 0:
            return min(max(i, 0), n - 1)
        if mode == 1:
            m = i % (2 * n)
            return 2 * n - 1 - m if m >= n else m
        return i % n

    x0 = i32(math.floor(float(np.float32(x - np.float32(0.5)))))
    return wrap(i32(math.floor(float(x))), size), wrap(x0, size), wrap(wrap_add1(x0), size)


def texel_cases():
    rng = np.random.default_rng(12)
    cases = []
    for size in (1, 2, 3, 7, 64, 255, 2047, 2048, 4096, 65535):
        for mode in (0, 1, 2):
            for u in (0.0, 0.49, 0.5, 1.0, -0.25, -1.0, 1e6, -1e6, 999999.9, -999999.9, 1e7, -1e7,
                      float("nan"), float("inf"), -float("inf"), 2147483648.0 / max(size, 1),
                      -2147483648.0 / max(size, 1)):
                cases.append((u, size, mode))
            for u in rng.uniform(-1e6, 1e6, 20):
                cases.append((float(u), size, mode))
    return cases


def test_oracle_texel_index_follows_the_reference_conversions():
    bad = [(u, n, m) for u, n, m in texel_cases() if oracle.texel_index(u, n, m) != _texel_ref(u, n, m)]
    assert not bad, bad[:5]


UDIV_DIVISORS = [1, 2, 3, 5, 6, 7, 63, 64, 65, 640, 4096, 1 << 16, (1 << 16) + 1, (1 << 31) - 1, 1 << 31,
                 (1 << 31) + 1, 0xFFFFFFFE, 0xFFFFFFFF] + [(1 << k) + d for k in range(2, 32) for d in (-1, 1)]


def _udiv_numerators(d):
    base = {0, 1, d - 1, d, d + 1, 2 * d - 1, 2 * d, 0xFFFFFFFF, 0xFFFFFFFE, 0x80000000, 0x7FFFFFFF}
    base |= {(q * d - 1) & 0xFFFFFFFF for q in (3, 1000, 65535)}
    return sorted(n for n in base if 0 <= n <= 0xFFFFFFFF)


def test_udiv_magic_host_mirror_is_exact():
    """The device quotient t = mulhi(n, m); q = (t + ((n - t) >> 1)) >> (l - 1)
    (l = 0: q = n), restated on the host with the library's (m, l): equal to
    n // d on every divisor's boundary numerators."""
    for d in UDIV_DIVISORS:
        m, l = oxpt.udiv_magic(d)
        for n in _udiv_numerators(d):
            if l == 0:
                q = n
            else:
                t = (n * m) >> 32
                q = (t + ((n - t) >> 1)) >> (l - 1)
            assert q == n // d, (n, d, m, l, q)
