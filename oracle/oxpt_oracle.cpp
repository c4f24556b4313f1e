// oxpt_oracle.cpp — CPU oracle of the path tracer. TEST INFRASTRUCTURE ONLY:
// loaded by tests/, the smoke check and the benchmark's CPU-baseline leg,
// never by the product library.
//
// Each function restates its counterpart in the reference kernel
// (oxideav-render-vulkan src/shaders/pathtrace.wgsl, cited as wgsl:LINE),
// which itself mirrors the CPU `oxideav_render::PathTracer` function by
// function. Scalar f32 arithmetic in the kernel's operation order; built
// with -ffp-contract=off so no fused multiply-add changes a rounding.
//
// Parity pins (what checks each part; see tests/test_oracle.py):
//  * RNG + Sobol'/Owen: closed-form integer code; net-stratification and
//    bijection invariants; direction numbers from the Joe-Kuo recurrence.
//  * traversal: brute-force closest hit over every leaf triangle.
//  * BSDF: sample pdf == eval pdf, pdf integrates to <= 1, reciprocity.
//  * lights: Arvo samples uniform in solid angle, solid angle vs quadrature.
//  * estimator: the reference's white furnace (tests/pathtrace.rs:295-317),
//    L = Le / (1 - rho) = 2 for all three strategies.
//  * resolve: closed forms and exact background bytes.
//  * f32_to_f16: the reference KAT (src/scene.rs:496-511) + numpy sweep.
#include "oxpt_oracle.h"

#include <atomic>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

namespace {

constexpr float PI = 3.14159265358979f;
constexpr float INF = 3.0e38f;
constexpr uint32_t NONE = 0xffffffffu;
constexpr float MIN_ALPHA = 1.0e-3f;
constexpr float MIN_SPHERICAL_SOLID_ANGLE = 1.0e-4f;

struct V3 {
    float x, y, z;
};
struct V4 {
    float x, y, z, w;
};

inline V3 v3(float a) { return {a, a, a}; }
inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator*(V3 a, V3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
inline V3 operator*(V3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
inline V3 operator*(float s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
inline V3 operator/(V3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
inline V3 operator/(V3 a, V3 b) { return {a.x / b.x, a.y / b.y, a.z / b.z}; }
inline V3 operator-(V3 a) { return {-a.x, -a.y, -a.z}; }
inline float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(V3 a, V3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
inline float comp(V3 a, uint32_t k) { return k == 0 ? a.x : (k == 1 ? a.y : a.z); }
inline V4 operator+(V4 a, V4 b) { return {a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w}; }
inline V4 operator*(V4 a, V4 b) { return {a.x * b.x, a.y * b.y, a.z * b.z, a.w * b.w}; }
inline V4 operator*(V4 a, float s) { return {a.x * s, a.y * s, a.z * s, a.w * s}; }
inline V3 xyz(V4 a) { return {a.x, a.y, a.z}; }

// WGSL min/max/clamp on floats: the non-NaN operand wins (fminf/fmaxf).
inline float fmin_(float a, float b) { return std::fmin(a, b); }
inline float fmax_(float a, float b) { return std::fmax(a, b); }
inline float clampf(float x, float lo, float hi) { return fmin_(fmax_(x, lo), hi); }
inline V3 vmin(V3 a, V3 b) { return {fmin_(a.x, b.x), fmin_(a.y, b.y), fmin_(a.z, b.z)}; }
inline V3 vmax(V3 a, V3 b) { return {fmax_(a.x, b.x), fmax_(a.y, b.y), fmax_(a.z, b.z)}; }
// WGSL mix(a, b, t) = a * (1 - t) + b * t.
inline float mixf(float a, float b, float t) { return a * (1.0f - t) + b * t; }
inline V4 mix4(V4 a, V4 b, float t) {
    return {mixf(a.x, b.x, t), mixf(a.y, b.y, t), mixf(a.z, b.z, t), mixf(a.w, b.w, t)};
}

inline uint32_t fbits(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    return u;
}
inline float bitsf(uint32_t u) {
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}
inline uint32_t rev32(uint32_t v) {
    v = ((v >> 1) & 0x55555555u) | ((v & 0x55555555u) << 1);
    v = ((v >> 2) & 0x33333333u) | ((v & 0x33333333u) << 2);
    v = ((v >> 4) & 0x0f0f0f0fu) | ((v & 0x0f0f0f0fu) << 4);
    v = ((v >> 8) & 0x00ff00ffu) | ((v & 0x00ff00ffu) << 8);
    return (v >> 16) | (v << 16);
}

// wgsl:134-156
inline bool finite(float x) { return (fbits(x) & 0x7f800000u) != 0x7f800000u; }
inline bool finite3(V3 v) { return finite(v.x) && finite(v.y) && finite(v.z); }
inline bool is_black(V3 a) { return !(a.x > 0.0f || a.y > 0.0f || a.z > 0.0f); }
inline float max3(V3 a) { return fmax_(fmax_(a.x, a.y), a.z); }
inline float mean3(V3 a) { return (a.x + a.y + a.z) * (1.0f / 3.0f); }
inline V3 nrm(V3 v) {
    float len = std::sqrt(dot(v, v));
    if (len < 1.1920929e-7f) return v3(0.0f);
    return v / len;
}
inline float sign_of(float x) { return (fbits(x) & 0x80000000u) ? -1.0f : 1.0f; }

// ---------------------------------------------------------------------
// RNG (wgsl:158-228).
// ---------------------------------------------------------------------

inline uint32_t pcg_hash(uint32_t x) {
    uint32_t state = x * 747796405u + 2891336453u;
    uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
    return (word >> 22u) ^ word;
}
inline uint32_t hash_mix(uint32_t a, uint32_t b) { return pcg_hash(a ^ pcg_hash(b)); }

inline uint32_t sobol(const uint32_t *tab, uint32_t index, uint32_t dim) {
    uint32_t r = 0, i = index, k = 0;
    while (i != 0) {
        if (i & 1u) r ^= tab[dim * 32u + k];
        i >>= 1u;
        k += 1u;
    }
    return r;
}
inline uint32_t laine_karras(uint32_t x0, uint32_t seed) {
    uint32_t x = x0 + seed;
    x ^= x * 0x6c50b47cu;
    x ^= x * 0xb82f1e52u;
    x ^= x * 0xc7afe638u;
    x ^= x * 0x8d22f6e6u;
    return x;
}
inline uint32_t nested_uniform_scramble(uint32_t x, uint32_t seed) {
    return rev32(laine_karras(rev32(x), seed));
}
inline float to_unit(uint32_t v) { return (float)(v >> 8u) * (1.0f / 16777216.0f); }

inline void sobol_owen_words(const uint32_t *tab, uint32_t index, uint32_t pattern_seed,
                             uint32_t out[4]) {
    uint32_t i = nested_uniform_scramble(index, pattern_seed);
    for (uint32_t d = 0; d < 4; d++)
        out[d] = nested_uniform_scramble(sobol(tab, i, d), hash_mix(pattern_seed, d));
}

struct Sampler {
    uint32_t pixel_seed, index, path_seed;
};
inline Sampler make_sampler(uint32_t seed, uint32_t x, uint32_t y, uint32_t index) {
    uint32_t ps = hash_mix(hash_mix(seed, y), x);
    return {ps, index, hash_mix(ps, index)};
}

// ---------------------------------------------------------------------
// Scene access.
// ---------------------------------------------------------------------

struct Ctx {
    const float *geom;
    const float *attrs;
    const uint32_t *tables;
    const uint16_t *atlas;
    uint32_t atlas_side;
    oxpt_params P;
    bool f64_lights;
    oxo_counts *ctr;
    void add(uint64_t b) const {
        if (ctr) ctr->bytes += b;
    }
};

inline V4 g4(const Ctx &c, uint32_t i) {
    const float *p = c.geom + 4ull * i;
    return {p[0], p[1], p[2], p[3]};
}
inline V4 a4(const Ctx &c, uint32_t i) {
    const float *p = c.attrs + 4ull * i;
    return {p[0], p[1], p[2], p[3]};
}
inline float tf(const Ctx &c, uint32_t i) { return bitsf(c.tables[i]); }
inline V3 tv3(const Ctx &c, uint32_t i) { return {tf(c, i), tf(c, i + 1), tf(c, i + 2)}; }
inline V4 tv4(const Ctx &c, uint32_t i) {
    return {tf(c, i), tf(c, i + 1), tf(c, i + 2), tf(c, i + 3)};
}

inline uint32_t tri_base(const Ctx &c, uint32_t slot) { return c.P.offs2[2] + 3u * slot; }
inline V3 tri_pos(const Ctx &c, uint32_t slot, uint32_t v) { return xyz(g4(c, tri_base(c, slot) + v)); }
inline uint32_t tri_global(const Ctx &c, uint32_t slot) { return fbits(g4(c, tri_base(c, slot)).w); }
inline uint32_t tri_light(const Ctx &c, uint32_t slot) { return fbits(g4(c, tri_base(c, slot) + 1).w); }
inline uint32_t tri_material(const Ctx &c, uint32_t slot) { return fbits(g4(c, tri_base(c, slot) + 2).w); }
inline V4 attr(const Ctx &c, uint32_t slot, uint32_t k) { return a4(c, slot * 12u + k); }
inline uint32_t attr_flags(const Ctx &c, uint32_t slot) { return fbits(a4(c, slot * 12u).w); }
inline V3 interp3a(const Ctx &c, uint32_t slot, uint32_t k, V3 b) {
    return xyz(attr(c, slot, k)) * b.x + xyz(attr(c, slot, k + 1)) * b.y + xyz(attr(c, slot, k + 2)) * b.z;
}
inline V4 interp4a(const Ctx &c, uint32_t slot, uint32_t k, V3 b) {
    return attr(c, slot, k) * b.x + attr(c, slot, k + 1) * b.y + attr(c, slot, k + 2) * b.z;
}
struct V2 {
    float x, y;
};
inline V2 vertex_uv(const Ctx &c, uint32_t slot, uint32_t uvset, uint32_t v) {
    V4 a = attr(c, slot, 6u + v);
    if (uvset == 0) return {a.x, a.y};
    return {a.z, a.w};
}
inline bool has_uv(const Ctx &c, uint32_t slot, uint32_t uvset) {
    uint32_t f = attr_flags(c, slot);
    if (uvset == 0) return (f & OXPT_F_UV0) != 0;
    if (uvset == 1) return (f & OXPT_F_UV1) != 0;
    return false;
}
inline uint32_t mat_off(const Ctx &c, uint32_t m) { return c.P.offs0[2] + m * OXPT_MAT_STRIDE; }
inline uint32_t mat_flags(const Ctx &c, uint32_t m) { return c.tables[mat_off(c, m) + OXPT_M_FLAGS]; }
inline bool mat_double_sided(const Ctx &c, uint32_t m) { return (mat_flags(c, m) & 4u) != 0; }

// ---------------------------------------------------------------------
// Textures (wgsl:310-507).
// ---------------------------------------------------------------------

struct Level {
    int32_t layer, x, y, w, h;
};

Level tex_level(const Ctx &c, uint32_t desc, uint32_t level) {
    uint32_t skip = c.tables[desc + 4], kept = c.tables[desc + 5];
    int32_t li = (int32_t)level - (int32_t)skip;
    uint32_t l = std::min((uint32_t)std::max(li, 0), kept - 1u);
    uint32_t o = desc + 8u + 4u * l;
    uint32_t wh = c.tables[o + 3];
    c.add(16);
    return {(int32_t)c.tables[o], (int32_t)c.tables[o + 1], (int32_t)c.tables[o + 2],
            (int32_t)(wh & 0xffffu), (int32_t)(wh >> 16u)};
}

// wrap_index (wgsl:335-347): clamp / mirrored repeat / repeat.
int32_t wrap_index(int32_t i, int32_t n, uint32_t mode) {
    if (mode == 0) return std::min(std::max(i, 0), n - 1);
    if (mode == 1) {
        int32_t m = ((i % (2 * n)) + 2 * n) % (2 * n);
        if (m >= n) return 2 * n - 1 - m;
        return m;
    }
    return ((i % n) + n) % n;
}

// i32(floor(x)) as the reference converts it (WGSL conversion rules, the CPU
// tracer's `as i32`): saturating at the i32 range. x is finite.
int32_t sat_i32(float f) {
    if (f >= 2147483648.0f) return INT32_MAX;
    if (f < -2147483648.0f) return INT32_MIN;
    return (int32_t)f;
}
// i + 1 in wrapping i32 arithmetic (WGSL integer addition).
int32_t next_i32(int32_t i) { return (int32_t)((uint32_t)i + 1u); }

float half_to_float(uint16_t h) {
    uint32_t s = (uint32_t)(h & 0x8000u) << 16, e = (h >> 10) & 0x1fu, m = h & 0x3ffu;
    if (e == 0) {
        float f = (float)m * (1.0f / 16777216.0f);
        return s ? -f : f;
    }
    if (e == 31) return bitsf(s | 0x7f800000u | (m << 13));
    return bitsf(s | ((e + 112u) << 23) | (m << 13));
}

V4 fetch(const Ctx &c, const Level &lv, int32_t x, int32_t y, uint32_t ws, uint32_t wt) {
    int32_t xi = wrap_index(x, lv.w, ws), yi = wrap_index(y, lv.h, wt);
    uint64_t side = c.atlas_side;
    uint64_t at = (((uint64_t)lv.layer * side + (uint64_t)(lv.y + yi)) * side + (uint64_t)(lv.x + xi)) * 4u;
    const uint16_t *t = c.atlas + at;
    if (c.ctr) c.ctr->tex_taps++;
    c.add(8);
    return {half_to_float(t[0]), half_to_float(t[1]), half_to_float(t[2]), half_to_float(t[3])};
}

V4 filter_level(const Ctx &c, const Level &lv, V2 uv, bool linear, uint32_t ws, uint32_t wt) {
    float u = uv.x, v = uv.y;
    if (!finite(u)) u = 0.0f;
    if (!finite(v)) v = 0.0f;
    u = clampf(u, -1.0e6f, 1.0e6f);
    v = clampf(v, -1.0e6f, 1.0e6f);
    float x = u * (float)lv.w, y = v * (float)lv.h;
    if (!linear) return fetch(c, lv, sat_i32(std::floor(x)), sat_i32(std::floor(y)), ws, wt);
    float xf = x - 0.5f, yf = y - 0.5f;
    float x0f = std::floor(xf), y0f = std::floor(yf);
    float fx = xf - x0f, fy = yf - y0f;
    int32_t x0 = sat_i32(x0f), y0 = sat_i32(y0f);
    V4 taps[4];
    for (int t = 0; t < 4; t++)
        taps[t] = fetch(c, lv, (t & 1) ? next_i32(x0) : x0, (t >> 1) ? next_i32(y0) : y0, ws, wt);
    return mix4(mix4(taps[0], taps[1], fx), mix4(taps[2], taps[3], fx), fy);
}

V4 sample_lod(const Ctx &c, uint32_t desc, V2 uv, float lod) {
    c.add(32);
    uint32_t levels = c.tables[desc], flags = c.tables[desc + 1];
    uint32_t ws = flags & 3u, wt = (flags >> 2) & 3u, minf = (flags >> 5) & 7u;
    bool lin = minf == 1 || minf == 3 || minf == 5;
    float max_level = (float)(levels - 1u);
    uint32_t l0 = 0, l1 = 0;
    float f = 0.0f;
    if (!finite(lod) || lod <= 0.0f) {
        lin = (flags & 16u) != 0;
    } else if (minf <= 1) {
        l0 = 0;
    } else if (minf <= 3) {
        l0 = (uint32_t)clampf(std::floor(lod + 0.5f), 0.0f, max_level);
    } else {
        float l = clampf(lod, 0.0f, max_level);
        l0 = (uint32_t)std::floor(l);
        l1 = std::min(l0 + 1u, levels - 1u);
        f = l - (float)l0;
    }
    uint32_t n = (f > 0.0f && l0 != l1) ? 2u : 1u;
    V4 r[2];
    for (uint32_t i = 0; i < n; i++) {
        uint32_t level = i == 1 ? l1 : l0;
        r[i] = filter_level(c, tex_level(c, desc, level), uv, lin, ws, wt);
    }
    if (n == 1) return r[0];
    return mix4(r[0], r[1], f);
}

inline uint32_t slot_word(const Ctx &c, uint32_t m, uint32_t s) { return mat_off(c, m) + OXPT_M_SLOTS + 8u * s; }
inline bool slot_present(const Ctx &c, uint32_t m, uint32_t s, uint32_t slot) {
    uint32_t o = slot_word(c, m, s);
    return c.tables[o] != NONE && has_uv(c, slot, c.tables[o + 1]);
}
inline V2 slot_xform(const Ctx &c, uint32_t m, uint32_t s, V2 uv) {
    uint32_t o = slot_word(c, m, s);
    return {tf(c, o + 2) * uv.x + tf(c, o + 3) * uv.y + tf(c, o + 4),
            tf(c, o + 5) * uv.x + tf(c, o + 6) * uv.y + tf(c, o + 7)};
}

struct Lod {
    float width;
    V3 dir;
};
inline Lod base_lod() { return {-1.0f, v3(0.0f)}; }

float cone_lod(const Ctx &c, uint32_t m, uint32_t s, uint32_t slot, const Lod &lod) {
    uint32_t o = slot_word(c, m, s);
    uint32_t desc = c.tables[o], uvset = c.tables[o + 1];
    V2 uv0 = slot_xform(c, m, s, vertex_uv(c, slot, uvset, 0));
    V2 uv1 = slot_xform(c, m, s, vertex_uv(c, slot, uvset, 1));
    V2 uv2 = slot_xform(c, m, s, vertex_uv(c, slot, uvset, 2));
    float ta = (float)c.tables[desc + 2] * (float)c.tables[desc + 3] *
               std::fabs((uv1.x - uv0.x) * (uv2.y - uv0.y) - (uv2.x - uv0.x) * (uv1.y - uv0.y));
    V3 p0 = tri_pos(c, slot, 0);
    V3 cr = cross(tri_pos(c, slot, 1) - p0, tri_pos(c, slot, 2) - p0);
    float pa = std::sqrt(dot(cr, cr));
    if (!(ta > 0.0f && pa > 0.0f && lod.width > 0.0f)) return 0.0f;
    V3 n = cr / pa;
    float cc = fmax_(std::fabs(dot(n, lod.dir)), 1.0e-3f);
    float l = 0.5f * std::log2(ta / pa) + std::log2(lod.width / cc);
    if (finite(l)) return l;
    return 0.0f;
}

V4 slot_sample(const Ctx &c, uint32_t m, uint32_t s, uint32_t slot, V3 b, const Lod &lod) {
    uint32_t o = slot_word(c, m, s);
    uint32_t uvset = c.tables[o + 1];
    V2 a0 = vertex_uv(c, slot, uvset, 0), a1 = vertex_uv(c, slot, uvset, 1), a2 = vertex_uv(c, slot, uvset, 2);
    V2 uv = {a0.x * b.x + a1.x * b.y + a2.x * b.z, a0.y * b.x + a1.y * b.y + a2.y * b.z};
    float l = 0.0f;
    if (lod.width >= 0.0f) l = cone_lod(c, m, s, slot, lod);
    return sample_lod(c, c.tables[o], slot_xform(c, m, s, uv), l);
}

float hit_alpha(const Ctx &c, uint32_t slot, V3 b) {
    uint32_t m = tri_material(c, slot);
    float a = tf(c, mat_off(c, m) + 3);
    if (slot_present(c, m, OXPT_S_BASE, slot)) a *= slot_sample(c, m, OXPT_S_BASE, slot, b, base_lod()).w;
    if (attr_flags(c, slot) & OXPT_F_COLORS) a *= interp4a(c, slot, 9, b).w;
    return a;
}

// ---------------------------------------------------------------------
// Rays and traversal (wgsl:513-735).
// ---------------------------------------------------------------------

struct Ray {
    V3 o, d, inv;
    uint32_t kx, ky, kz;
    V3 shear;
    bool valid;
};

Ray make_ray(V3 o, V3 d) {
    Ray r;
    r.o = o;
    r.d = d;
    float dk[3] = {d.x, d.y, d.z}, inv[3];
    for (int k = 0; k < 3; k++) inv[k] = std::fabs(dk[k]) < 1.0e-30f ? sign_of(dk[k]) * 1.0e30f : 1.0f / dk[k];
    r.inv = {inv[0], inv[1], inv[2]};
    V3 ad = {std::fabs(d.x), std::fabs(d.y), std::fabs(d.z)};
    uint32_t kz = 2;
    if (ad.x >= ad.y && ad.x >= ad.z) kz = 0;
    else if (ad.y >= ad.z) kz = 1;
    uint32_t kx = (kz + 1) % 3, ky = (kx + 1) % 3;
    if (dk[kz] < 0.0f) std::swap(kx, ky);
    r.kx = kx;
    r.ky = ky;
    r.kz = kz;
    r.shear = v3(0.0f);
    if (dk[kz] != 0.0f) r.shear = {dk[kx] / dk[kz], dk[ky] / dk[kz], 1.0f / dk[kz]};
    r.valid = finite3(o) && finite3(d) && (d.x != 0.0f || d.y != 0.0f || d.z != 0.0f);
    return r;
}

float slab(const Ray &r, V3 bmin, V3 bmax, float t_max) {
    V3 t0s = (bmin - r.o) * r.inv, t1s = (bmax - r.o) * r.inv;
    V3 tn = vmin(t0s, t1s), tf_ = vmax(t0s, t1s);
    float t0 = fmax_(fmax_(0.0f, tn.x), fmax_(tn.y, tn.z));
    float t1 = fmin_(t_max, fmin_(tf_.x, fmin_(tf_.y, tf_.z)) * 1.0000004f);
    return t0 <= t1 ? t0 : -1.0f;
}

struct Hit {
    uint32_t slot;
    float t;
    V3 b;
    bool front;
};

Hit intersect(const Ctx &c, const Ray &r, uint32_t slot, float t_max) {
    Hit h;
    h.slot = NONE;
    V3 p0 = tri_pos(c, slot, 0), p1 = tri_pos(c, slot, 1), p2 = tri_pos(c, slot, 2);
    V3 a = p0 - r.o, b = p1 - r.o, cc = p2 - r.o;
    float ax = comp(a, r.kx) - r.shear.x * comp(a, r.kz);
    float ay = comp(a, r.ky) - r.shear.y * comp(a, r.kz);
    float bx = comp(b, r.kx) - r.shear.x * comp(b, r.kz);
    float by = comp(b, r.ky) - r.shear.y * comp(b, r.kz);
    float cx = comp(cc, r.kx) - r.shear.x * comp(cc, r.kz);
    float cy = comp(cc, r.ky) - r.shear.y * comp(cc, r.kz);
    float uu = cx * by - cy * bx;
    float vv = ax * cy - ay * cx;
    float ww = bx * ay - by * ax;
    if ((uu < 0.0f || vv < 0.0f || ww < 0.0f) && (uu > 0.0f || vv > 0.0f || ww > 0.0f)) return h;
    float det = uu + vv + ww;
    if (det == 0.0f || !finite(det)) return h;
    float az = r.shear.z * comp(a, r.kz), bz = r.shear.z * comp(b, r.kz), cz = r.shear.z * comp(cc, r.kz);
    float rcp = 1.0f / det;
    float t = (uu * az + vv * bz + ww * cz) * rcp;
    if (!(t >= 0.0f && t <= t_max)) return h;
    float u = vv * rcp, v = ww * rcp;
    h.slot = slot;
    h.t = t;
    h.b = {1.0f - u - v, u, v};
    h.front = dot(r.d, cross(p1 - p0, p2 - p0)) < 0.0f;
    return h;
}

inline float coin(const Sampler &s, uint32_t ray, uint32_t tri) {
    return to_unit(hash_mix(hash_mix(s.path_seed, ray), tri));
}

bool accept(const Ctx &c, const Hit &h, const Sampler &s, uint32_t ray_id, bool primary) {
    uint32_t m = tri_material(c, h.slot);
    uint32_t fl = mat_flags(c, m);
    if (primary && !h.front && (fl & 4u) == 0) return false;
    uint32_t mode = fl & 3u;
    if (mode == 0) return true;
    c.add(4);
    if (mode == 1) return hit_alpha(c, h.slot, h.b) >= tf(c, mat_off(c, m) + 11);
    return coin(s, ray_id, tri_global(c, h.slot)) < hit_alpha(c, h.slot, h.b);
}

Hit traverse(const Ctx &c, const Ray &r, float t_max_in, const Sampler &s, uint32_t ray_id,
             bool primary, bool any, uint32_t skip) {
    Hit best;
    best.slot = NONE;
    if (c.P.env[3] == 0 || !r.valid) return best;
    if (c.ctr) c.ctr->rays++;
    float t_max = t_max_in;
    c.add(32);
    if (slab(r, xyz(g4(c, 0)), xyz(g4(c, 1)), t_max) < 0.0f) return best;
    uint32_t stack[64];
    float stack_t[64];
    uint32_t sp = 0, idx = 0;
    for (;;) {
        V4 lo = g4(c, 2 * idx), hi = g4(c, 2 * idx + 1);
        uint32_t count = fbits(hi.w), first = fbits(lo.w);
        bool descended = false;
        if (count > 0) {
            bool done = false;
            for (uint32_t slot = first; slot < first + count; slot++) {
                if (c.ctr) c.ctr->tris++;
                c.add(48);
                Hit h = intersect(c, r, slot, t_max);
                if (h.slot != NONE && tri_global(c, slot) != skip && accept(c, h, s, ray_id, primary)) {
                    best = h;
                    if (any) {
                        done = true;
                        break;
                    }
                    t_max = h.t;
                }
            }
            if (done) break;
        } else {
            uint32_t l = first;
            if (c.ctr) c.ctr->nodes++;
            c.add(64);
            float tl = slab(r, xyz(g4(c, 2 * l)), xyz(g4(c, 2 * l + 1)), t_max);
            float tr = slab(r, xyz(g4(c, 2 * (l + 1))), xyz(g4(c, 2 * (l + 1) + 1)), t_max);
            if (tl >= 0.0f && tr >= 0.0f) {
                uint32_t near_ = l, far_ = l + 1;
                float t_far = tr;
                if (tl > tr) {
                    near_ = l + 1;
                    far_ = l;
                    t_far = tl;
                }
                if (sp < 64) {
                    stack[sp] = far_;
                    stack_t[sp] = t_far;
                    sp++;
                }
                idx = near_;
                descended = true;
            } else if (tl >= 0.0f) {
                idx = l;
                descended = true;
            } else if (tr >= 0.0f) {
                idx = l + 1;
                descended = true;
            }
        }
        if (descended) continue;
        bool found = false;
        while (sp != 0) {
            sp--;
            if (stack_t[sp] <= t_max) {
                idx = stack[sp];
                found = true;
                break;
            }
        }
        if (!found) break;
    }
    return best;
}

V3 offset_ray_origin(V3 p, V3 n) {
    float pk[3] = {p.x, p.y, p.z}, nk[3] = {n.x, n.y, n.z}, out[3];
    for (int k = 0; k < 3; k++) {
        int32_t of_i = (int32_t)(256.0f * nk[k]);
        int32_t bits = (int32_t)fbits(pk[k]);
        int32_t moved = pk[k] < 0.0f ? (int32_t)((uint32_t)bits - (uint32_t)of_i)
                                     : (int32_t)((uint32_t)bits + (uint32_t)of_i);
        float p_i = bitsf((uint32_t)moved);
        if (std::fabs(pk[k]) < 1.0f / 32.0f) out[k] = pk[k] + (1.0f / 65536.0f) * nk[k];
        else if (finite(p_i)) out[k] = p_i;
        else out[k] = pk[k];
    }
    return {out[0], out[1], out[2]};
}

// ---------------------------------------------------------------------
// Surface + material (wgsl:765-916).
// ---------------------------------------------------------------------

struct Surface {
    V3 position, ng, sn;
};

Surface surface(const Ctx &c, const Hit &h) {
    V3 p0 = tri_pos(c, h.slot, 0), p1 = tri_pos(c, h.slot, 1), p2 = tri_pos(c, h.slot, 2);
    V3 b = h.b;
    Surface s;
    s.position = p0 * b.x + p1 * b.y + p2 * b.z;
    V3 cr = cross(p1 - p0, p2 - p0);
    float len = std::sqrt(dot(cr, cr));
    s.ng = len > 0.0f ? cr / len : V3{0.0f, 0.0f, 1.0f};
    s.sn = s.ng;
    if (attr_flags(c, h.slot) & OXPT_F_NORMALS) {
        V3 n = nrm(interp3a(c, h.slot, 0, b));
        if (dot(n, n) > 0.5f) s.sn = n;
    }
    return s;
}

struct Mat {
    V4 base;
    float metallic, roughness;
    V3 normal;
    bool unlit;
    V3 emission;
    float ior, specular;
    V3 specular_color;
    float transmission, thickness;
    V3 atten_color;
    float atten_dist, clearcoat, cc_rough;
    V3 cc_normal, sheen;
    float sheen_rough;
};

V3 perturb(const Ctx &c, uint32_t slot, V3 b, V3 n, V4 t, float scale) {
    if ((attr_flags(c, slot) & OXPT_F_TANGENTS) == 0) return n;
    V4 tg = interp4a(c, slot, 3, b);
    float w = attr(c, slot, 3).w < 0.0f ? -1.0f : 1.0f;
    V3 t3 = xyz(tg);
    float d = dot(n, t3);
    V3 tt = nrm(t3 - n * d);
    if (dot(tt, tt) < 0.5f) return n;
    V3 bt = cross(n, tt) * w;
    V3 ts = {(t.x * 2.0f - 1.0f) * scale, (t.y * 2.0f - 1.0f) * scale, t.z * 2.0f - 1.0f};
    V3 nn = nrm(tt * ts.x + bt * ts.y + n * ts.z);
    if (dot(nn, nn) > 0.5f) return nn;
    return n;
}

Mat material(const Ctx &c, const Hit &h, const Surface &surf, const Lod &lod, bool flip) {
    V3 sn = flip ? -surf.sn : surf.sn;
    uint32_t m = tri_material(c, h.slot), o = mat_off(c, m), slot = h.slot;
    V3 b = h.b;
    V4 tex[13];
    uint32_t present = 0;
    for (uint32_t si = 0; si < 13; si++) {
        tex[si] = {1.0f, 1.0f, 1.0f, 1.0f};
        if (slot_present(c, m, si, slot)) {
            present |= 1u << si;
            Lod l = si == OXPT_S_EMISSIVE ? base_lod() : lod;
            tex[si] = slot_sample(c, m, si, slot, b, l);
        }
    }
    Mat r;
    V4 col = tv4(c, o) * tex[OXPT_S_BASE];
    if (attr_flags(c, slot) & OXPT_F_COLORS) col = col * interp4a(c, slot, 9, b);
    r.base = col;
    r.metallic = clampf(tf(c, o + 8) * tex[OXPT_S_MR].z, 0.0f, 1.0f);
    r.roughness = clampf(tf(c, o + 9) * tex[OXPT_S_MR].y, 0.0f, 1.0f);
    r.normal = sn;
    if (present & (1u << OXPT_S_NORMAL)) r.normal = perturb(c, slot, b, sn, tex[OXPT_S_NORMAL], tf(c, o + 10));
    r.unlit = (c.tables[o + OXPT_M_FLAGS] & 8u) != 0;
    r.emission = v3(0.0f);
    V3 e = tv3(c, o + 4);
    if (!r.unlit && !is_black(e)) r.emission = e * xyz(tex[OXPT_S_EMISSIVE]);
    r.ior = tf(c, o + 12);
    r.specular = tf(c, o + 13) * tex[OXPT_S_SPEC].w;
    r.specular_color = tv3(c, o + 14) * xyz(tex[OXPT_S_SPEC_COLOR]);
    r.transmission = tf(c, o + 17) * tex[OXPT_S_TRANSMISSION].x;
    r.thickness = tf(c, o + 18) * tex[OXPT_S_THICKNESS].y;
    r.atten_color = tv3(c, o + 19);
    r.atten_dist = tf(c, o + 22);
    r.clearcoat = tf(c, o + 23) * tex[OXPT_S_CC].x;
    r.cc_rough = tf(c, o + 24) * tex[OXPT_S_CC_ROUGH].y;
    r.cc_normal = sn;
    if (present & (1u << OXPT_S_CC_NORMAL)) r.cc_normal = perturb(c, slot, b, sn, tex[OXPT_S_CC_NORMAL], tf(c, o + 25));
    r.sheen = tv3(c, o + 26) * xyz(tex[OXPT_S_SHEEN_COLOR]);
    r.sheen_rough = tf(c, o + 29) * tex[OXPT_S_SHEEN_ROUGH].w;
    return r;
}

V3 emission(const Ctx &c, const Hit &h) {
    uint32_t m = tri_material(c, h.slot), o = mat_off(c, m);
    V3 e = tv3(c, o + 4);
    if ((c.tables[o + OXPT_M_FLAGS] & 8u) != 0 || is_black(e)) return v3(0.0f);
    if (slot_present(c, m, OXPT_S_EMISSIVE, h.slot))
        return e * xyz(slot_sample(c, m, OXPT_S_EMISSIVE, h.slot, h.b, base_lod()));
    return e;
}

// ---------------------------------------------------------------------
// BSDF (wgsl:922-1380).
// ---------------------------------------------------------------------

float d_ggx(float n_dot_h, float a) {
    if (n_dot_h <= 0.0f) return 0.0f;
    float a2 = a * a;
    float d = n_dot_h * n_dot_h * (a2 - 1.0f) + 1.0f;
    return a2 / (PI * d * d);
}
float lambda(float c_in, float a) {
    float c = fmax_(std::fabs(c_in), 1.0e-7f);
    float a2 = a * a;
    return (std::sqrt(a2 + (1.0f - a2) * c * c) / c - 1.0f) * 0.5f;
}
float g1(float c, float a) { return 1.0f / (1.0f + lambda(c, a)); }
float g2(float cv, float cl, float a) { return 1.0f / (1.0f + lambda(cv, a) + lambda(cl, a)); }
V3 schlick(V3 f0, float c) {
    float x = 1.0f - clampf(c, 0.0f, 1.0f);
    float k = x * x * x * x * x;
    return f0 + (v3(1.0f) - f0) * k;
}
V3 sample_vndf(V3 v, float a, float u0, float u1) {
    V3 vh = nrm({a * v.x, a * v.y, v.z});
    float lensq = vh.x * vh.x + vh.y * vh.y;
    V3 t1 = {1.0f, 0.0f, 0.0f};
    if (lensq > 0.0f) {
        float il = 1.0f / std::sqrt(lensq);
        t1 = {-vh.y * il, vh.x * il, 0.0f};
    }
    V3 t2 = cross(vh, t1);
    float r = std::sqrt(u0);
    float phi = 2.0f * PI * u1;
    float p1 = r * std::cos(phi);
    float p2 = r * std::sin(phi);
    float s = 0.5f * (1.0f + vh.z);
    p2 = (1.0f - s) * std::sqrt(fmax_(1.0f - p1 * p1, 0.0f)) + s * p2;
    float pz = std::sqrt(fmax_(1.0f - p1 * p1 - p2 * p2, 0.0f));
    V3 nh = t1 * p1 + t2 * p2 + vh * pz;
    return nrm({a * nh.x, a * nh.y, fmax_(nh.z, 1.0e-6f)});
}
float vndf_reflect_pdf(V3 n, V3 v, V3 h, float a) {
    float nv = dot(n, v), nh = dot(n, h);
    if (nv <= 0.0f || nh <= 0.0f || dot(v, h) <= 0.0f) return 0.0f;
    return g1(nv, a) * d_ggx(nh, a) / (4.0f * nv);
}
V3 reflect_h(V3 v, V3 h) { return 2.0f * dot(v, h) * h - v; }
float d_charlie(float n_dot_h, float a) {
    float inv = 1.0f / fmax_(a, 1.0e-3f);
    float sin2 = fmax_(1.0f - n_dot_h * n_dot_h, 0.0f);
    if (sin2 <= 0.0f) return 0.0f;
    return (2.0f + inv) * std::pow(sin2, inv * 0.5f) / (2.0f * PI);
}
float v_neubelt(float nl, float nv) { return 1.0f / fmax_(4.0f * (nl + nv - nl * nv), 1.0e-6f); }

float sheen_albedo(const float *lut, float cc, float rough) {
    const uint32_t SC = OXPT_SHEEN_COS, SR = OXPT_SHEEN_ROUGH;
    float x = clampf(clampf(cc, 0.0f, 1.0f) * (float)SC - 0.5f, 0.0f, (float)(SC - 1));
    float y = clampf(clampf(rough, 0.0f, 1.0f) * (float)(SR - 1), 0.0f, (float)(SR - 1));
    uint32_t x0 = (uint32_t)std::floor(x), y0 = (uint32_t)std::floor(y);
    uint32_t x1 = std::min(x0 + 1u, SC - 1u), y1 = std::min(y0 + 1u, SR - 1u);
    float fx = x - (float)x0, fy = y - (float)y0;
    float a00 = lut[y0 * SC + x0], a10 = lut[y0 * SC + x1], a01 = lut[y1 * SC + x0], a11 = lut[y1 * SC + x1];
    float a = a00 + (a10 - a00) * fx;
    float bb = a01 + (a11 - a01) * fx;
    return a + (bb - a) * fy;
}

V3 basis_t(V3 n) {
    float sign = sign_of(n.z);
    float a = -1.0f / (sign + n.z);
    float b = n.x * n.y * a;
    return {1.0f + sign * n.x * n.x * a, sign * b, -sign * n.x};
}
V3 basis_b(V3 n) {
    float sign = sign_of(n.z);
    float a = -1.0f / (sign + n.z);
    float b = n.x * n.y * a;
    return {b, sign + n.y * n.y * a, -n.y};
}
V3 cosine_hemisphere(float u0, float u1) {
    float a = 2.0f * u0 - 1.0f, b = 2.0f * u1 - 1.0f;
    float r = 0.0f, phi = 0.0f;
    if (a == 0.0f && b == 0.0f) {
        r = 0.0f;
        phi = 0.0f;
    } else if (std::fabs(a) > std::fabs(b)) {
        r = a;
        phi = (PI / 4.0f) * (b / a);
    } else {
        r = b;
        phi = PI / 2.0f - (PI / 4.0f) * (a / b);
    }
    float x = r * std::cos(phi), y = r * std::sin(phi);
    return {x, y, std::sqrt(fmax_(1.0f - x * x - y * y, 0.0f))};
}

struct Bsdf {
    V3 v, n, ng, ncc, base;
    float metallic, alpha, specular;
    V3 f0d;
    float transmission;
    bool thin;
    float eta_i, eta_t, clearcoat, cc_alpha;
    V3 sheen;
    float sheen_rough, sheen_alpha, w_cc;
    float p[4];
    float p4;
};

V3 fresnel_d(const Bsdf &bs, float c) {
    if (bs.eta_i > bs.eta_t) {
        float eta = bs.eta_i / bs.eta_t;
        float s2 = eta * eta * fmax_(1.0f - c * c, 0.0f);
        if (s2 >= 1.0f) return v3(1.0f);
        return schlick(bs.f0d, std::sqrt(1.0f - s2));
    }
    return schlick(bs.f0d, c);
}

Bsdf make_bsdf(const float *lut, const Mat &m, V3 v, V3 ng_in, bool front) {
    float fl = front ? 1.0f : -1.0f;
    V3 ng = ng_in * fl;
    V3 n = m.normal * fl;
    if (dot(n, v) <= 1.0e-4f) n = ng;
    V3 ncc = m.cc_normal * fl;
    if (dot(ncc, v) <= 1.0e-4f) ncc = n;
    bool thin = m.thickness <= 0.0f;
    float ior = 1.5f;
    if (finite(m.ior) && m.ior >= 1.0f) ior = m.ior;
    float eta_i = ior, eta_t = 1.0f;
    if (thin || front) {
        eta_i = 1.0f;
        eta_t = ior;
    }
    bool inside = !thin && !front;
    float q = (ior - 1.0f) / (ior + 1.0f);
    float r0 = q * q;
    Bsdf b;
    b.v = v;
    b.n = n;
    b.ng = ng;
    b.ncc = ncc;
    b.base = xyz(m.base);
    b.metallic = m.metallic;
    b.alpha = fmax_(m.roughness * m.roughness, MIN_ALPHA);
    b.specular = clampf(m.specular, 0.0f, 1.0f);
    b.f0d = vmin(r0 * m.specular_color, v3(1.0f));
    b.transmission = m.transmission;
    b.thin = thin;
    b.eta_i = eta_i;
    b.eta_t = eta_t;
    b.clearcoat = m.clearcoat;
    b.sheen = m.sheen;
    if (inside) {
        b.clearcoat = 0.0f;
        b.sheen = v3(0.0f);
    }
    b.cc_alpha = fmax_(m.cc_rough * m.cc_rough, MIN_ALPHA);
    b.sheen_rough = m.sheen_rough;
    b.sheen_alpha = fmax_(m.sheen_rough * m.sheen_rough, 1.0e-3f);
    float cv = fmax_(dot(n, v), 1.0e-4f);
    b.w_cc = 0.0f;
    if (b.clearcoat > 0.0f) b.w_cc = b.clearcoat * schlick(v3(0.04f), dot(ncc, v)).x;
    float sheen_max = max3(b.sheen);
    float e_v = 0.0f;
    if (sheen_max > 0.0f) e_v = sheen_albedo(lut, cv, b.sheen_rough);
    float s_scale = (1.0f - b.w_cc) * (1.0f - sheen_max * e_v);
    float fd = b.specular * max3(fresnel_d(b, cv));
    float fm = mean3(schlick(b.base, cv));
    float mm = b.metallic;
    float base_mean = fmax_(mean3(b.base), 0.0f);
    float p[4] = {s_scale * (1.0f - mm) * (1.0f - b.transmission) * (1.0f - fd) * base_mean,
                  s_scale * ((1.0f - mm) * fd + mm * fm),
                  s_scale * (1.0f - mm) * b.transmission * (1.0f - fd) * base_mean, b.w_cc};
    float p4 = (1.0f - b.w_cc) * sheen_max * e_v;
    float sum = p[0] + p[1] + p[2] + p[3] + p4;
    if (sum > 0.0f && finite(sum)) {
        for (int i = 0; i < 4; i++) p[i] = p[i] / sum;
        p4 = p4 / sum;
    } else {
        for (int i = 0; i < 4; i++) p[i] = 0.0f;
        p4 = 0.0f;
    }
    for (int i = 0; i < 4; i++) b.p[i] = p[i];
    b.p4 = p4;
    return b;
}

bool bsdf_black(const Bsdf &b) {
    return b.p[0] <= 0.0f && b.p[1] <= 0.0f && b.p[2] <= 0.0f && b.p[3] <= 0.0f && b.p4 <= 0.0f;
}
float lobe_p(const Bsdf &b, uint32_t i) { return i == 4 ? b.p4 : b.p[i]; }

struct Eval {
    V3 f;
    float pdf;
};

Eval bsdf_eval(const float *lut, const Bsdf &b, V3 l) {
    V3 v = b.v, n = b.n;
    float nv = fmax_(dot(n, v), 1.0e-4f);
    if (dot(b.ng, l) > 0.0f) {
        float nl = dot(n, l);
        V3 h = nrm(v + l);
        float vh = fmax_(dot(v, h), 0.0f);
        float nh = dot(n, h);
        V3 f = v3(0.0f);
        float pdf = 0.0f;
        if (nl > 0.0f) {
            V3 fd = fresnel_d(b, vh);
            V3 fm = schlick(b.base, vh);
            float spec_c = d_ggx(nh, b.alpha) * g2(nv, nl, b.alpha) / (4.0f * nv * nl);
            float m = b.metallic;
            float diff_k = (1.0f - m) * (1.0f - b.transmission) * (1.0f - b.specular * max3(fd)) / PI;
            V3 layer = spec_c * ((1.0f - m) * b.specular * fd + m * fm) + diff_k * b.base;
            float smax = max3(b.sheen);
            if (smax > 0.0f) {
                float scale = fmin_(1.0f - smax * sheen_albedo(lut, nv, b.sheen_rough),
                                    1.0f - smax * sheen_albedo(lut, nl, b.sheen_rough));
                float sh = d_charlie(nh, b.sheen_alpha) * v_neubelt(nl, nv);
                layer = layer * scale + b.sheen * sh;
            }
            f = layer * (1.0f - b.w_cc);
            pdf += (b.p[0] + b.p4) * nl / PI;
            pdf += b.p[1] * vndf_reflect_pdf(n, v, h, b.alpha);
        }
        if (b.clearcoat > 0.0f) {
            float cnv = fmax_(dot(b.ncc, v), 1.0e-4f);
            float cnl = dot(b.ncc, l);
            if (cnl > 0.0f) {
                float cnh = dot(b.ncc, h);
                float cv = b.w_cc * d_ggx(cnh, b.cc_alpha) * g2(cnv, cnl, b.cc_alpha) / (4.0f * cnv * cnl);
                f = f + v3(cv);
                pdf += b.p[3] * vndf_reflect_pdf(b.ncc, v, h, b.cc_alpha);
            }
        }
        return {f * fmax_(nl, 0.0f), pdf};
    }
    if (b.transmission <= 0.0f || b.metallic >= 1.0f) return {v3(0.0f), 0.0f};
    float nl = -dot(n, l);
    if (nl <= 0.0f) return {v3(0.0f), 0.0f};
    float smax = max3(b.sheen);
    float sheen_scale = 1.0f;
    if (smax > 0.0f) sheen_scale = 1.0f - smax * sheen_albedo(lut, nv, b.sheen_rough);
    float w = (1.0f - b.metallic) * b.transmission * (1.0f - b.w_cc) * sheen_scale;
    float val = 0.0f, pdf_t = 0.0f;
    if (b.thin) {
        V3 lm = l + n * (2.0f * nl);
        V3 h = nrm(v + lm);
        float vh = dot(v, h), nh = dot(n, h);
        if (vh <= 0.0f || nh <= 0.0f) return {v3(0.0f), 0.0f};
        float fd = b.specular * max3(fresnel_d(b, vh));
        val = d_ggx(nh, b.alpha) * g2(nv, nl, b.alpha) / (4.0f * nv * nl) * (1.0f - fd);
        pdf_t = vndf_reflect_pdf(n, v, h, b.alpha);
    } else {
        float ei = b.eta_i, et = b.eta_t;
        V3 h = nrm(-(ei * v + et * l));
        if (dot(h, n) < 0.0f) h = -h;
        float vh = dot(v, h), lh = dot(l, h), nh = dot(n, h);
        if (vh <= 0.0f || lh >= 0.0f || nh <= 0.0f) return {v3(0.0f), 0.0f};
        float denom = ei * vh + et * lh;
        if (std::fabs(denom) < 1.0e-6f) return {v3(0.0f), 0.0f};
        float d = d_ggx(nh, b.alpha);
        float fd = b.specular * max3(fresnel_d(b, vh));
        float jac = et * et * std::fabs(lh) / (denom * denom);
        val = vh * std::fabs(lh) * d * g2(nv, nl, b.alpha) * et * et * (1.0f - fd) / (nv * nl * denom * denom);
        pdf_t = g1(nv, b.alpha) * vh * d / nv * jac;
    }
    return {b.base * (w * val * nl), b.p[2] * pdf_t};
}

struct BSample {
    bool ok;
    V3 l, f;
    float pdf;
};

bool refract_h(V3 v, V3 h, float eta, V3 *out) {
    float ci = dot(v, h);
    float s2 = eta * eta * fmax_(1.0f - ci * ci, 0.0f);
    if (s2 >= 1.0f) return false;
    float ct = std::sqrt(1.0f - s2);
    float k = eta * ci - ct;
    *out = nrm(-eta * v + k * h);
    return true;
}

BSample bsdf_sample(const float *lut, const Bsdf &b, V3 u) {
    BSample res;
    res.ok = false;
    float acc = 0.0f;
    uint32_t lobe = NONE;
    for (uint32_t i = 0; i < 5; i++) {
        float p = lobe_p(b, i);
        acc += p;
        if (p > 0.0f && u.z < acc) {
            lobe = i;
            break;
        }
    }
    if (lobe == NONE) {
        for (uint32_t i = 0; i < 5; i++)
            if (lobe_p(b, i) > 0.0f) lobe = i;
        if (lobe == NONE) return res;
    }
    V3 v = b.v, l;
    if (lobe == 0 || lobe == 4) {
        V3 t = basis_t(b.n), bb = basis_b(b.n);
        V3 c = cosine_hemisphere(u.x, u.y);
        l = t * c.x + bb * c.y + b.n * c.z;
    } else {
        V3 nn = b.n;
        float a = b.alpha;
        if (lobe == 3) {
            nn = b.ncc;
            a = b.cc_alpha;
        }
        V3 t = basis_t(nn), bb = basis_b(nn);
        V3 vl = {dot(v, t), dot(v, bb), dot(v, nn)};
        if (vl.z <= 0.0f) return res;
        V3 hl = sample_vndf(vl, a, u.x, u.y);
        V3 h = t * hl.x + bb * hl.y + nn * hl.z;
        if (lobe != 2) {
            l = reflect_h(v, h);
        } else if (b.thin) {
            V3 r = reflect_h(v, h);
            float d = dot(r, b.n);
            l = r + b.n * (-2.0f * d);
        } else {
            V3 rl;
            if (!refract_h(v, h, b.eta_i / b.eta_t, &rl)) return res;
            l = rl;
        }
    }
    if (!finite3(l) || dot(l, l) < 0.5f) return res;
    bool reflection = dot(b.ng, l) > 0.0f;
    if (reflection == (lobe == 2)) return res;
    Eval e = bsdf_eval(lut, b, l);
    if (e.pdf <= 0.0f || !finite(e.pdf) || is_black(e.f) || !finite3(e.f)) return res;
    res.ok = true;
    res.l = l;
    res.f = e.f;
    res.pdf = e.pdf;
    return res;
}

// ---------------------------------------------------------------------
// Lights (wgsl:1386-1654).
// ---------------------------------------------------------------------

struct LightSample {
    bool ok;
    V3 l;
    float distance;
    V3 radiance;
};

LightSample punctual_sample(const Ctx &c, uint32_t j, V3 p) {
    uint32_t o = c.P.offs0[3] + 16u * j;
    c.add(64);
    LightSample r;
    r.ok = false;
    uint32_t kind = c.tables[o + 3];
    V3 base = tv3(c, o + 8), dir = tv3(c, o + 4);
    if (kind == 0) {
        r.ok = true;
        r.l = -dir;
        r.distance = -1.0f;
        r.radiance = base;
        return r;
    }
    V3 to = tv3(c, o) - p;
    float d2 = dot(to, to);
    if (!(d2 > 1.0e-12f)) return r;
    float d = std::sqrt(d2);
    V3 l = to / d;
    float att = 1.0f / d2;
    float range = tf(c, o + 7);
    if (range > 0.0f) {
        float x = d / range;
        att *= clampf(1.0f - x * x * x * x, 0.0f, 1.0f);
    }
    if (kind == 2) {
        float cd = dot(dir, -l);
        float a = clampf(cd * tf(c, o + 12) + tf(c, o + 13), 0.0f, 1.0f);
        att *= a * a;
    }
    if (att <= 0.0f) return r;
    r.ok = true;
    r.l = l;
    r.distance = d;
    r.radiance = base * att;
    return r;
}

uint32_t partition_point(const Ctx &c, uint32_t o, uint32_t n, float u) {
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        uint32_t mid = (lo + hi) / 2u;
        c.add(4);
        if (tf(c, o + mid) <= u) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

inline float elight_pmf(const Ctx &c, uint32_t li) { return tf(c, c.P.offs1[0] + 4u * li + 1); }
inline uint32_t elight_slot(const Ctx &c, uint32_t li) { return c.tables[c.P.offs1[0] + 4u * li + 3]; }

uint32_t pick_elight(const Ctx &c, float u) {
    uint32_t n = c.P.cfg[3];
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        uint32_t mid = (lo + hi) / 2u;
        c.add(16);
        if (tf(c, c.P.offs1[0] + 4u * mid) <= u) lo = mid + 1;
        else hi = mid;
    }
    return std::min(lo, n - 1u);
}

V3 nrm_any(V3 v) {
    float l = std::sqrt(dot(v, v));
    if (l > 0.0f) return v / l;
    return v3(0.0f);
}

float triangle_solid_angle_f32(V3 p, V3 a0, V3 b0, V3 c0) {
    V3 a = nrm_any(a0 - p), b = nrm_any(b0 - p), c = nrm_any(c0 - p);
    float num = std::fabs(dot(a, cross(b, c)));
    float den = 1.0f + dot(a, b) + dot(b, c) + dot(c, a);
    float o = 2.0f * std::atan2(num, den);
    if (finite(o)) return fmax_(o, 0.0f);
    return 0.0f;
}

V3 sample_spherical_triangle_f32(V3 p, V3 a0, V3 b0, V3 c0, float u0, float u1) {
    V3 zero = v3(0.0f);
    V3 a = nrm_any(a0 - p), b = nrm_any(b0 - p), c = nrm_any(c0 - p);
    V3 nab = nrm_any(cross(a, b)), nbc = nrm_any(cross(b, c)), nca = nrm_any(cross(c, a));
    if (dot(a, a) == 0.0f || dot(b, b) == 0.0f || dot(c, c) == 0.0f || dot(nab, nab) == 0.0f ||
        dot(nbc, nbc) == 0.0f || dot(nca, nca) == 0.0f)
        return zero;
    float alpha = std::acos(clampf(-dot(nab, nca), -1.0f, 1.0f));
    float beta = std::acos(clampf(-dot(nbc, nab), -1.0f, 1.0f));
    float gamma = std::acos(clampf(-dot(nca, nbc), -1.0f, 1.0f));
    float area = alpha + beta + gamma - PI;
    if (!(area > 0.0f)) return zero;
    float ap = u0 * area;
    float s = std::sin(ap - alpha), t = std::cos(ap - alpha);
    float sa = std::sin(alpha), ca = std::cos(alpha);
    float cos_c = dot(a, b);
    float uu = t - ca, vv = s + sa * cos_c;
    float den = (vv * s + uu * t) * sa;
    if (std::fabs(den) < 1.0e-30f) return zero;
    float q = clampf(((vv * t - uu * s) * ca - vv) / den, -1.0f, 1.0f);
    V3 ca_perp = nrm_any(c - dot(c, a) * a);
    if (dot(ca_perp, ca_perp) == 0.0f) return zero;
    float r = std::sqrt(fmax_(1.0f - q * q, 0.0f));
    V3 cp = q * a + r * ca_perp;
    float z = 1.0f - u1 * (1.0f - dot(cp, b));
    V3 cb_perp = nrm_any(cp - dot(cp, b) * b);
    float w = std::sqrt(fmax_(1.0f - z * z, 0.0f));
    V3 d = b;
    if (dot(cb_perp, cb_perp) > 0.0f) d = z * b + w * cb_perp;
    return nrm_any(d);
}

// f64 twins: the reference's CPU tracer evaluates these in double
// (README.md:117-119); same steps, double arithmetic.
struct D3 {
    double x, y, z;
};
inline D3 dd(V3 a) { return {a.x, a.y, a.z}; }
inline D3 operator-(D3 a, D3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline D3 operator+(D3 a, D3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline D3 operator*(D3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
inline D3 operator*(double s, D3 a) { return {a.x * s, a.y * s, a.z * s}; }
inline double dotd(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline D3 crossd(D3 a, D3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline D3 nrm_anyd(D3 v) {
    double l = std::sqrt(dotd(v, v));
    if (l > 0.0) return {v.x / l, v.y / l, v.z / l};
    return {0, 0, 0};
}
inline double clampd(double x, double lo, double hi) { return std::fmin(std::fmax(x, lo), hi); }

float triangle_solid_angle_f64(V3 pf, V3 a0, V3 b0, V3 c0) {
    D3 p = dd(pf);
    D3 a = nrm_anyd(dd(a0) - p), b = nrm_anyd(dd(b0) - p), c = nrm_anyd(dd(c0) - p);
    double num = std::fabs(dotd(a, crossd(b, c)));
    double den = 1.0 + dotd(a, b) + dotd(b, c) + dotd(c, a);
    double o = 2.0 * std::atan2(num, den);
    if (std::isfinite(o)) return (float)std::fmax(o, 0.0);
    return 0.0f;
}

V3 sample_spherical_triangle_f64(V3 pf, V3 a0, V3 b0, V3 c0, float u0, float u1) {
    V3 zero = v3(0.0f);
    D3 p = dd(pf);
    D3 a = nrm_anyd(dd(a0) - p), b = nrm_anyd(dd(b0) - p), c = nrm_anyd(dd(c0) - p);
    D3 nab = nrm_anyd(crossd(a, b)), nbc = nrm_anyd(crossd(b, c)), nca = nrm_anyd(crossd(c, a));
    if (dotd(a, a) == 0.0 || dotd(b, b) == 0.0 || dotd(c, c) == 0.0 || dotd(nab, nab) == 0.0 ||
        dotd(nbc, nbc) == 0.0 || dotd(nca, nca) == 0.0)
        return zero;
    const double PId = 3.14159265358979323846;
    double alpha = std::acos(clampd(-dotd(nab, nca), -1.0, 1.0));
    double beta = std::acos(clampd(-dotd(nbc, nab), -1.0, 1.0));
    double gamma = std::acos(clampd(-dotd(nca, nbc), -1.0, 1.0));
    double area = alpha + beta + gamma - PId;
    if (!(area > 0.0)) return zero;
    double ap = (double)u0 * area;
    double s = std::sin(ap - alpha), t = std::cos(ap - alpha);
    double sa = std::sin(alpha), ca = std::cos(alpha);
    double cos_c = dotd(a, b);
    double uu = t - ca, vv = s + sa * cos_c;
    double den = (vv * s + uu * t) * sa;
    if (std::fabs(den) < 1.0e-300) return zero;
    double q = clampd(((vv * t - uu * s) * ca - vv) / den, -1.0, 1.0);
    D3 ca_perp = nrm_anyd(c - dotd(c, a) * a);
    if (dotd(ca_perp, ca_perp) == 0.0) return zero;
    double r = std::sqrt(std::fmax(1.0 - q * q, 0.0));
    D3 cp = q * a + r * ca_perp;
    double z = 1.0 - (double)u1 * (1.0 - dotd(cp, b));
    D3 cb_perp = nrm_anyd(cp - dotd(cp, b) * b);
    double w = std::sqrt(std::fmax(1.0 - z * z, 0.0));
    D3 d = b;
    if (dotd(cb_perp, cb_perp) > 0.0) d = z * b + w * cb_perp;
    D3 n = nrm_anyd(d);
    return {(float)n.x, (float)n.y, (float)n.z};
}

inline float triangle_solid_angle(const Ctx &c, V3 p, V3 a, V3 b, V3 cc) {
    return c.f64_lights ? triangle_solid_angle_f64(p, a, b, cc) : triangle_solid_angle_f32(p, a, b, cc);
}
inline V3 sample_spherical_triangle(const Ctx &c, V3 p, V3 a, V3 b, V3 cc, float u0, float u1) {
    return c.f64_lights ? sample_spherical_triangle_f64(p, a, b, cc, u0, u1)
                        : sample_spherical_triangle_f32(p, a, b, cc, u0, u1);
}

inline float p_env(const Ctx &c) { return c.P.fparams[2]; }

float tri_pdf(const Ctx &c, uint32_t li, uint32_t slot, V3 p, V3 q) {
    V3 a = tri_pos(c, slot, 0), b = tri_pos(c, slot, 1), cc = tri_pos(c, slot, 2);
    float sel = (1.0f - p_env(c)) * elight_pmf(c, li);
    float omega = triangle_solid_angle(c, p, a, b, cc);
    if (omega >= MIN_SPHERICAL_SOLID_ANGLE) return sel / omega;
    V3 cr = cross(b - a, cc - a);
    float len = std::sqrt(dot(cr, cr));
    V3 to = q - p;
    float d2 = dot(to, to);
    if (len <= 0.0f || d2 <= 0.0f) return 0.0f;
    float cs = std::fabs(dot(cr, to) / (len * std::sqrt(d2)));
    if (cs <= 0.0f) return 0.0f;
    return sel * d2 / (0.5f * len * cs);
}

struct Texel {
    uint32_t x, y;
    float st;
};

Texel env_texel(const Ctx &c, V3 d) {
    uint32_t w = c.P.env[0], h = c.P.env[1];
    float u = 0.5f + std::atan2(d.x, -d.z) / (2.0f * PI);
    float th = std::acos(clampf(d.y, -1.0f, 1.0f));
    float v = th / PI;
    uint32_t x = std::min((uint32_t)fmax_(u * (float)w, 0.0f), w - 1u);
    uint32_t y = std::min((uint32_t)fmax_(v * (float)h, 0.0f), h - 1u);
    return {x, y, std::sin(th)};
}
V3 env_radiance(const Ctx &c, V3 d) {
    Texel t = env_texel(c, d);
    c.add(12);
    return tv3(c, c.P.offs2[1] + 3u * (t.y * c.P.env[0] + t.x));
}
float env_pdf(const Ctx &c, V3 d) {
    Texel t = env_texel(c, d);
    if (t.st <= 1.0e-6f) return 0.0f;
    uint32_t w = c.P.env[0], h = c.P.env[1];
    c.add(4);
    return tf(c, c.P.offs2[0] + t.y * w + t.x) * (float)(w * h) / (2.0f * PI * PI * t.st);
}

struct EnvSample {
    bool ok;
    V3 d, le;
    float pdf;
};

EnvSample env_sample(const Ctx &c, float u0, float u1) {
    EnvSample r;
    r.ok = false;
    uint32_t w = c.P.env[0], h = c.P.env[1];
    uint32_t marg = c.P.offs1[2];
    uint32_t y = std::min(std::max(partition_point(c, marg, h + 1u, u0), 1u), h) - 1u;
    uint32_t row = c.P.offs1[3] + y * (w + 1u);
    uint32_t x = std::min(std::max(partition_point(c, row, w + 1u, u1), 1u), w) - 1u;
    c.add(16);
    float my = (u0 - tf(c, marg + y)) / fmax_(tf(c, marg + y + 1) - tf(c, marg + y), 1.0e-12f);
    float mx = (u1 - tf(c, row + x)) / fmax_(tf(c, row + x + 1) - tf(c, row + x), 1.0e-12f);
    float u = ((float)x + clampf(mx, 0.0f, 0.99999f)) / (float)w;
    float v = ((float)y + clampf(my, 0.0f, 0.99999f)) / (float)h;
    float phi = 2.0f * PI * (u - 0.5f);
    float th = PI * v;
    float st = std::sin(th), ct = std::cos(th);
    V3 d = {st * std::sin(phi), ct, -st * std::cos(phi)};
    float pdf = env_pdf(c, d);
    if (pdf <= 0.0f || !finite(pdf)) return r;
    r.ok = true;
    r.d = d;
    r.le = env_radiance(c, d);
    r.pdf = pdf;
    return r;
}

// ---------------------------------------------------------------------
// Integrator (wgsl:1660-2045).
// ---------------------------------------------------------------------

float power_heuristic(float a, float b) {
    float a2 = a * a, b2 = b * b;
    if (a2 + b2 > 0.0f && finite(a2 + b2)) return a2 / (a2 + b2);
    if (!finite(a) && a > 0.0f) return 1.0f;
    return 0.0f;
}
V3 clamp_c(const Ctx &c, V3 col) {
    float m = max3(col);
    float cl = c.P.fparams[1];
    if (cl > 0.0f && m > cl) return col * (cl / m);
    return col;
}
V3 spawn(V3 p, V3 ng, V3 dir) {
    if (dot(dir, ng) >= 0.0f) return offset_ray_origin(p, ng);
    return offset_ray_origin(p, -ng);
}
float light_pdf(const Ctx &c, const Hit &h, V3 p, V3 q) {
    uint32_t li = tri_light(c, h.slot);
    if (li == NONE) return 0.0f;
    return tri_pdf(c, li, h.slot, p, q);
}

struct PathResult {
    bool covered;
    V3 radiance;
};

inline V4 pattern(const Ctx &c, const uint32_t *sob, const Sampler &s, uint32_t p) {
    uint32_t w[4];
    sobol_owen_words(sob, s.index, hash_mix(s.pixel_seed, p), w);
    (void)c;
    return {to_unit(w[0]), to_unit(w[1]), to_unit(w[2]), to_unit(w[3])};
}

PathResult trace_sample(const Ctx &c, const uint32_t *sob, const float *lut, uint32_t x, uint32_t y,
                        uint32_t index) {
    const oxpt_params &P = c.P;
    Sampler s = make_sampler(P.dims[2], x, y, index);
    V4 p0 = pattern(c, sob, s, 0);
    float width = (float)P.dims[0], height = (float)P.dims[1];
    float ndc_x = (((float)x + p0.x) / fmax_(width, 1.0f)) * 2.0f - 1.0f;
    float ndc_y = 1.0f - (((float)y + p0.y) / fmax_(height, 1.0f)) * 2.0f;
    bool orthographic = P.eye[3] > 0.5f;
    float half_w = P.forward[3], half_h = P.side[3];
    V3 eye = {P.eye[0], P.eye[1], P.eye[2]};
    V3 fwd = {P.forward[0], P.forward[1], P.forward[2]};
    V3 side = {P.side[0], P.side[1], P.side[2]};
    V3 up = {P.up[0], P.up[1], P.up[2]};
    V3 o, d;
    if (orthographic) {
        o = eye + (side * (ndc_x * half_w) + up * (ndc_y * half_h));
        d = fwd;
    } else {
        o = eye;
        d = nrm(side * (ndc_x * half_w) + up * (ndc_y * half_h) + fwd);
    }
    V3 beta = v3(1.0f), radiance = v3(0.0f);
    float prev_pdf = 0.0f;
    V3 prev_pos = v3(0.0f);
    bool in_medium = false;
    V3 sigma = v3(0.0f);
    uint32_t strategy = P.cfg[1];
    bool has_env = P.env[2] != 0;
    uint32_t n_elights = P.cfg[3], n_punct = P.cfg[2];
    uint32_t k = 0;

    bool shadow = false;
    V3 r_o = o, r_d = d;
    float r_tmax = INF;
    uint32_t r_id = 0, r_skip = NONE;
    V3 pending = v3(0.0f);
    Bsdf bsdf{};
    V3 p = v3(0.0f), ng = v3(0.0f);
    float thickness = 0.0f, transmission = 0.0f;
    V3 atten_color = v3(1.0f);
    float atten_dist = -1.0f;
    bool front = true;
    V4 pl_u = {0, 0, 0, 0};
    uint32_t job = 0;

    for (;;) {
        Hit hit = traverse(c, make_ray(r_o, r_d), r_tmax, s, r_id, !shadow && k == 0, shadow, r_skip);
        if (shadow) {
            if (hit.slot == NONE) radiance = radiance + pending;
            job += 1;
        } else {
            if (in_medium) {
                if (hit.slot == NONE) break;
                V3 e = -sigma * hit.t;
                beta = beta * V3{std::exp(e.x), std::exp(e.y), std::exp(e.z)};
            }
            if (hit.slot == NONE) {
                if (k == 0) return {false, v3(0.0f)};
                V3 le = v3(P.fparams[0]);
                float w = 1.0f;
                if (has_env) {
                    le = env_radiance(c, d);
                    if (strategy == 0) w = power_heuristic(prev_pdf, p_env(c) * env_pdf(c, d));
                    else if (strategy == 1) w = 0.0f;
                }
                if (w > 0.0f) radiance = radiance + clamp_c(c, beta * le * w);
                break;
            }
            if (c.ctr) c.ctr->hits++;
            c.add(192 + 544 + 48);
            Surface surf = surface(c, hit);
            uint32_t mflags = mat_flags(c, tri_material(c, hit.slot));
            bool double_sided = (mflags & 4u) != 0;
            Lod lod = base_lod();
            if (k == 0) {
                float cw = P.up[3];
                if (!orthographic) cw = P.up[3] * hit.t;
                lod = {cw, d};
            }
            bool flip = !hit.front && double_sided;
            Mat mat = material(c, hit, surf, lod, flip);
            if (flip) {
                mat.normal = -mat.normal;
                mat.cc_normal = -mat.cc_normal;
            }
            if ((hit.front || double_sided) && !is_black(mat.emission)) {
                V3 le = mat.emission;
                if (k == 0) {
                    radiance = radiance + beta * le;
                } else {
                    float pl = light_pdf(c, hit, prev_pos, surf.position);
                    float w = 1.0f;
                    if (pl > 0.0f) {
                        if (strategy == 0) w = power_heuristic(prev_pdf, pl);
                        else if (strategy == 1) w = 0.0f;
                    }
                    if (w > 0.0f) radiance = radiance + clamp_c(c, beta * le * w);
                }
            }
            if (mat.unlit) {
                V3 contrib = beta * xyz(mat.base);
                radiance = radiance + (k == 0 ? contrib : clamp_c(c, contrib));
                break;
            }
            if (k >= P.dims[3]) break;
            bsdf = make_bsdf(lut, mat, -d, surf.ng, hit.front);
            if (bsdf_black(bsdf)) break;
            p = surf.position;
            ng = bsdf.ng;
            thickness = mat.thickness;
            transmission = mat.transmission;
            atten_color = mat.atten_color;
            atten_dist = mat.atten_dist;
            front = hit.front;
            pl_u = pattern(c, sob, s, 2u + 2u * k);
            job = 0;
        }

        bool armed = false;
        uint32_t ray_id = k << 16u;
        while (job <= n_punct) {
            bool ok = false;
            V3 l = v3(0.0f);
            float tmax = INF;
            uint32_t rid = ray_id | 0xffffu;
            uint32_t skip = NONE;
            V3 le = v3(0.0f);
            float pl = -1.0f;
            if (job < n_punct) {
                LightSample ls = punctual_sample(c, job, p);
                if (ls.ok) {
                    ok = true;
                    l = ls.l;
                    if (ls.distance >= 0.0f) tmax = ls.distance * (1.0f - 1.0e-4f);
                    rid = ray_id | 0x8000u | (job & 0x7fffu);
                    le = ls.radiance;
                }
            } else if (strategy != 2 && (has_env || n_elights > 0)) {
                float pe = p_env(c);
                if (pl_u.z < pe) {
                    if (has_env) {
                        EnvSample es = env_sample(c, pl_u.x, pl_u.y);
                        if (es.ok && pe * es.pdf > 0.0f) {
                            ok = true;
                            l = es.d;
                            le = es.le;
                            pl = pe * es.pdf;
                        }
                    }
                } else if (n_elights > 0) {
                    float us = pl_u.z;
                    if (pe > 0.0f) us = clampf((pl_u.z - pe) / (1.0f - pe), 0.0f, 0.999999f);
                    uint32_t li = pick_elight(c, us);
                    uint32_t lslot = elight_slot(c, li);
                    c.add(16 + 48 + 544);
                    V3 a = tri_pos(c, lslot, 0), b = tri_pos(c, lslot, 1), cc = tri_pos(c, lslot, 2);
                    bool has_target = false;
                    V3 q = v3(0.0f);
                    if (triangle_solid_angle(c, p, a, b, cc) >= MIN_SPHERICAL_SOLID_ANGLE) {
                        V3 sl = sample_spherical_triangle(c, p, a, b, cc, pl_u.x, pl_u.y);
                        if (dot(sl, sl) > 0.0f) {
                            V3 n = cross(b - a, cc - a);
                            float dn = dot(sl, n);
                            if (dn != 0.0f) {
                                float t = dot(a - p, n) / dn;
                                if (t > 0.0f && finite(t)) {
                                    q = p + sl * t;
                                    has_target = true;
                                }
                            }
                        }
                    } else {
                        float su = std::sqrt(pl_u.x);
                        V3 w3 = {1.0f - su, su * (1.0f - pl_u.y), su * pl_u.y};
                        q = a * w3.x + b * w3.y + cc * w3.z;
                        has_target = true;
                    }
                    if (has_target) {
                        V3 e1 = b - a, e2 = cc - a, rr = q - a;
                        float d11 = dot(e1, e1), d12 = dot(e1, e2), d22 = dot(e2, e2);
                        float r1 = dot(rr, e1), r2 = dot(rr, e2);
                        float det = d11 * d22 - d12 * d12;
                        V3 to = q - p;
                        float dist2 = dot(to, to);
                        if (finite(det) && std::fabs(det) > 1.17549435e-38f && dist2 > 1.0e-12f) {
                            float bu = (d22 * r1 - d12 * r2) / det;
                            float bv = (d11 * r2 - d12 * r1) / det;
                            Hit lh;
                            lh.slot = lslot;
                            lh.t = std::sqrt(dist2);
                            V3 bb = {1.0f - bu - bv, bu, bv};
                            lh.b = vmin(vmax(bb, v3(0.0f)), v3(1.0f));
                            lh.front = true;
                            V3 lng = surface(c, lh).ng;
                            bool lds = mat_double_sided(c, tri_material(c, lslot));
                            float dist = std::sqrt(dist2);
                            V3 ld = to / dist;
                            float cos_l = -dot(lng, ld);
                            float tp = tri_pdf(c, li, lslot, p, q);
                            if ((cos_l > 0.0f || (lds && cos_l < 0.0f)) && tp > 0.0f && finite(tp)) {
                                ok = true;
                                l = ld;
                                le = emission(c, lh);
                                pl = tp;
                                tmax = fmax_(dist * (1.0f - 1.0e-4f), 0.0f);
                                skip = tri_global(c, lslot);
                            }
                        }
                    }
                }
            }
            if (ok) {
                Eval e = bsdf_eval(lut, bsdf, l);
                if (!is_black(e.f) && !is_black(le)) {
                    if (pl < 0.0f) {
                        pending = clamp_c(c, beta * e.f * le);
                    } else {
                        float w = 1.0f;
                        if (strategy == 0) w = power_heuristic(pl, e.pdf);
                        pending = clamp_c(c, beta * e.f * le * (w / pl));
                    }
                    r_o = spawn(p, ng, l);
                    r_d = l;
                    r_tmax = tmax;
                    r_id = rid;
                    r_skip = skip;
                    armed = true;
                    break;
                }
            }
            job += 1;
        }
        if (armed) {
            shadow = true;
            continue;
        }

        V4 pb_u = pattern(c, sob, s, 1u + 2u * k);
        BSample bs = bsdf_sample(lut, bsdf, {pb_u.x, pb_u.y, pb_u.z});
        if (!bs.ok) break;
        beta = beta * (bs.f * (1.0f / bs.pdf));
        if (!finite3(beta)) break;
        prev_pdf = bs.pdf;
        prev_pos = p;
        if (dot(bs.l, ng) < 0.0f && thickness > 0.0f && transmission > 0.0f) {
            if (front) {
                if (atten_dist >= 0.0f) {
                    V3 lc = vmax(atten_color, v3(1.0e-6f));
                    sigma = V3{-std::log(lc.x), -std::log(lc.y), -std::log(lc.z)} / atten_dist;
                    in_medium = true;
                } else {
                    in_medium = false;
                }
            } else {
                in_medium = false;
            }
        }
        if (k + 1u >= P.cfg[0]) {
            float q = clampf(max3(beta), 0.05f, 1.0f);
            if (q < 1.0f) {
                if (pb_u.w >= q) break;
                beta = beta * (1.0f / q);
            }
        }
        k += 1;
        d = bs.l;
        shadow = false;
        r_o = spawn(p, ng, bs.l);
        r_d = d;
        r_tmax = INF;
        r_id = k << 16u;
        r_skip = NONE;
    }
    if (finite3(radiance)) return {true, radiance};
    return {true, v3(0.0f)};
}

Ctx make_ctx(const oxo_scene *s, const oxpt_params *P, int flags, oxo_counts *ctr) {
    Ctx c;
    c.geom = s->geom;
    c.attrs = s->attrs;
    c.tables = s->tables;
    c.atlas = s->atlas;
    c.atlas_side = s->atlas_side;
    c.P = *P;
    c.f64_lights = (flags & OXO_F64_LIGHTS) != 0;
    c.ctr = ctr;
    return c;
}

// sRGB transfer + tone maps (resolve.wgsl:38-75).
float srgb_to_linear(float c) {
    if (c <= 0.04045f) return c / 12.92f;
    return std::pow((c + 0.055f) / 1.055f, 2.4f);
}
float linear_to_srgb(float c) {
    float x = clampf(c, 0.0f, 1.0f);
    if (x <= 0.0031308f) return 12.92f * x;
    return 1.055f * std::pow(x, 1.0f / 2.4f) - 0.055f;
}
float aces(float x) {
    float y = x * 0.6f;
    return clampf((y * (2.51f * y + 0.03f)) / (y * (2.43f * y + 0.59f) + 0.14f), 0.0f, 1.0f);
}
V3 tone_map(uint32_t op, V3 rgb_in) {
    V3 rgb = vmax(rgb_in, v3(0.0f));
    if (op == 1) {
        float l = dot(rgb, V3{0.2126f, 0.7152f, 0.0722f});
        if (l <= 0.0f) return v3(0.0f);
        return vmin(rgb / (1.0f + l), v3(1.0f));
    }
    if (op == 2) return {aces(rgb.x), aces(rgb.y), aces(rgb.z)};
    return vmin(rgb, v3(1.0f));
}
// Unorm8 store: round to nearest (the render target's conversion).
inline uint8_t unorm8(float v) {
    float x = clampf(v, 0.0f, 1.0f) * 255.0f + 0.5f;
    return (uint8_t)std::floor(x);
}

}  // namespace

// ---------------------------------------------------------------------
// C ABI.
// ---------------------------------------------------------------------

extern "C" {

void oxo_sobol_directions(uint32_t out[OXPT_SOBOL_WORDS]) {
    // Dimension 0: van der Corput (bit-reversed identity). Dimensions 1-3:
    // Joe & Kuo (2008) primitive polynomials x+1, x^2+x+1, x^3+x+1 with
    // initial m = {1}, {1,3}, {1,3,1}; recurrence
    // v_k = v_{k-s} ^ (v_{k-s} >> s) ^ XOR_j a_j v_{k-j}.
    for (int k = 0; k < 32; k++) out[k] = 1u << (31 - k);
    struct Poly {
        int s;
        uint32_t a;
        uint32_t m[3];
    } polys[3] = {{1, 0, {1, 0, 0}}, {2, 1, {1, 3, 0}}, {3, 1, {1, 3, 1}}};
    for (int d = 0; d < 3; d++) {
        const Poly &p = polys[d];
        uint32_t *v = out + 32 * (d + 1);
        for (int k = 0; k < p.s; k++) v[k] = p.m[k] << (31 - k);
        for (int k = p.s; k < 32; k++) {
            uint32_t x = v[k - p.s] ^ (v[k - p.s] >> p.s);
            for (int j = 1; j < p.s; j++)
                if ((p.a >> (p.s - 1 - j)) & 1u) x ^= v[k - j];
            v[k] = x;
        }
    }
}

void oxo_sheen_table(float out[OXPT_SHEEN_WORDS], int n) {
    // E(mu, r) = integral over the hemisphere of D_charlie(n.h; max(r^2,1e-3))
    // * V_neubelt(n.l, n.v) * (n.l) dw_l, by the midpoint rule in
    // (cos theta_l, phi_l); mu = (i + 0.5)/32, r = j/15.
    if (n <= 0) n = 256;
    const double PId = 3.14159265358979323846;
    for (uint32_t j = 0; j < OXPT_SHEEN_ROUGH; j++) {
        double r = (double)j / (double)(OXPT_SHEEN_ROUGH - 1);
        double a = std::fmax(r * r, 1.0e-3);
        double inv = 1.0 / std::fmax(a, 1.0e-3);
        for (uint32_t i = 0; i < OXPT_SHEEN_COS; i++) {
            double mu = ((double)i + 0.5) / (double)OXPT_SHEEN_COS;
            double vx = std::sqrt(std::fmax(1.0 - mu * mu, 0.0)), vz = mu;
            double sum = 0.0;
            for (int a_i = 0; a_i < n; a_i++) {
                double cl = ((double)a_i + 0.5) / (double)n;
                double sl = std::sqrt(std::fmax(1.0 - cl * cl, 0.0));
                for (int p_i = 0; p_i < 2 * n; p_i++) {
                    double phi = 2.0 * PId * ((double)p_i + 0.5) / (double)(2 * n);
                    double lx = sl * std::cos(phi), ly = sl * std::sin(phi), lz = cl;
                    double hx = vx + lx, hy = ly, hz = vz + lz;
                    double hl = std::sqrt(hx * hx + hy * hy + hz * hz);
                    double nh = hz / hl;
                    double sin2 = std::fmax(1.0 - nh * nh, 0.0);
                    double D = sin2 <= 0.0 ? 0.0 : (2.0 + inv) * std::pow(sin2, inv * 0.5) / (2.0 * PId);
                    double V = 1.0 / std::fmax(4.0 * (cl + mu - cl * mu), 1.0e-6);
                    sum += D * V * cl;
                }
            }
            double dw = (1.0 / n) * (2.0 * PId / (2 * n));
            out[j * OXPT_SHEEN_COS + i] = (float)(sum * dw);
        }
    }
}

uint32_t oxo_pcg_hash(uint32_t x) { return pcg_hash(x); }
uint32_t oxo_hash_mix(uint32_t a, uint32_t b) { return hash_mix(a, b); }
uint32_t oxo_nested_uniform_scramble(uint32_t x, uint32_t seed) { return nested_uniform_scramble(x, seed); }

void oxo_pattern_words(const uint32_t sob[OXPT_SOBOL_WORDS], uint32_t seed, uint32_t x, uint32_t y,
                       uint32_t index, uint32_t p, uint32_t out[4]) {
    Sampler s = make_sampler(seed, x, y, index);
    sobol_owen_words(sob, s.index, hash_mix(s.pixel_seed, p), out);
}

void oxo_pattern_words_batch(const uint32_t sob[OXPT_SOBOL_WORDS], uint32_t seed, size_t n,
                             const uint32_t *x, const uint32_t *y, const uint32_t *index,
                             const uint32_t *p, uint32_t *out4) {
    for (size_t i = 0; i < n; i++) oxo_pattern_words(sob, seed, x[i], y[i], index[i], p[i], out4 + 4 * i);
}

void oxo_light_picks(const oxo_scene *s, const oxpt_params *P, size_t n, const float *u, uint32_t *out3) {
    Ctx c = make_ctx(s, P, 0, nullptr);
    uint32_t w = P->env[0], h = P->env[1];
    for (size_t i = 0; i < n; i++) {
        out3[3 * i] = P->cfg[3] > 0 ? pick_elight(c, u[i]) : NONE;
        if (P->env[2]) {
            uint32_t y = std::min(std::max(partition_point(c, P->offs1[2], h + 1u, u[i]), 1u), h) - 1u;
            uint32_t row = P->offs1[3] + y * (w + 1u);
            uint32_t x = std::min(std::max(partition_point(c, row, w + 1u, u[i]), 1u), w) - 1u;
            out3[3 * i + 1] = y;
            out3[3 * i + 2] = x;
        } else {
            out3[3 * i + 1] = NONE;
            out3[3 * i + 2] = NONE;
        }
    }
}

int oxo_trace_sample(const oxo_scene *s, const oxpt_params *P, uint32_t x, uint32_t y, uint32_t index,
                     int flags, float radiance[3]) {
    Ctx c = make_ctx(s, P, flags, nullptr);
    const uint32_t *sob = s->tables + P->offs0[0];
    const float *lut = reinterpret_cast<const float *>(s->tables + P->offs0[1]);
    PathResult r = trace_sample(c, sob, lut, x, y, index);
    radiance[0] = r.radiance.x;
    radiance[1] = r.radiance.y;
    radiance[2] = r.radiance.z;
    return r.covered ? 1 : 0;
}

int oxo_render(const oxo_scene *s, const oxpt_params *P, uint32_t first, uint32_t count, uint32_t x0,
               uint32_t y0, uint32_t x1, uint32_t y1, float *accum, int threads, int flags,
               oxo_counts *counts) {
    uint32_t w = P->dims[0], h = P->dims[1];
    x1 = std::min(x1, w);
    y1 = std::min(y1, h);
    if (x0 >= x1 || y0 >= y1 || count == 0) return 0;
    if (threads <= 0) threads = (int)std::min(16u, std::max(1u, std::thread::hardware_concurrency()));
    const uint32_t *sob = s->tables + P->offs0[0];
    const float *lut = reinterpret_cast<const float *>(s->tables + P->offs0[1]);
    std::atomic<uint32_t> next_row{y0};
    std::vector<oxo_counts> per(threads);
    std::memset(per.data(), 0, sizeof(oxo_counts) * per.size());
    auto worker = [&](int ti) {
        Ctx c = make_ctx(s, P, flags, counts ? &per[ti] : nullptr);
        for (;;) {
            uint32_t y = next_row.fetch_add(1);
            if (y >= y1) break;
            for (uint32_t x = x0; x < x1; x++) {
                float *a = accum + 4ull * ((uint64_t)y * w + x);
                float acc[4] = {a[0], a[1], a[2], a[3]};
                for (uint32_t i = 0; i < count; i++) {
                    PathResult r = trace_sample(c, sob, lut, x, y, first + i);
                    if (c.ctr) c.ctr->paths++;
                    if (r.covered) {
                        acc[0] += r.radiance.x;
                        acc[1] += r.radiance.y;
                        acc[2] += r.radiance.z;
                        acc[3] += 1.0f;
                    }
                }
                a[0] = acc[0];
                a[1] = acc[1];
                a[2] = acc[2];
                a[3] = acc[3];
                c.add(32);  // accumulator read + write per pixel per launch
            }
        }
    };
    if (threads == 1) {
        worker(0);
    } else {
        std::vector<std::thread> pool;
        for (int t = 0; t < threads; t++) pool.emplace_back(worker, t);
        for (auto &t : pool) t.join();
    }
    if (counts) {
        for (auto &p : per) {
            counts->paths += p.paths;
            counts->rays += p.rays;
            counts->nodes += p.nodes;
            counts->tris += p.tris;
            counts->hits += p.hits;
            counts->tex_taps += p.tex_taps;
            counts->bytes += p.bytes;
        }
    }
    return 0;
}

void oxo_closest_hit(const oxo_scene *s, const oxpt_params *P, const float o[3], const float d[3], int brute,
                     uint32_t *slot, float *t) {
    Ctx c = make_ctx(s, P, 0, nullptr);
    Sampler smp = make_sampler(0, 0, 0, 0);
    Ray r = make_ray({o[0], o[1], o[2]}, {d[0], d[1], d[2]});
    Hit best;
    best.slot = NONE;
    best.t = 0.0f;
    if (!brute) {
        best = traverse(c, r, INF, smp, 0, false, false, NONE);
    } else if (P->env[3] != 0 && r.valid) {
        // Every leaf slot in ascending order, keeping the nearest (ties: the
        // first slot, which a BVH walk may order differently).
        uint32_t nslots = 0;
        for (uint32_t i = 0; i < P->env[3]; i++) {
            uint32_t cnt = fbits(g4(c, 2 * i + 1).w), first = fbits(g4(c, 2 * i).w);
            if (cnt > 0) nslots = std::max(nslots, first + cnt);
        }
        float tmax = INF;
        for (uint32_t sl = 0; sl < nslots; sl++) {
            Hit h = intersect(c, r, sl, tmax);
            if (h.slot != NONE && h.t < tmax) {
                best = h;
                tmax = h.t;
            }
        }
    }
    *slot = best.slot;
    *t = best.slot == NONE ? 0.0f : best.t;
}

void oxo_resolve(const float *accum, uint32_t w, uint32_t h, const oxpt_resolve_params *rp, void *out) {
    bool hdr = rp->cfg[3] == 1;
    V3 bg_lin = {srgb_to_linear(rp->background[0]), srgb_to_linear(rp->background[1]),
                 srgb_to_linear(rp->background[2])};
    float ba = rp->background[3];
    float n = rp->exposure[1];
    for (uint64_t idx = 0; idx < (uint64_t)w * h; idx++) {
        const float *a = accum + 4 * idx;
        V4 o;
        if (a[3] <= 0.0f || n <= 0.0f) {
            o = {bg_lin.x, bg_lin.y, bg_lin.z, ba};
        } else {
            V3 m = V3{a[0], a[1], a[2]} / a[3];
            float f = clampf(a[3] / n, 0.0f, 1.0f);
            V3 cc = m;
            if (!hdr) cc = tone_map(rp->cfg[2], m * rp->exposure[0]);
            if (f >= 1.0f) {
                o = {cc.x, cc.y, cc.z, 1.0f};
            } else {
                float a_sum = f + (1.0f - f) * ba;
                if (a_sum > 0.0f) {
                    V3 r = (f * cc + (1.0f - f) * bg_lin * ba) / a_sum;
                    o = {r.x, r.y, r.z, a_sum};
                } else {
                    V3 r = f * cc + (1.0f - f) * bg_lin;
                    o = {r.x, r.y, r.z, 0.0f};
                }
            }
        }
        if (hdr) {
            float *dst = (float *)out + 4 * idx;
            dst[0] = o.x;
            dst[1] = o.y;
            dst[2] = o.z;
            dst[3] = o.w;
        } else {
            uint8_t *dst = (uint8_t *)out + 4 * idx;
            dst[0] = unorm8(linear_to_srgb(o.x));
            dst[1] = unorm8(linear_to_srgb(o.y));
            dst[2] = unorm8(linear_to_srgb(o.z));
            dst[3] = unorm8(o.w);
        }
    }
}

uint16_t oxo_f32_to_f16(float v) {
    // Restates src/scene.rs:244-272 (RNE, overflow to inf, subnormals, NaN).
    uint32_t bits = fbits(v);
    uint16_t sign = (uint16_t)((bits >> 16) & 0x8000u);
    int32_t exp = (int32_t)((bits >> 23) & 0xffu);
    uint32_t mant = bits & 0x007fffffu;
    if (exp == 0xff) return (uint16_t)(sign | 0x7c00u | (mant != 0 ? 0x200u : 0u));
    int32_t e = exp - 127 + 15;
    if (e >= 0x1f) return (uint16_t)(sign | 0x7c00u);
    if (e <= 0) {
        if (e < -10) return sign;
        uint32_t m = mant | 0x00800000u;
        uint32_t shift = (uint32_t)(14 - e);
        uint32_t half = m >> shift;
        uint32_t rem = m & ((1u << shift) - 1u);
        uint32_t mid = 1u << (shift - 1);
        bool round = rem > mid || (rem == mid && (half & 1u) == 1u);
        return (uint16_t)(sign | (half + (round ? 1u : 0u)));
    }
    uint32_t half = ((uint32_t)e << 10) | (mant >> 13);
    uint32_t rem = mant & 0x1fffu;
    bool round = rem > 0x1000u || (rem == 0x1000u && (half & 1u) == 1u);
    return (uint16_t)(sign | (half + (round ? 1u : 0u)));
}

void oxo_bsdf_probe(const uint32_t *mat, const float lut[OXPT_SHEEN_WORDS], const float n[3], const float v[3],
                    int front, const float u[3], const float l[3], float out[13]) {
    Mat m;
    auto f = [&](uint32_t i) { return bitsf(mat[i]); };
    m.base = {f(0), f(1), f(2), f(3)};
    m.metallic = clampf(f(8), 0.0f, 1.0f);
    m.roughness = clampf(f(9), 0.0f, 1.0f);
    m.normal = {n[0], n[1], n[2]};
    m.unlit = false;
    m.emission = v3(0.0f);
    m.ior = f(12);
    m.specular = f(13);
    m.specular_color = {f(14), f(15), f(16)};
    m.transmission = f(17);
    m.thickness = f(18);
    m.atten_color = {f(19), f(20), f(21)};
    m.atten_dist = f(22);
    m.clearcoat = f(23);
    m.cc_rough = f(24);
    m.cc_normal = m.normal;
    m.sheen = {f(26), f(27), f(28)};
    m.sheen_rough = f(29);
    V3 nn = {n[0], n[1], n[2]};
    Bsdf b = make_bsdf(lut, m, {v[0], v[1], v[2]}, nn, front != 0);
    BSample s = bsdf_sample(lut, b, {u[0], u[1], u[2]});
    out[0] = s.ok ? 1.0f : 0.0f;
    out[1] = s.l.x;
    out[2] = s.l.y;
    out[3] = s.l.z;
    out[4] = s.f.x;
    out[5] = s.f.y;
    out[6] = s.f.z;
    out[7] = s.pdf;
    out[8] = s.ok ? bsdf_eval(lut, b, s.l).pdf : 0.0f;
    Eval e = bsdf_eval(lut, b, {l[0], l[1], l[2]});
    out[9] = e.f.x;
    out[10] = e.f.y;
    out[11] = e.f.z;
    out[12] = e.pdf;
}

float oxo_solid_angle(const float p[3], const float a[3], const float b[3], const float c[3], int f64) {
    V3 P = {p[0], p[1], p[2]}, A = {a[0], a[1], a[2]}, B = {b[0], b[1], b[2]}, C = {c[0], c[1], c[2]};
    return f64 ? triangle_solid_angle_f64(P, A, B, C) : triangle_solid_angle_f32(P, A, B, C);
}

void oxo_texel_index(float u, int32_t size, uint32_t mode, int32_t out[3]) {
    // The index steps of filter_level (wgsl:355-380) for one coordinate.
    if (!finite(u)) u = 0.0f;
    u = clampf(u, -1.0e6f, 1.0e6f);
    float x = u * (float)size;
    int32_t x0 = sat_i32(std::floor(x - 0.5f));
    out[0] = wrap_index(sat_i32(std::floor(x)), size, mode);
    out[1] = wrap_index(x0, size, mode);
    out[2] = wrap_index(next_i32(x0), size, mode);
}

float oxo_ndf(int kind, float n_dot_h, float a) { return kind == 0 ? d_ggx(n_dot_h, a) : d_charlie(n_dot_h, a); }

void oxo_sample_spherical_triangle(const float p[3], const float a[3], const float b[3], const float c[3],
                                   float u0, float u1, int f64, float out[3]) {
    V3 P = {p[0], p[1], p[2]}, A = {a[0], a[1], a[2]}, B = {b[0], b[1], b[2]}, C = {c[0], c[1], c[2]};
    V3 r = f64 ? sample_spherical_triangle_f64(P, A, B, C, u0, u1) : sample_spherical_triangle_f32(P, A, B, C, u0, u1);
    out[0] = r.x;
    out[1] = r.y;
    out[2] = r.z;
}

}  // extern "C"
