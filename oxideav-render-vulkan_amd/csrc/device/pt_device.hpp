// Device-side estimator for gfx950: the reference kernel's functions
// (oxideav-render-vulkan src/shaders/pathtrace.wgsl, cited as wgsl:LINE)
// written for CDNA4 wave64. Scene words are read straight from HBM as
// float4 / u32 (include/oxpt/layout.h); the Sobol' matrices and the sheen
// LUT are staged in LDS once per workgroup (wgsl:99-100,2052-2058).
//
// Arithmetic follows the reference's operation order in f32 and the
// library is built with -ffp-contract=off, so the CPU oracle and this
// code round alike; transcendental functions are the device library's.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "oxpt/layout.h"

namespace oxd {

constexpr float PI = 3.14159265358979f;
constexpr float INF = 3.0e38f;
constexpr uint32_t NONE = 0xffffffffu;
constexpr float MIN_ALPHA = 1.0e-3f;
constexpr float MIN_SPHERICAL_SOLID_ANGLE = 1.0e-4f;
constexpr uint32_t STACK_DEPTH = 64;  // wgsl:661-662; overflow drops the far child

struct F3 {
    float x, y, z;
};
__device__ __forceinline__ F3 f3(float a) { return {a, a, a}; }
__device__ __forceinline__ F3 operator+(F3 a, F3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ F3 operator-(F3 a, F3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ F3 operator*(F3 a, F3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
__device__ __forceinline__ F3 operator*(F3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ F3 operator*(float s, F3 a) { return {s * a.x, s * a.y, s * a.z}; }
__device__ __forceinline__ F3 operator/(F3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
__device__ __forceinline__ F3 operator-(F3 a) { return {-a.x, -a.y, -a.z}; }
__device__ __forceinline__ float dot(F3 a, F3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ F3 cross(F3 a, F3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ float comp(F3 a, uint32_t k) { return k == 0 ? a.x : (k == 1 ? a.y : a.z); }
__device__ __forceinline__ F3 xyz(float4 a) { return {a.x, a.y, a.z}; }
__device__ __forceinline__ float4 add4(float4 a, float4 b) { return make_float4(a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w); }
__device__ __forceinline__ float4 mul4(float4 a, float4 b) { return make_float4(a.x * b.x, a.y * b.y, a.z * b.z, a.w * b.w); }
__device__ __forceinline__ float4 scale4(float4 a, float s) { return make_float4(a.x * s, a.y * s, a.z * s, a.w * s); }
__device__ __forceinline__ float mixf(float a, float b, float t) { return a * (1.0f - t) + b * t; }
__device__ __forceinline__ float4 mix4(float4 a, float4 b, float t) {
    return make_float4(mixf(a.x, b.x, t), mixf(a.y, b.y, t), mixf(a.z, b.z, t), mixf(a.w, b.w, t));
}
__device__ __forceinline__ float clampf(float x, float lo, float hi) { return fminf(fmaxf(x, lo), hi); }
__device__ __forceinline__ F3 vmin(F3 a, F3 b) { return {fminf(a.x, b.x), fminf(a.y, b.y), fminf(a.z, b.z)}; }
__device__ __forceinline__ F3 vmax(F3 a, F3 b) { return {fmaxf(a.x, b.x), fmaxf(a.y, b.y), fmaxf(a.z, b.z)}; }
__device__ __forceinline__ uint32_t fbits(float f) { return __float_as_uint(f); }
__device__ __forceinline__ float bitsf(uint32_t u) { return __uint_as_float(u); }
__device__ __forceinline__ bool finite(float x) { return (fbits(x) & 0x7f800000u) != 0x7f800000u; }
__device__ __forceinline__ bool finite3(F3 v) { return finite(v.x) && finite(v.y) && finite(v.z); }
__device__ __forceinline__ bool is_black(F3 a) { return !(a.x > 0.0f || a.y > 0.0f || a.z > 0.0f); }
__device__ __forceinline__ float max3(F3 a) { return fmaxf(fmaxf(a.x, a.y), a.z); }
__device__ __forceinline__ float mean3(F3 a) { return (a.x + a.y + a.z) * (1.0f / 3.0f); }
__device__ __forceinline__ F3 nrm(F3 v) {
    float len = sqrtf(dot(v, v));
    if (len < 1.1920929e-7f) return f3(0.0f);
    return v / len;
}
__device__ __forceinline__ F3 nrm_any(F3 v) {
    float l = sqrtf(dot(v, v));
    if (l > 0.0f) return v / l;
    return f3(0.0f);
}
__device__ __forceinline__ float sign_of(float x) { return (fbits(x) & 0x80000000u) ? -1.0f : 1.0f; }

// ---- RNG (wgsl:158-228) -------------------------------------------------

__device__ __forceinline__ uint32_t pcg_hash(uint32_t x) {
    uint32_t state = x * 747796405u + 2891336453u;
    uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
    return (word >> 22u) ^ word;
}
__device__ __forceinline__ uint32_t hash_mix(uint32_t a, uint32_t b) { return pcg_hash(a ^ pcg_hash(b)); }

// XOR of the direction words of the set bits of `index` (wgsl:168-183),
// from the LDS copy of the matrices.
__device__ __forceinline__ uint32_t sobol(const uint32_t *lds_dir, uint32_t index) {
    uint32_t r = 0;
    while (index != 0) {
        uint32_t k = __builtin_ctz(index);
        r ^= lds_dir[k];
        index &= index - 1u;
    }
    return r;
}
__device__ __forceinline__ uint32_t laine_karras(uint32_t x0, uint32_t seed) {
    uint32_t x = x0 + seed;
    x ^= x * 0x6c50b47cu;
    x ^= x * 0xb82f1e52u;
    x ^= x * 0xc7afe638u;
    x ^= x * 0x8d22f6e6u;
    return x;
}
__device__ __forceinline__ uint32_t nested_uniform_scramble(uint32_t x, uint32_t seed) {
    return __builtin_bitreverse32(laine_karras(__builtin_bitreverse32(x), seed));
}
__device__ __forceinline__ float to_unit(uint32_t v) { return (float)(v >> 8u) * (1.0f / 16777216.0f); }

__device__ __forceinline__ void sobol_owen_words(const uint32_t *lds_sobol, uint32_t index, uint32_t pattern_seed,
                                                 uint32_t out[4]) {
    uint32_t i = nested_uniform_scramble(index, pattern_seed);
#pragma unroll
    for (uint32_t d = 0; d < 4; d++)
        out[d] = nested_uniform_scramble(sobol(lds_sobol + 32u * d, i), hash_mix(pattern_seed, d));
}
__device__ __forceinline__ float4 pattern(const uint32_t *lds_sobol, uint32_t pixel_seed, uint32_t index, uint32_t p) {
    uint32_t w[4];
    sobol_owen_words(lds_sobol, index, hash_mix(pixel_seed, p), w);
    return make_float4(to_unit(w[0]), to_unit(w[1]), to_unit(w[2]), to_unit(w[3]));
}
__device__ __forceinline__ float coin(uint32_t path_seed, uint32_t ray, uint32_t tri) {
    return to_unit(hash_mix(hash_mix(path_seed, ray), tri));
}

// ---- Scene view --------------------------------------------------------

// Scene buffers are typed as global-memory (address_space(1)) pointers so
// every access compiles to global_load even when the pointer itself was
// read from memory (a job table) rather than from the kernel arguments;
// generic pointers loaded from memory would otherwise be accessed with
// slower flat instructions.
#define OXD_GLOBAL __attribute__((address_space(1)))
typedef float f32x4_t __attribute__((ext_vector_type(4)));
typedef uint32_t u32x4_t __attribute__((ext_vector_type(4)));
typedef uint32_t u32x2_t __attribute__((ext_vector_type(2)));
typedef const OXD_GLOBAL f32x4_t *gf4_ptr;
typedef const OXD_GLOBAL uint32_t *gu32_ptr;
typedef const OXD_GLOBAL uint16_t *gu16_ptr;

__device__ __forceinline__ float4 ldg4(gf4_ptr p, uint32_t i) {
    f32x4_t v = p[i];
    return make_float4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ uint4 ldgu4(gu32_ptr p) {
    u32x4_t v = *(const OXD_GLOBAL u32x4_t *)p;
    return make_uint4(v.x, v.y, v.z, v.w);
}

struct Scene {
    gf4_ptr geom;
    gf4_ptr attrs;
    gu32_ptr tables;
    gu16_ptr atlas;
    uint32_t atlas_side;
    uint32_t tri_base;    // P.offs2.z
    uint32_t wide_base;   // P.offs2.w: first 4-wide node float4 (0: walk the binary tree)
    uint32_t node_count;  // P.env.w
    uint32_t off_mat;     // P.offs0.z
    uint32_t off_plights; // P.offs0.w
    uint32_t off_elights; // P.offs1.x
    uint32_t n_emissive;  // P.cfg.w
    uint32_t env_w, env_h;
    uint32_t env_marg, env_cond, env_mass, env_rad;
    float p_env;
    const float *lds_sheen;
    // LDS copies of the light-selection CDFs (null: read them from tables):
    // the environment marginal (env_h + 1 floats) and the emissive CDF
    // (one float per emissive triangle).
    const float *lds_marg;
    const float *lds_ecdf;
    // geom + tri_base and geom + wide_base: the walk's loads index these
    // directly, so the offsets need not stay live beside the base pointer.
    gf4_ptr tris;
    gf4_ptr wide;
};

__device__ __forceinline__ float tf(const Scene &S, uint32_t i) { return bitsf(S.tables[i]); }
__device__ __forceinline__ F3 tv3(const Scene &S, uint32_t i) { return {tf(S, i), tf(S, i + 1), tf(S, i + 2)}; }
__device__ __forceinline__ float4 tv4(const Scene &S, uint32_t i) {
    return make_float4(tf(S, i), tf(S, i + 1), tf(S, i + 2), tf(S, i + 3));
}
__device__ __forceinline__ float4 tri_word(const Scene &S, uint32_t slot, uint32_t v) { return ldg4(S.tris, 3u * slot + v); }
__device__ __forceinline__ F3 tri_pos(const Scene &S, uint32_t slot, uint32_t v) { return xyz(tri_word(S, slot, v)); }
__device__ __forceinline__ uint32_t tri_global(const Scene &S, uint32_t slot) { return fbits(tri_word(S, slot, 0).w); }
__device__ __forceinline__ uint32_t tri_light(const Scene &S, uint32_t slot) { return fbits(tri_word(S, slot, 1).w); }
__device__ __forceinline__ uint32_t tri_material(const Scene &S, uint32_t slot) { return fbits(tri_word(S, slot, 2).w); }
__device__ __forceinline__ float4 attr(const Scene &S, uint32_t slot, uint32_t k) { return ldg4(S.attrs, slot * 12u + k); }
__device__ __forceinline__ uint32_t attr_flags(const Scene &S, uint32_t slot) { return fbits(ldg4(S.attrs, slot * 12u).w); }
__device__ __forceinline__ F3 interp3a(const Scene &S, uint32_t slot, uint32_t k, F3 b) {
    return xyz(attr(S, slot, k)) * b.x + xyz(attr(S, slot, k + 1)) * b.y + xyz(attr(S, slot, k + 2)) * b.z;
}
__device__ __forceinline__ float4 interp4a(const Scene &S, uint32_t slot, uint32_t k, F3 b) {
    return add4(add4(scale4(attr(S, slot, k), b.x), scale4(attr(S, slot, k + 1), b.y)), scale4(attr(S, slot, k + 2), b.z));
}
__device__ __forceinline__ float2 vertex_uv(const Scene &S, uint32_t slot, uint32_t uvset, uint32_t v) {
    float4 a = attr(S, slot, 6u + v);
    return uvset == 0 ? make_float2(a.x, a.y) : make_float2(a.z, a.w);
}
__device__ __forceinline__ bool has_uv(const Scene &S, uint32_t slot, uint32_t uvset) {
    uint32_t f = attr_flags(S, slot);
    if (uvset == 0) return (f & OXPT_F_UV0) != 0;
    if (uvset == 1) return (f & OXPT_F_UV1) != 0;
    return false;
}
__device__ __forceinline__ uint32_t mat_off(const Scene &S, uint32_t m) { return S.off_mat + m * OXPT_MAT_STRIDE; }
__device__ __forceinline__ uint32_t mat_flags(const Scene &S, uint32_t m) { return S.tables[mat_off(S, m) + OXPT_M_FLAGS]; }

// ---- Textures (wgsl:310-507) -------------------------------------------

struct Level {
    int32_t layer, x, y, w, h;
};

__device__ __forceinline__ Level tex_level(const Scene &S, uint32_t desc, uint32_t level) {
    uint32_t skip = S.tables[desc + 4], kept = S.tables[desc + 5];
    int32_t li = (int32_t)level - (int32_t)skip;
    uint32_t l = min((uint32_t)max(li, 0), kept - 1u);
    uint4 r = ldgu4(S.tables + desc + 8u + 4u * l);
    return {(int32_t)r.x, (int32_t)r.y, (int32_t)r.z, (int32_t)(r.w & 0xffffu), (int32_t)(r.w >> 16u)};
}
// binary16 bits -> f32 (exact).
__device__ __forceinline__ float h2f(uint32_t bits) {
    return (float)__builtin_bit_cast(_Float16, (uint16_t)bits);
}
// i mod n in [0, n), exactly, for n >= 1 and |i / n| <= 2^21 (texel
// indices: |u| <= 1e6 times the level size). The float quotient is within
// 1 of floor(i / n) there (|i| rounds to float with relative error 2^-24,
// the fast divide adds 2 ulp: < 0.5 in all), and one correction step
// makes the integer remainder exact — the value ((i % n) + n) % n gives,
// without two integer divisions.
__device__ __forceinline__ int32_t pos_mod(int32_t i, int32_t n) {
    const float q = floorf(__fdividef((float)i, (float)n));
    // i - q n in wrapping 32-bit arithmetic: the true remainder is within
    // (-n, 2n), so the wrapped difference is that remainder.
    int32_t r = (int32_t)((uint32_t)i - (uint32_t)(int32_t)q * (uint32_t)n);
    r = r < 0 ? r + n : r;
    r = r >= n ? r - n : r;
    return r;
}
// The reference's i32(f) of a floored texel coordinate: saturating at the
// i32 range (WGSL conversion rules; the CPU tracer's `as i32`). f is finite.
__device__ __forceinline__ int32_t sat_i32(float f) {
    return f >= 2147483648.0f ? 2147483647 : f < -2147483648.0f ? (-2147483647 - 1) : (int32_t)f;
}
// i + 1 in wrapping i32 arithmetic (the bilinear footprint's right / lower tap).
__device__ __forceinline__ int32_t next_i32(int32_t i) { return (int32_t)((uint32_t)i + 1u); }
__device__ __forceinline__ int32_t wrap_index(int32_t i, int32_t n, uint32_t mode) {
    if (mode == 0) return min(max(i, 0), n - 1);
    if (mode == 1) {
        int32_t m = pos_mod(i, 2 * n);
        return m >= n ? 2 * n - 1 - m : m;
    }
    return pos_mod(i, n);
}
__device__ __forceinline__ float4 fetch(const Scene &S, const Level &lv, int32_t x, int32_t y, uint32_t ws, uint32_t wt) {
    int32_t xi = wrap_index(x, lv.w, ws), yi = wrap_index(y, lv.h, wt);
    uint64_t side = S.atlas_side;
    uint64_t at = (((uint64_t)lv.layer * side + (uint64_t)(lv.y + yi)) * side + (uint64_t)(lv.x + xi));
    u32x2_t raw = ((const OXD_GLOBAL u32x2_t *)S.atlas)[at];
    return make_float4(h2f(raw.x & 0xffffu), h2f(raw.x >> 16), h2f(raw.y & 0xffffu), h2f(raw.y >> 16));
}
__device__ float4 filter_level(const Scene &S, const Level &lv, float2 uv, bool linear, uint32_t ws, uint32_t wt) {
    float u = uv.x, v = uv.y;
    if (!finite(u)) u = 0.0f;
    if (!finite(v)) v = 0.0f;
    u = clampf(u, -1.0e6f, 1.0e6f);
    v = clampf(v, -1.0e6f, 1.0e6f);
    float x = u * (float)lv.w, y = v * (float)lv.h;
    if (!linear) return fetch(S, lv, sat_i32(floorf(x)), sat_i32(floorf(y)), ws, wt);
    float xf = x - 0.5f, yf = y - 0.5f;
    float x0f = floorf(xf), y0f = floorf(yf);
    float fx = xf - x0f, fy = yf - y0f;
    int32_t x0 = sat_i32(x0f), y0 = sat_i32(y0f);
    const int32_t x1 = next_i32(x0), y1 = next_i32(y0);
    float4 t0 = fetch(S, lv, x0, y0, ws, wt), t1 = fetch(S, lv, x1, y0, ws, wt);
    float4 t2 = fetch(S, lv, x0, y1, ws, wt), t3 = fetch(S, lv, x1, y1, ws, wt);
    return mix4(mix4(t0, t1, fx), mix4(t2, t3, fx), fy);
}
__device__ float4 sample_lod(const Scene &S, uint32_t desc, float2 uv, float lod) {
    uint32_t levels = S.tables[desc], flags = S.tables[desc + 1];
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
        l0 = (uint32_t)clampf(floorf(lod + 0.5f), 0.0f, max_level);
    } else {
        float l = clampf(lod, 0.0f, max_level);
        l0 = (uint32_t)floorf(l);
        l1 = min(l0 + 1u, levels - 1u);
        f = l - (float)l0;
    }
    float4 r0 = filter_level(S, tex_level(S, desc, l0), uv, lin, ws, wt);
    if (!(f > 0.0f && l0 != l1)) return r0;
    float4 r1 = filter_level(S, tex_level(S, desc, l1), uv, lin, ws, wt);
    return mix4(r0, r1, f);
}
__device__ __forceinline__ uint32_t slot_word(const Scene &S, uint32_t m, uint32_t s) { return mat_off(S, m) + OXPT_M_SLOTS + 8u * s; }
__device__ __forceinline__ bool slot_present(const Scene &S, uint32_t m, uint32_t s, uint32_t slot) {
    uint32_t o = slot_word(S, m, s);
    return S.tables[o] != NONE && has_uv(S, slot, S.tables[o + 1]);
}
__device__ __forceinline__ float2 slot_xform(const Scene &S, uint32_t m, uint32_t s, float2 uv) {
    uint32_t o = slot_word(S, m, s);
    return make_float2(tf(S, o + 2) * uv.x + tf(S, o + 3) * uv.y + tf(S, o + 4),
                       tf(S, o + 5) * uv.x + tf(S, o + 6) * uv.y + tf(S, o + 7));
}
struct Lod {
    float width;
    F3 dir;
};
__device__ __forceinline__ Lod base_lod() { return {-1.0f, f3(0.0f)}; }

__device__ float cone_lod(const Scene &S, uint32_t m, uint32_t s, uint32_t slot, const Lod &lod) {
    uint32_t o = slot_word(S, m, s);
    uint32_t desc = S.tables[o], uvset = S.tables[o + 1];
    float2 uv0 = slot_xform(S, m, s, vertex_uv(S, slot, uvset, 0));
    float2 uv1 = slot_xform(S, m, s, vertex_uv(S, slot, uvset, 1));
    float2 uv2 = slot_xform(S, m, s, vertex_uv(S, slot, uvset, 2));
    float ta = (float)S.tables[desc + 2] * (float)S.tables[desc + 3] *
               fabsf((uv1.x - uv0.x) * (uv2.y - uv0.y) - (uv2.x - uv0.x) * (uv1.y - uv0.y));
    F3 p0 = tri_pos(S, slot, 0);
    F3 cr = cross(tri_pos(S, slot, 1) - p0, tri_pos(S, slot, 2) - p0);
    float pa = sqrtf(dot(cr, cr));
    if (!(ta > 0.0f && pa > 0.0f && lod.width > 0.0f)) return 0.0f;
    F3 n = cr / pa;
    float c = fmaxf(fabsf(dot(n, lod.dir)), 1.0e-3f);
    float l = 0.5f * log2f(ta / pa) + log2f(lod.width / c);
    return finite(l) ? l : 0.0f;
}
__device__ float4 slot_sample(const Scene &S, uint32_t m, uint32_t s, uint32_t slot, F3 b, const Lod &lod) {
    uint32_t o = slot_word(S, m, s);
    uint32_t uvset = S.tables[o + 1];
    float2 a0 = vertex_uv(S, slot, uvset, 0), a1 = vertex_uv(S, slot, uvset, 1), a2 = vertex_uv(S, slot, uvset, 2);
    float2 uv = make_float2(a0.x * b.x + a1.x * b.y + a2.x * b.z, a0.y * b.x + a1.y * b.y + a2.y * b.z);
    float l = 0.0f;
    if (lod.width >= 0.0f) l = cone_lod(S, m, s, slot, lod);
    return sample_lod(S, S.tables[o], slot_xform(S, m, s, uv), l);
}
__device__ float hit_alpha(const Scene &S, uint32_t slot, F3 b) {
    uint32_t m = tri_material(S, slot);
    float a = tf(S, mat_off(S, m) + 3);
    if (slot_present(S, m, OXPT_S_BASE, slot)) a *= slot_sample(S, m, OXPT_S_BASE, slot, b, base_lod()).w;
    if (attr_flags(S, slot) & OXPT_F_COLORS) a *= interp4a(S, slot, 9, b).w;
    return a;
}

// ---- Rays and traversal (wgsl:513-735) --------------------------------

struct Ray {
    F3 o, d, inv;
    uint32_t kx, ky, kz;
    F3 shear;
    bool valid;
    // 4-wide node rows (float4 index in the node) holding each axis's near
    // planes: lo (rows 0, 1, 2) where the direction is positive, else hi
    // (rows 3, 4, 5); the far planes are the other row of the pair.
    uint32_t nx, ny, nz;
};

__device__ __forceinline__ Ray make_ray(F3 o, F3 d) {
    Ray r;
    r.o = o;
    r.d = d;
    r.inv.x = fabsf(d.x) < 1.0e-30f ? sign_of(d.x) * 1.0e30f : 1.0f / d.x;
    r.inv.y = fabsf(d.y) < 1.0e-30f ? sign_of(d.y) * 1.0e30f : 1.0f / d.y;
    r.inv.z = fabsf(d.z) < 1.0e-30f ? sign_of(d.z) * 1.0e30f : 1.0f / d.z;
    float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
    uint32_t kz = 2;
    if (ax >= ay && ax >= az) kz = 0;
    else if (ay >= az) kz = 1;
    uint32_t kx = kz == 2 ? 0u : kz + 1u;
    uint32_t ky = kx == 2 ? 0u : kx + 1u;
    float dz = comp(d, kz);
    if (dz < 0.0f) {
        uint32_t t = kx;
        kx = ky;
        ky = t;
    }
    r.kx = kx;
    r.ky = ky;
    r.kz = kz;
    r.shear = f3(0.0f);
    if (dz != 0.0f) r.shear = {comp(d, kx) / dz, comp(d, ky) / dz, 1.0f / dz};
    r.valid = finite3(o) && finite3(d) && (d.x != 0.0f || d.y != 0.0f || d.z != 0.0f);
    r.nx = r.inv.x >= 0.0f ? 0u : 3u;
    r.ny = r.inv.y >= 0.0f ? 1u : 4u;
    r.nz = r.inv.z >= 0.0f ? 2u : 5u;
    return r;
}

__device__ __forceinline__ float slab(const Ray &r, float4 lo, float4 hi, float t_max) {
    float t0x = (lo.x - r.o.x) * r.inv.x, t1x = (hi.x - r.o.x) * r.inv.x;
    float t0y = (lo.y - r.o.y) * r.inv.y, t1y = (hi.y - r.o.y) * r.inv.y;
    float t0z = (lo.z - r.o.z) * r.inv.z, t1z = (hi.z - r.o.z) * r.inv.z;
    float tnx = fminf(t0x, t1x), tny = fminf(t0y, t1y), tnz = fminf(t0z, t1z);
    float tfx = fmaxf(t0x, t1x), tfy = fmaxf(t0y, t1y), tfz = fmaxf(t0z, t1z);
    float t0 = fmaxf(fmaxf(0.0f, tnx), fmaxf(tny, tnz));
    float t1 = fminf(t_max, fminf(tfx, fminf(tfy, tfz)) * 1.0000004f);
    return t0 <= t1 ? t0 : -1.0f;
}

// slab() on a box given as its near and far planes along the ray (near =
// lo where the direction component is positive, else hi): for lo <= hi the
// entry and exit distances of each axis are already the min and max slab()
// takes (rounding is monotone and the sign of inv orders them), so the
// result is slab()'s to the bit with six fewer operations.
__device__ __forceinline__ float slab_nf(const Ray &r, float nx, float ny, float nz, float fx, float fy, float fz,
                                         float t_max) {
    float tnx = (nx - r.o.x) * r.inv.x, tfx = (fx - r.o.x) * r.inv.x;
    float tny = (ny - r.o.y) * r.inv.y, tfy = (fy - r.o.y) * r.inv.y;
    float tnz = (nz - r.o.z) * r.inv.z, tfz = (fz - r.o.z) * r.inv.z;
    float t0 = fmaxf(fmaxf(0.0f, tnx), fmaxf(tny, tnz));
    float t1 = fminf(t_max, fminf(tfx, fminf(tfy, tfz)) * 1.0000004f);
    return t0 <= t1 ? t0 : -1.0f;
}

struct Hit {
    uint32_t slot;
    float t;
    F3 b;
    bool front;
};

// Woop-Benthin-Wald watertight test on a triangle already in registers.
__device__ __forceinline__ bool intersect_tri(const Ray &r, F3 p0, F3 p1, F3 p2, float t_max, float &t_out,
                                              F3 &b_out) {
    F3 a = p0 - r.o, b = p1 - r.o, c = p2 - r.o;
    float akz = comp(a, r.kz), bkz = comp(b, r.kz), ckz = comp(c, r.kz);
    float ax = comp(a, r.kx) - r.shear.x * akz;
    float ay = comp(a, r.ky) - r.shear.y * akz;
    float bx = comp(b, r.kx) - r.shear.x * bkz;
    float by = comp(b, r.ky) - r.shear.y * bkz;
    float cx = comp(c, r.kx) - r.shear.x * ckz;
    float cy = comp(c, r.ky) - r.shear.y * ckz;
    float uu = cx * by - cy * bx;
    float vv = ax * cy - ay * cx;
    float ww = bx * ay - by * ax;
    if ((uu < 0.0f || vv < 0.0f || ww < 0.0f) && (uu > 0.0f || vv > 0.0f || ww > 0.0f)) return false;
    float det = uu + vv + ww;
    if (det == 0.0f || !finite(det)) return false;
    float az = r.shear.z * akz, bz = r.shear.z * bkz, cz = r.shear.z * ckz;
    float rcp = 1.0f / det;
    float t = (uu * az + vv * bz + ww * cz) * rcp;
    if (!(t >= 0.0f && t <= t_max)) return false;
    float u = vv * rcp, v = ww * rcp;
    t_out = t;
    b_out = {1.0f - u - v, u, v};
    return true;
}

// Whether the ray meets the triangle's front face (wgsl:617, with the
// intersection above). The walks evaluate it per accepted candidate; built
// with Defer, only for camera rays (whose acceptance culls back faces), and
// for other closest-hit rays once, for the hit the walk returns (any-hit
// queries do not use it). Deferring saves the batched form work; the
// single-frame form at 6 waves per SIMD spills more with it (the 1080p
// frame 1027 -> 925 Mpaths/s, the uneven batch 1053 -> 1062).
__device__ __forceinline__ bool tri_front(const Ray &r, F3 p0, F3 p1, F3 p2) {
    return dot(r.d, cross(p1 - p0, p2 - p0)) < 0.0f;
}
// The front flag of the hit a closest-hit walk returns (camera rays have it
// from their acceptance test already).
__device__ __forceinline__ void set_front(const Scene &S, const Ray &r, bool primary, bool any, Hit &best) {
    if (best.slot != NONE && !primary && !any)
        best.front = tri_front(r, tri_pos(S, best.slot, 0), tri_pos(S, best.slot, 1), tri_pos(S, best.slot, 2));
}

// Any-hit acceptance (wgsl:622-636) given the material flags word.
__device__ __forceinline__ bool accept(const Scene &S, uint32_t slot, F3 b, bool front, uint32_t mat, uint32_t fl,
                                       uint32_t path_seed, uint32_t ray_id, bool primary, uint32_t gid) {
    if (primary && !front && (fl & 4u) == 0) return false;
    uint32_t mode = fl & 3u;
    if (mode == 0) return true;
    if (mode == 1) return hit_alpha(S, slot, b) >= tf(S, mat_off(S, mat) + 11);
    return coin(path_seed, ray_id, gid) < hit_alpha(S, slot, b);
}

// Acceptance of a triangle hit that passed the skip test. Opaque walks are
// built for scenes whose every material has alpha mode OPAQUE (flags bits
// 0-1 clear): accept() then only culls back faces of camera rays on
// single-sided materials, so the material word is read for those alone and
// the alpha lookup (a texture sample) is not compiled into the walk loop.
template <bool Opaque>
__device__ __forceinline__ bool accept_hit(const Scene &S, uint32_t slot, F3 b, bool front, uint32_t mat,
                                           uint32_t path_seed, uint32_t ray_id, bool primary, uint32_t gid) {
    if constexpr (Opaque) {
        return !(primary && !front && (mat_flags(S, mat) & 4u) == 0);
    } else {
        return accept(S, slot, b, front, mat, mat_flags(S, mat), path_seed, ray_id, primary, gid);
    }
}

// Traversal stack: 64 entries (the reference's depth, wgsl:661-662), each
// (t_far bits << 32 | node). The bottom `lds_depth` entries of every lane
// live in LDS, laid out [entry][lane] in the wave's own slice so lanes at
// the same depth touch consecutive words; deeper entries go to private
// memory. LdsDepth > 0 fixes that share at compile time, LdsDepth < 0 takes
// it from the run-time `lds_depth` (0..64), LdsDepth == 0 keeps the whole
// stack private. The LDS pointer is typed address_space(3) so its accesses
// compile to ds_read / ds_write rather than flat accesses.
typedef __attribute__((address_space(3))) uint64_t lds_u64;

__device__ __forceinline__ lds_u64 *to_lds(uint64_t *p) { return (lds_u64 *)p; }

// Stack entry: bit 63 clear -> bits 62..32 are |t_far| and the low word is
// the node record itself (first << 5 | count, count <= 30, first < 2^27),
// so a pop needs no memory access; bit 63 set -> the low word is the node
// index and the record is read back (leaves past 30 slots or indices past
// 2^27, which only very large or degenerate scenes produce). t_far >= 0, so
// dropping its sign bit (+0 for -0) leaves every comparison unchanged.
__device__ __forceinline__ uint64_t stack_entry(float t_far, uint32_t first, uint32_t count, uint32_t node) {
    const uint64_t t = (uint64_t)(fbits(t_far) & 0x7fffffffu) << 32;
    if (count <= 30u && first < (1u << 27)) return t | (first << 5 | count);
    return t | (1ull << 63) | node;
}

// Ordered walk (wgsl:651-735): near child first, far child pushed with its
// entry distance, pops culled by t_max. The current node's record (first,
// count) travels with it — read from the parent's child pair on descent,
// from the stack entry on a pop — so each inner-node step issues only the
// one 64-byte child-pair load.
template <int LdsDepth, bool Opaque = false, bool Defer = false>
__device__ __forceinline__ Hit traverse2(const Scene &S, const Ray &r, float t_max_in, uint32_t path_seed,
                                         uint32_t ray_id, bool primary, bool any, uint32_t skip, lds_u64 *lds_stack,
                                         uint32_t lds_depth) {
    Hit best;
    best.slot = NONE;
    best.t = 0.0f;
    best.b = f3(0.0f);
    best.front = false;
    if (S.node_count == 0 || !r.valid) return best;
    float t_max = t_max_in;
    uint32_t first, count;
    {
        const float4 lo = ldg4(S.geom, 0), hi = ldg4(S.geom, 1);
        if (slab(r, lo, hi, t_max) < 0.0f) return best;
        first = fbits(lo.w);
        count = fbits(hi.w);
    }
    const uint32_t in_lds = LdsDepth > 0 ? (uint32_t)LdsDepth : (LdsDepth < 0 ? lds_depth : 0u);
    lds_u64 *const lds = lds_stack + __lane_id();
    uint64_t priv[LdsDepth > 0 ? STACK_DEPTH - LdsDepth : STACK_DEPTH];
    constexpr uint32_t priv_off = LdsDepth > 0 ? (uint32_t)LdsDepth : 0u;
    uint32_t sp = 0;
    for (;;) {
        bool descended = false;
        if (count > 0) {
            bool done = false;
            for (uint32_t slot = first; slot < first + count; slot++) {
                float4 w0 = tri_word(S, slot, 0), w1 = tri_word(S, slot, 1), w2 = tri_word(S, slot, 2);
                float t;
                F3 b;
                if (!intersect_tri(r, xyz(w0), xyz(w1), xyz(w2), t_max, t, b)) continue;
                uint32_t gid = fbits(w0.w);
                if (gid == skip) continue;
                const bool front = (!Defer || primary) && tri_front(r, xyz(w0), xyz(w1), xyz(w2));
                if (!accept_hit<Opaque>(S, slot, b, front, fbits(w2.w), path_seed, ray_id, primary, gid)) continue;
                best.slot = slot;
                best.t = t;
                best.b = b;
                best.front = front;
                if (any) {
                    done = true;
                    break;
                }
                t_max = t;
            }
            if (done) break;
        } else {
            const uint32_t l = first;
            const float4 llo = ldg4(S.geom, 2u * l), lhi = ldg4(S.geom, 2u * l + 1u),
                         rlo = ldg4(S.geom, 2u * l + 2u), rhi = ldg4(S.geom, 2u * l + 3u);
            const float tl = slab(r, llo, lhi, t_max);
            const float tr = slab(r, rlo, rhi, t_max);
            if (tl >= 0.0f && tr >= 0.0f) {
                const bool swap = tl > tr;
                const uint32_t far_ = swap ? l : l + 1u;
                const float t_far = swap ? tl : tr;
                const uint32_t far_first = fbits(swap ? llo.w : rlo.w), far_count = fbits(swap ? lhi.w : rhi.w);
                if (sp < STACK_DEPTH) {
                    const uint64_t e = stack_entry(t_far, far_first, far_count, far_);
                    if (sp < in_lds) lds[sp * 64u] = e;
                    else priv[sp - priv_off] = e;
                    sp++;
                }
                first = fbits(swap ? rlo.w : llo.w);
                count = fbits(swap ? rhi.w : lhi.w);
                descended = true;
            } else if (tl >= 0.0f) {
                first = fbits(llo.w);
                count = fbits(lhi.w);
                descended = true;
            } else if (tr >= 0.0f) {
                first = fbits(rlo.w);
                count = fbits(rhi.w);
                descended = true;
            }
        }
        if (descended) continue;
        bool found = false;
        while (sp != 0) {
            sp--;
            const uint64_t e = sp < in_lds ? lds[sp * 64u] : priv[sp - priv_off];
            if (bitsf((uint32_t)(e >> 32) & 0x7fffffffu) <= t_max) {
                const uint32_t w = (uint32_t)e;
                if (e >> 63) {
                    first = fbits(ldg4(S.geom, 2u * w).w);
                    count = fbits(ldg4(S.geom, 2u * w + 1u).w);
                } else {
                    first = w >> 5;
                    count = w & 31u;
                }
                found = true;
                break;
            }
        }
        if (!found) break;
    }
    if constexpr (Defer) set_front(S, r, primary, any, best);
    return best;
}

// Leaf: test slots [first, first + count) in order (wgsl:683-699); returns
// true when an any-hit query is answered.
template <bool Opaque, bool Defer = false>
__device__ __forceinline__ bool leaf_hits(const Scene &S, const Ray &r, uint32_t first, uint32_t count, float &t_max,
                                          Hit &best, uint32_t path_seed, uint32_t ray_id, bool primary, bool any,
                                          uint32_t skip) {
    for (uint32_t slot = first; slot < first + count; slot++) {
        float4 w0 = tri_word(S, slot, 0), w1 = tri_word(S, slot, 1), w2 = tri_word(S, slot, 2);
        float t;
        F3 b;
        if (!intersect_tri(r, xyz(w0), xyz(w1), xyz(w2), t_max, t, b)) continue;
        uint32_t gid = fbits(w0.w);
        if (gid == skip) continue;
        const bool front = (!Defer || primary) && tri_front(r, xyz(w0), xyz(w1), xyz(w2));
        if (!accept_hit<Opaque>(S, slot, b, front, fbits(w2.w), path_seed, ray_id, primary, gid)) continue;
        best.slot = slot;
        best.t = t;
        best.b = b;
        best.front = front;
        if (any) return true;
        t_max = t;
    }
    return false;
}

// Walk of the 4-wide tree (the binary tree's boxes and leaves, two levels
// per node; see bvh.cpp). Per node: one 128-byte load, four slab tests,
// the hit children ordered by entry distance (ties by slot), the nearest
// entered and the others pushed farthest first. Every box on the way to
// a leaf is the binary walk's box, so the leaves reached and the closest
// accepted hit are the same; only the order between leaves differs, which
// decides nothing but exact ties in t. A node has at most 3 pushes, and
// the builder bounds 3 * depth by WIDE_STACK, so no entry is ever dropped.
constexpr uint32_t WIDE_STACK = 96;
constexpr uint32_t WIDE_EMPTY = 0xffffffffu;

// One step of the 4-wide walk. A lane holds a current node (first, count)
// and a pending leaf (lf, lc: its untested slots). A step tests the pending
// leaf's next triangle and, in the same step, takes the current node: an
// inner node's hit children are ordered and the nearest entered; a leaf
// becomes the pending leaf once none is pending (until then the lane only
// drains the pending one); with no current node (first = WIDE_EMPTY) one
// stack entry is popped. So a lane walks on to its next node while it tests
// a leaf's triangles, and the triangle and node work of a wave overlap
// instead of alternating. Each leaf's triangles are tested in slot order and
// every leaf entered before t_max is tested, so the closest accepted hit is
// the ordered walk's; only the order between leaves differs, which decides
// nothing but exact ties in t. Returns false when the walk is over (nothing
// pending, no current node, empty stack, or an any-hit query answered).
// AnyUnordered: any-hit queries skip the ordering (the answer is the same;
// faster where shadow rays to punctual lights dominate, slower on open
// scenes whose shadow rays find near occluders sooner in distance order).
template <bool AnyUnordered = false, bool Opaque = false, bool Defer = false>
__device__ __forceinline__ bool wide_step(const Scene &S, const Ray &r, float &t_max, Hit &best, uint32_t &first,
                                          uint32_t &count, uint32_t &lf, uint32_t &lc, uint32_t &sp,
                                          uint32_t path_seed, uint32_t ray_id, bool primary, bool any, uint32_t skip,
                                          lds_u64 *lds, uint64_t *priv, uint32_t in_lds, uint32_t priv_off) {
    // The pending leaf's next triangle (its triangles in slot order, as
    // leaf_hits takes them, one per step).
    if (lc > 0) {
        if (leaf_hits<Opaque, Defer>(S, r, lf, 1, t_max, best, path_seed, ray_id, primary, any, skip)) return false;
        lf++;
        lc--;
    }
    if (count > 0) {
        // The current node is a leaf: it becomes the pending leaf once that
        // is drained, and the walk goes on to the next node meanwhile.
        if (lc > 0) return true;
        lf = first;
        lc = count;
        first = WIDE_EMPTY;
        count = 0;
    } else if (first != WIDE_EMPTY) {
        const gf4_ptr nd = S.wide + 8u * first;
        const uint4 rf = ldgu4((gu32_ptr)(nd + 6)), cn = ldgu4((gu32_ptr)(nd + 7));
        // Near and far planes of the four children per axis (the node's lo
        // and hi rows, ordered by the ray's direction signs). An empty
        // slot's box is lo = +inf, hi = -inf (bvh.cpp): its near plane is
        // +inf and its far plane -inf whatever the direction, so slab_nf
        // misses it by itself.
        const float4 nx = ldg4(nd, r.nx), ny = ldg4(nd, r.ny), nz = ldg4(nd, r.nz);
        const float4 fx = ldg4(nd, 3u - r.nx), fy = ldg4(nd, 5u - r.ny), fz = ldg4(nd, 7u - r.nz);
        const float t0 = slab_nf(r, nx.x, ny.x, nz.x, fx.x, fy.x, fz.x, t_max);
        const float t1 = slab_nf(r, nx.y, ny.y, nz.y, fx.y, fy.y, fz.y, t_max);
        const float t2 = slab_nf(r, nx.z, ny.z, nz.z, fx.z, fy.z, fz.z, t_max);
        const float t3 = slab_nf(r, nx.w, ny.w, nz.w, fx.w, fy.w, fz.w, t_max);
        if (AnyUnordered && any) {
            // Any-hit query: t_max stays put until the walk ends at the
            // first accepted hit, so the order children are visited in
            // cannot change the answer (is some triangle on the segment
            // accepted) — only how soon it is found. Enter the first hit
            // child in slot order and push the others unsorted.
            const float tj[4] = {t0, t1, t2, t3};
            const uint32_t rj[4] = {rf.x, rf.y, rf.z, rf.w}, cj[4] = {cn.x, cn.y, cn.z, cn.w};
            bool entered = false;
            first = WIDE_EMPTY;
            count = 0;
#pragma unroll
            for (int j = 0; j < 4; j++) {
                if (!(tj[j] >= 0.0f)) continue;
                if (!entered) {
                    first = rj[j];
                    count = cj[j];
                    entered = true;
                } else if (sp < WIDE_STACK) {
                    const uint64_t e = ((uint64_t)(fbits(tj[j]) & 0x7fffffffu) << 32) | (rj[j] << 5 | cj[j]);
                    if (sp < in_lds) lds[sp * 64u] = e;
                    else priv[sp - priv_off] = e;
                    sp++;
                }
            }
            if (count > 0 && lc == 0) {
                lf = first;
                lc = count;
                first = WIDE_EMPTY;
                count = 0;
            } else if (entered) {
                return true;
            }
        } else {
        // Sort keys: (|t| bits << 32 | child record) for hit children, all
        // ones for misses. A key is the stack entry itself (stack_entry's
        // short form: every wide child record fits, first < 2^27, count <=
        // 30), so a push stores it and the entered child's record is read
        // back without selecting among the children. Equal entry distances
        // are ordered by record, which only decides ties between exactly
        // equal triangle distances, as any visiting order does.
        uint64_t k[4];
        k[0] = t0 >= 0.0f ? ((uint64_t)(fbits(t0) & 0x7fffffffu) << 32) | (rf.x << 5 | cn.x) : ~0ull;
        k[1] = t1 >= 0.0f ? ((uint64_t)(fbits(t1) & 0x7fffffffu) << 32) | (rf.y << 5 | cn.y) : ~0ull;
        k[2] = t2 >= 0.0f ? ((uint64_t)(fbits(t2) & 0x7fffffffu) << 32) | (rf.z << 5 | cn.z) : ~0ull;
        k[3] = t3 >= 0.0f ? ((uint64_t)(fbits(t3) & 0x7fffffffu) << 32) | (rf.w << 5 | cn.w) : ~0ull;
#define OXD_CSWAP(a, b)                                 \
    {                                                   \
        const uint64_t lo_ = k[a] < k[b] ? k[a] : k[b]; \
        const uint64_t hi_ = k[a] < k[b] ? k[b] : k[a]; \
        k[a] = lo_;                                     \
        k[b] = hi_;                                     \
    }
        OXD_CSWAP(0, 1) OXD_CSWAP(2, 3) OXD_CSWAP(0, 2) OXD_CSWAP(1, 3) OXD_CSWAP(1, 2)
#undef OXD_CSWAP
        if (k[0] != ~0ull) {
            // Push the farther hit children, farthest first.
#pragma unroll
            for (int j = 3; j >= 1; j--) {
                if (k[j] == ~0ull) continue;
                if (sp < WIDE_STACK) {
                    if (sp < in_lds) lds[sp * 64u] = k[j];
                    else priv[sp - priv_off] = k[j];
                    sp++;
                }
            }
            first = (uint32_t)k[0] >> 5;
            count = (uint32_t)k[0] & 31u;
            if (count == 0 || lc > 0) return true;
            // The nearest child is a leaf and none is pending: it becomes
            // the pending leaf and the next node is popped in this step.
            lf = first;
            lc = count;
        }
        first = WIDE_EMPTY;
        count = 0;
        }
    }
    // One pop per step as well: an entry beyond t_max leaves the lane with
    // no current node (first = WIDE_EMPTY), and the next step pops again,
    // so a lane discarding entries does not hold its wave in a pop loop.
    if (sp == 0) return lc > 0;
    sp--;
    const uint64_t e = sp < in_lds ? lds[sp * 64u] : priv[sp - priv_off];
    if (bitsf((uint32_t)(e >> 32)) <= t_max) {
        first = (uint32_t)e >> 5;
        count = (uint32_t)e & 31u;
    } else {
        first = WIDE_EMPTY;
        count = 0;
    }
    return true;
}

template <int LdsDepth, bool AnyUnordered = false, bool Opaque = false, bool Defer = false>
__device__ __forceinline__ Hit traverse4(const Scene &S, const Ray &r, float t_max_in, uint32_t path_seed,
                                         uint32_t ray_id, bool primary, bool any, uint32_t skip, lds_u64 *lds_stack,
                                         uint32_t lds_depth) {
    Hit best;
    best.slot = NONE;
    best.t = 0.0f;
    best.b = f3(0.0f);
    best.front = false;
    if (S.node_count == 0 || !r.valid) return best;
    float t_max = t_max_in;
    if (slab(r, ldg4(S.geom, 0), ldg4(S.geom, 1), t_max) < 0.0f) return best;
    const uint32_t in_lds = LdsDepth > 0 ? (uint32_t)LdsDepth : (LdsDepth < 0 ? lds_depth : 0u);
    lds_u64 *const lds = lds_stack + __lane_id();
    uint64_t priv[LdsDepth > 0 ? WIDE_STACK - LdsDepth : WIDE_STACK];
    constexpr uint32_t priv_off = LdsDepth > 0 ? (uint32_t)LdsDepth : 0u;
    uint32_t sp = 0, first = 0, count = 0;  // wide node 0, inner
    uint32_t lf = 0, lc = 0;                // no pending leaf
    while (wide_step<AnyUnordered, Opaque, Defer>(S, r, t_max, best, first, count, lf, lc, sp, path_seed, ray_id, primary, any, skip,
                                   lds, priv, in_lds, priv_off)) {
    }
    if constexpr (Defer) set_front(S, r, primary, any, best);
    return best;
}

// Closest (or, with `any`, some) accepted hit of the ray up to t_max_in:
// the 4-wide walk when the scene carries one, else the binary walk.
template <int LdsDepth = 0, bool AnyUnordered = false, bool Opaque = false, bool Defer = false>
__device__ __forceinline__ Hit traverse(const Scene &S, const Ray &r, float t_max_in, uint32_t path_seed,
                                        uint32_t ray_id, bool primary, bool any, uint32_t skip,
                                        lds_u64 *lds_stack = nullptr, uint32_t lds_depth = 0) {
    if (S.wide_base != 0)
        return traverse4<LdsDepth, AnyUnordered, Opaque, Defer>(S, r, t_max_in, path_seed, ray_id, primary, any, skip, lds_stack,
                                                 lds_depth);
    return traverse2<LdsDepth, Opaque, Defer>(S, r, t_max_in, path_seed, ray_id, primary, any, skip, lds_stack, lds_depth);
}

__device__ __forceinline__ float offset_one(float p, float n) {
    int32_t of_i = (int32_t)(256.0f * n);
    int32_t bits = (int32_t)fbits(p);
    int32_t moved = p < 0.0f ? (int32_t)((uint32_t)bits - (uint32_t)of_i) : (int32_t)((uint32_t)bits + (uint32_t)of_i);
    float p_i = bitsf((uint32_t)moved);
    if (fabsf(p) < 1.0f / 32.0f) return p + (1.0f / 65536.0f) * n;
    if (finite(p_i)) return p_i;
    return p;
}
// Wachter & Binder 2019 ray-origin offset (wgsl:738-759).
__device__ __forceinline__ F3 offset_ray_origin(F3 p, F3 n) {
    return {offset_one(p.x, n.x), offset_one(p.y, n.y), offset_one(p.z, n.z)};
}

}  // namespace oxd
