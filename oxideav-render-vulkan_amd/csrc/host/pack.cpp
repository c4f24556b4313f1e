// Scene -> kernel words: the host half of GpuPathTracer::upload and
// ensure_binding (oxideav-render-vulkan src/pathtrace.rs:847-1239,
// 1275-1627). Word layouts follow include/oxpt/layout.h.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <stdexcept>
#include <thread>

#include "scene.hpp"

namespace oxpt {
namespace {

inline uint32_t fb(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    return u;
}
inline float luminance(const float *c) { return 0.2126f * c[0] + 0.7152f * c[1] + 0.0722f * c[2]; }
inline bool finitef(float v) { return std::isfinite(v); }

float srgb_to_linear(float c) {
    if (c <= 0.04045f) return c / 12.92f;
    return std::pow((c + 0.055f) / 1.055f, 2.4f);
}

// One mip level of RGBA f32 texels.
struct Mip {
    uint32_t w, h;
    std::vector<float> t;  // 4 per texel
};

// Mip chain of a texture in one colour space: level 0 decoded (sRGB
// channels through the sRGB EOTF, alpha linear), each further level a
// 2x2 box filter (edge texels repeated on odd sizes), down to 1x1.
std::vector<Mip> build_mips(const Texture &tx, bool srgb) {
    std::vector<Mip> out;
    Mip m0{tx.w, tx.h, std::vector<float>((size_t)tx.w * tx.h * 4)};
    float lut[256];
    for (int i = 0; i < 256; i++) lut[i] = srgb ? srgb_to_linear((float)i / 255.0f) : (float)i / 255.0f;
    for (size_t i = 0; i < (size_t)tx.w * tx.h; i++) {
        for (int c = 0; c < 3; c++) m0.t[4 * i + c] = lut[tx.rgba8[4 * i + c]];
        m0.t[4 * i + 3] = (float)tx.rgba8[4 * i + 3] / 255.0f;
    }
    out.push_back(std::move(m0));
    while (out.back().w > 1 || out.back().h > 1) {
        const Mip &p = out.back();
        uint32_t w = std::max(1u, p.w / 2), h = std::max(1u, p.h / 2);
        Mip m{w, h, std::vector<float>((size_t)w * h * 4)};
        for (uint32_t y = 0; y < h; y++)
            for (uint32_t x = 0; x < w; x++)
                for (int c = 0; c < 4; c++) {
                    float s = 0.0f;
                    for (uint32_t dy = 0; dy < 2; dy++)
                        for (uint32_t dx = 0; dx < 2; dx++) {
                            uint32_t sx = std::min(2 * x + dx, p.w - 1), sy = std::min(2 * y + dy, p.h - 1);
                            s += p.t[4 * ((size_t)sy * p.w + sx) + c];
                        }
                    m.t[4 * ((size_t)y * w + x) + c] = 0.25f * s;
                }
        out.push_back(std::move(m));
    }
    return out;
}

// Level sizes of build_mips' chain: halve (at least 1) down to 1x1.
std::vector<std::pair<uint32_t, uint32_t>> mip_dims(uint32_t w, uint32_t h) {
    std::vector<std::pair<uint32_t, uint32_t>> d{{w, h}};
    while (d.back().first > 1 || d.back().second > 1)
        d.push_back({std::max(1u, d.back().first / 2), std::max(1u, d.back().second / 2)});
    return d;
}

uint32_t sampler_flags(const oxpt_sampler &s) {
    return (s.wrap_s & 3u) | (s.wrap_t & 3u) << 2 | (s.mag & 1u) << 4 | (s.min & 7u) << 5;
}

// Colour space of each material texture slot (src/pathtrace.rs:1042-1091):
// base, emissive, specular colour and sheen colour are sRGB.
constexpr bool kSlotSrgb[OXPT_N_TEX_SLOTS] = {true, false, false, true, false, true, false,
                                               false, false, false, false, true, false};

// material_words (src/pathtrace.rs:1338-1423): scalar part.
std::vector<uint32_t> material_words(const oxpt_material &m) {
    std::vector<uint32_t> w(OXPT_MAT_STRIDE, 0);
    auto put = [&](size_t i, float v) { w[i] = fb(v); };
    for (int k = 0; k < 4; k++) put(k, m.base_color[k]);
    for (int k = 0; k < 3; k++) put(4 + k, m.emissive[k]);
    uint32_t mode = m.alpha_mode == 1 ? 1u : (m.alpha_mode == 2 ? 2u : 0u);
    float cutoff = mode == 1 ? m.alpha_cutoff : 0.0f;
    put(8, m.metallic);
    put(9, m.roughness);
    put(10, m.normal_scale);
    put(11, cutoff);
    put(12, m.ior);
    auto fin = [](float v, float d) { return finitef(v) ? v : d; };
    float specular = 1.0f, sc[3] = {1.0f, 1.0f, 1.0f};
    if (m.has_specular) {
        specular = std::clamp(fin(m.specular, 1.0f), 0.0f, 1.0f);
        for (int k = 0; k < 3; k++) sc[k] = std::max(fin(m.specular_color[k], 1.0f), 0.0f);
    }
    put(13, specular);
    for (int k = 0; k < 3; k++) put(14 + k, sc[k]);
    put(17, m.has_transmission ? std::clamp(fin(m.transmission, 0.0f), 0.0f, 1.0f) : 0.0f);
    float thickness = 0.0f, ac[3] = {1.0f, 1.0f, 1.0f}, ad = -1.0f;
    if (m.has_volume) {
        thickness = std::max(fin(m.thickness, 0.0f), 0.0f);
        for (int k = 0; k < 3; k++) ac[k] = std::clamp(fin(m.attenuation_color[k], 1.0f), 0.0f, 1.0f);
        float d = m.attenuation_distance;
        if (finitef(d) && d > 0.0f) ad = d;
    }
    put(18, thickness);
    for (int k = 0; k < 3; k++) put(19 + k, ac[k]);
    put(22, ad);
    float cc = 0.0f, ccr = 0.0f, ccs = 1.0f;
    if (m.has_clearcoat) {
        cc = std::clamp(fin(m.clearcoat, 0.0f), 0.0f, 1.0f);
        ccr = std::clamp(fin(m.clearcoat_roughness, 0.0f), 0.0f, 1.0f);
        ccs = fin(m.clearcoat_normal_scale, 1.0f);
    }
    put(23, cc);
    put(24, ccr);
    put(25, ccs);
    float shc[3] = {0.0f, 0.0f, 0.0f}, shr = 0.0f;
    if (m.has_sheen) {
        for (int k = 0; k < 3; k++) shc[k] = std::clamp(fin(m.sheen_color[k], 0.0f), 0.0f, 1.0f);
        shr = std::clamp(fin(m.sheen_roughness, 0.0f), 0.0f, 1.0f);
    }
    for (int k = 0; k < 3; k++) put(26 + k, shc[k]);
    put(29, shr);
    bool emissive_lit = m.emissive[0] > 0.0f || m.emissive[1] > 0.0f || m.emissive[2] > 0.0f;
    w[7] = mode | (m.double_sided ? 1u : 0u) << 2 | (m.unlit ? 1u : 0u) << 3 | (emissive_lit ? 1u : 0u) << 4;
    return w;
}

// Atlas: every kept mip level shelf-packed (tallest first) into square
// layers (AtlasBuilder::pack, src/pathtrace.rs:1513-1592).
struct Rect {
    uint32_t layer, x, y, w, h;
};
struct Placed {
    uint32_t skip = 0;
    std::vector<Rect> rects;
    uint32_t levels = 0;
};

}  // namespace

void sobol_directions(uint32_t out[OXPT_SOBOL_WORDS]) {
    // Dimension 0 is the van der Corput sequence; dimensions 1-3 use the
    // Joe-Kuo (2008) direction numbers for the primitive polynomials
    // x + 1, x^2 + x + 1 and x^3 + x + 1 (initial m = 1 | 1,3 | 1,3,1),
    // extended with Bratley-Fox's recurrence on the integers m_k:
    // m_k = 2 a_1 m_{k-1} ^ 4 a_2 m_{k-2} ^ ... ^ 2^s m_{k-s} ^ m_{k-s},
    // then v_k = m_k << (32 - k).
    for (uint32_t k = 0; k < 32; k++) out[k] = 0x80000000u >> k;
    const uint32_t deg[3] = {1, 2, 3};
    const uint32_t coef[3] = {0, 1, 1};  // a_1..a_{s-1}, most significant first
    const uint32_t init[3][3] = {{1, 0, 0}, {1, 3, 0}, {1, 3, 1}};
    for (int d = 0; d < 3; d++) {
        uint32_t s = deg[d];
        uint64_t m[33] = {0};
        for (uint32_t k = 1; k <= s; k++) m[k] = init[d][k - 1];
        for (uint32_t k = s + 1; k <= 32; k++) {
            uint64_t v = m[k - s] ^ (m[k - s] << s);
            for (uint32_t j = 1; j < s; j++)
                if ((coef[d] >> (s - 1 - j)) & 1u) v ^= m[k - j] << j;
            m[k] = v;
        }
        for (uint32_t k = 1; k <= 32; k++) out[32 * (d + 1) + (k - 1)] = (uint32_t)(m[k] << (32 - k));
    }
}

namespace {
void compute_sheen_table(float out[OXPT_SHEEN_WORDS]) {
    // Directional albedo of the Charlie sheen lobe with Neubelt visibility,
    // E(mu, r) = int D(n.h; max(r^2, 1e-3)) V(n.l, n.v) (n.l) dw_l, on the
    // grid the kernel reads (mu = (i + 0.5) / 32, r = j / 15): midpoint rule
    // over (theta_l, phi_l) with the sin(theta_l) Jacobian, phi symmetric.
    const double PI = 3.14159265358979323846;
    const int NT = 192, NP = 192;
    for (uint32_t j = 0; j < OXPT_SHEEN_ROUGH; j++) {
        double r = (double)j / 15.0;
        double a = std::max(r * r, 1.0e-3);
        double inv = 1.0 / std::max(a, 1.0e-3);
        double norm = (2.0 + inv) / (2.0 * PI);
        for (uint32_t i = 0; i < OXPT_SHEEN_COS; i++) {
            double mu = (i + 0.5) / 32.0;
            double vs = std::sqrt(1.0 - mu * mu);
            double acc = 0.0;
            for (int ti = 0; ti < NT; ti++) {
                double th = (ti + 0.5) * (0.5 * PI / NT);
                double cl = std::cos(th), sl = std::sin(th);
                double vis = 1.0 / std::max(4.0 * (cl + mu - cl * mu), 1.0e-6);
                double row = 0.0;
                for (int pi = 0; pi < NP; pi++) {
                    double ph = (pi + 0.5) * (PI / NP);  // [0, pi], doubled below
                    double hx = vs + sl * std::cos(ph), hy = sl * std::sin(ph), hz = mu + cl;
                    double nh2 = hz * hz / (hx * hx + hy * hy + hz * hz);
                    double s2 = std::max(1.0 - nh2, 0.0);
                    row += s2 > 0.0 ? norm * std::pow(s2, 0.5 * inv) : 0.0;
                }
                acc += row * vis * cl * sl;
            }
            out[j * 32 + i] = (float)(acc * (0.5 * PI / NT) * (2.0 * PI / NP));
        }
    }
}
}  // namespace

void sheen_table(float out[OXPT_SHEEN_WORDS]) {
    // Computed once per process (thread-safe static initialisation).
    static const std::vector<float> lut = [] {
        std::vector<float> v(OXPT_SHEEN_WORDS);
        compute_sheen_table(v.data());
        return v;
    }();
    std::memcpy(out, lut.data(), OXPT_SHEEN_WORDS * sizeof(float));
}

std::vector<float> material_emission(const HostScene &s) {
    std::vector<float> out(s.materials.size(), 0.0f);
    for (size_t i = 0; i < s.materials.size(); i++) {
        const oxpt_material &m = s.materials[i];
        bool emits = !m.unlit && (m.emissive[0] > 0.0f || m.emissive[1] > 0.0f || m.emissive[2] > 0.0f);
        out[i] = emits ? luminance(m.emissive) : 0.0f;
    }
    return out;
}

void texel_decode_lut(float out[512]) {
    for (int i = 0; i < 256; i++) {
        out[i] = srgb_to_linear((float)i / 255.0f);
        out[256 + i] = (float)i / 255.0f;
    }
}

uint16_t f32_to_f16(float v) {
    // binary32 -> binary16, round to nearest even, overflow to infinity,
    // subnormals kept, NaN stays NaN (src/scene.rs:244-272).
    uint32_t x = fb(v);
    uint32_t sign = (x >> 16) & 0x8000u;
    uint32_t e8 = (x >> 23) & 0xffu, frac = x & 0x7fffffu;
    if (e8 == 0xffu) return (uint16_t)(sign | 0x7c00u | (frac ? 0x200u : 0u));
    int e = (int)e8 - 112;  // rebias 127 -> 15
    if (e >= 31) return (uint16_t)(sign | 0x7c00u);
    if (e <= 0) {
        if (e < -10) return (uint16_t)sign;
        uint32_t full = frac | 0x800000u;
        uint32_t sh = (uint32_t)(14 - e);
        uint32_t q = full >> sh, r = full & ((1u << sh) - 1u), halfway = 1u << (sh - 1);
        q += (r > halfway || (r == halfway && (q & 1u))) ? 1u : 0u;
        return (uint16_t)(sign | q);
    }
    uint32_t q = ((uint32_t)e << 10) | (frac >> 13), r = frac & 0x1fffu;
    q += (r > 0x1000u || (r == 0x1000u && (q & 1u))) ? 1u : 0u;
    return (uint16_t)(sign | q);
}

CameraSource camera_source(const HostScene &s) {
    CameraSource cs;
    cs.has_camera = s.has_camera;
    cs.camera = s.camera;
    cs.empty = s.soup.pos.empty();
    for (int k = 0; k < 3; k++) {
        cs.lo[k] = INFINITY;
        cs.hi[k] = -INFINITY;
    }
    for (size_t i = 0; i < s.soup.pos.size(); i += 3)
        for (int k = 0; k < 3; k++) {
            cs.lo[k] = std::min(cs.lo[k], s.soup.pos[i + k]);
            cs.hi[k] = std::max(cs.hi[k], s.soup.pos[i + k]);
        }
    return cs;
}

CameraFrame resolve_camera(const CameraSource &s, uint32_t w, uint32_t h) {
    oxpt_camera c = s.camera;
    if (!s.has_camera) {
        // Frame the scene bounds from +z (or a unit view of the origin).
        float ctr[3] = {0, 0, 0}, rad = 1.0f;
        if (!s.empty) {
            float r2 = 0;
            for (int k = 0; k < 3; k++) {
                ctr[k] = 0.5f * (s.lo[k] + s.hi[k]);
                r2 += 0.25f * (s.hi[k] - s.lo[k]) * (s.hi[k] - s.lo[k]);
            }
            rad = std::max(std::sqrt(r2), 1e-3f);
        }
        c.yfov = 0.8f;
        float dist = rad / std::sin(0.5f * c.yfov);
        c.eye[0] = ctr[0];
        c.eye[1] = ctr[1];
        c.eye[2] = ctr[2] + dist;
        for (int k = 0; k < 3; k++) c.target[k] = ctr[k];
        c.up[0] = 0;
        c.up[1] = 1;
        c.up[2] = 0;
        c.orthographic = 0;
    }
    CameraFrame f;
    float fwd[3] = {c.target[0] - c.eye[0], c.target[1] - c.eye[1], c.target[2] - c.eye[2]};
    float fl = std::sqrt(fwd[0] * fwd[0] + fwd[1] * fwd[1] + fwd[2] * fwd[2]);
    if (!(fl > 0.0f)) {
        fwd[0] = 0;
        fwd[1] = 0;
        fwd[2] = -1;
        fl = 1;
    }
    for (int k = 0; k < 3; k++) fwd[k] /= fl;
    float sd[3] = {fwd[1] * c.up[2] - fwd[2] * c.up[1], fwd[2] * c.up[0] - fwd[0] * c.up[2],
                   fwd[0] * c.up[1] - fwd[1] * c.up[0]};
    float sl = std::sqrt(sd[0] * sd[0] + sd[1] * sd[1] + sd[2] * sd[2]);
    if (!(sl > 1e-12f)) {
        sd[0] = 1;
        sd[1] = 0;
        sd[2] = 0;
        sl = 1;
    }
    for (int k = 0; k < 3; k++) sd[k] /= sl;
    float up[3] = {sd[1] * fwd[2] - sd[2] * fwd[1], sd[2] * fwd[0] - sd[0] * fwd[2], sd[0] * fwd[1] - sd[1] * fwd[0]};
    for (int k = 0; k < 3; k++) {
        f.eye[k] = c.eye[k];
        f.forward[k] = fwd[k];
        f.side[k] = sd[k];
        f.up[k] = up[k];
    }
    f.ortho = c.orthographic != 0;
    f.half_h = f.ortho ? c.ymag : std::tan(0.5f * c.yfov);
    f.half_w = f.half_h * (float)w / (float)std::max(h, 1u);
    return f;
}

Packed pack_scene(const HostScene &s, bool geometry, std::vector<uint32_t> *light_of_out, bool atlas,
                  bool emissive, const std::vector<Texture> *textures) {
    Packed pk;
    const std::vector<Texture> &tex_list = textures ? *textures : s.textures;
    const Soup &soup = s.soup;
    size_t ntri = soup.count();
    // Without `emissive` the lights are found by the device builder, which
    // appends their records to the tables.
    const size_t n_scan = emissive ? ntri : 0;
    // ---- Emissive lights (EmissiveLights::build): every triangle whose
    // material emits (not unlit, non-black factor) with positive area;
    // selection weight = area x luminance(emissive factor).
    std::vector<uint32_t> light_of(ntri, OXPT_NONE), elight_tris;
    std::vector<double> weights;
    for (size_t g = 0; g < n_scan; g++) {
        const oxpt_material &m = s.materials.at(soup.material[g]);
        bool emits = !m.unlit && (m.emissive[0] > 0.0f || m.emissive[1] > 0.0f || m.emissive[2] > 0.0f);
        if (!emits) continue;
        const float *p = &soup.pos[9 * g];
        double e1[3], e2[3];
        for (int k = 0; k < 3; k++) {
            e1[k] = (double)p[3 + k] - p[k];
            e2[k] = (double)p[6 + k] - p[k];
        }
        double cx = e1[1] * e2[2] - e1[2] * e2[1], cy = e1[2] * e2[0] - e1[0] * e2[2], cz = e1[0] * e2[1] - e1[1] * e2[0];
        double area = 0.5 * std::sqrt(cx * cx + cy * cy + cz * cz);
        double wgt = area * (double)luminance(m.emissive);
        if (!(wgt > 0.0)) continue;
        light_of[g] = (uint32_t)elight_tris.size();
        elight_tris.push_back((uint32_t)g);
        weights.push_back(wgt);
    }
    double total = 0.0;
    for (double w : weights) total += w;
    std::vector<float> cdf(weights.size()), pmf(weights.size());
    double run = 0.0;
    for (size_t i = 0; i < weights.size(); i++) {
        float prev = i ? cdf[i - 1] : 0.0f;
        run += weights[i];
        cdf[i] = i + 1 == weights.size() ? 1.0f : (float)(run / total);
        pmf[i] = cdf[i] - prev;
    }

    if (light_of_out) *light_of_out = light_of;
    // ---- BVH + triangles in leaf-slot order (or left to the device
    // builder: then the emissive records' slot words are OXPT_NONE until
    // the caller fills them in).
    std::vector<uint32_t> node_words, order;
    if (geometry) pk.bvh_depth = build_bvh(soup, node_words, order);
    pk.node_count = (uint32_t)(node_words.size() / 8);
    pk.tri_base = pk.node_count * 2;
    std::vector<uint32_t> slot_of(ntri, OXPT_NONE);
    for (size_t sl = 0; sl < order.size(); sl++) slot_of[order[sl]] = (uint32_t)sl;
    pk.geom.resize(node_words.size() + order.size() * 12);
    std::memcpy(pk.geom.data(), node_words.data(), node_words.size() * 4);
    pk.attrs.resize(order.size() * 48);
    auto fill = [&](size_t b, size_t e) {
        for (size_t sl = b; sl < e; sl++) {
            uint32_t g = order[sl];
            const float *p = &soup.pos[9 * (size_t)g];
            uint32_t ws[3] = {g, light_of[g], soup.material[g]};
            float *gw = &pk.geom[node_words.size() + 12 * sl];
            for (int v = 0; v < 3; v++) {
                gw[4 * v + 0] = p[3 * v + 0];
                gw[4 * v + 1] = p[3 * v + 1];
                gw[4 * v + 2] = p[3 * v + 2];
                std::memcpy(&gw[4 * v + 3], &ws[v], 4);
            }
            // push_attrs (src/pathtrace.rs:1275-1315).
            uint32_t fl = soup.flags[g];
            float *a = &pk.attrs[48 * sl];
            for (int v = 0; v < 3; v++) {
                for (int k = 0; k < 3; k++) a[4 * v + k] = (fl & OXPT_F_NORMALS) ? soup.normals[9 * (size_t)g + 3 * v + k] : 0.0f;
                float wv = 0.0f;
                if (v == 0) std::memcpy(&wv, &fl, 4);
                a[4 * v + 3] = wv;
            }
            for (int v = 0; v < 3; v++)
                for (int k = 0; k < 4; k++) a[12 + 4 * v + k] = (fl & OXPT_F_TANGENTS) ? soup.tangents[12 * (size_t)g + 4 * v + k] : 0.0f;
            for (int v = 0; v < 3; v++) {
                a[24 + 4 * v + 0] = (fl & OXPT_F_UV0) ? soup.uv0[6 * (size_t)g + 2 * v] : 0.0f;
                a[24 + 4 * v + 1] = (fl & OXPT_F_UV0) ? soup.uv0[6 * (size_t)g + 2 * v + 1] : 0.0f;
                a[24 + 4 * v + 2] = (fl & OXPT_F_UV1) ? soup.uv1[6 * (size_t)g + 2 * v] : 0.0f;
                a[24 + 4 * v + 3] = (fl & OXPT_F_UV1) ? soup.uv1[6 * (size_t)g + 2 * v + 1] : 0.0f;
            }
            for (int v = 0; v < 3; v++)
                for (int k = 0; k < 4; k++) a[36 + 4 * v + k] = (fl & OXPT_F_COLORS) ? soup.colors[12 * (size_t)g + 4 * v + k] : 1.0f;
        }
    };
    unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    if (order.size() > 65536 && hw > 1) {
        std::vector<std::thread> pool;
        size_t chunk = (order.size() + hw - 1) / hw;
        for (unsigned i = 0; i < hw; i++) {
            size_t b = i * chunk, e = std::min(order.size(), b + chunk);
            if (b < e) pool.emplace_back(fill, b, e);
        }
        for (auto &t : pool) t.join();
    } else {
        fill(0, order.size());
    }
    // ---- 4-wide nodes after the triangles, 128-byte aligned.
    if (!order.empty() && geometry) {
        std::vector<float> wide;
        pk.wide_depth = build_wide(node_words, wide);
        if (!wide.empty()) {
            size_t base = (pk.geom.size() / 4 + 7) / 8 * 8;
            pk.wide_base = (uint32_t)base;
            pk.geom.resize(4 * base, 0.0f);
            pk.geom.insert(pk.geom.end(), wide.begin(), wide.end());
        }
    }
    if (geometry) {
        if (pk.geom.empty()) pk.geom.resize(8, 0.0f);
        if (pk.attrs.empty()) pk.attrs.resize(48, 0.0f);
    }

    // ---- Tables.
    std::vector<uint32_t> &t = pk.tables;
    pk.off_sobol = (uint32_t)t.size();
    t.resize(t.size() + OXPT_SOBOL_WORDS);
    sobol_directions(&t[pk.off_sobol]);
    pk.off_sheen = (uint32_t)t.size();
    {
        float lut[OXPT_SHEEN_WORDS];
        sheen_table(lut);
        for (float v : lut) t.push_back(fb(v));
    }
    // Textures: one chain per (texture, colour space) a slot binds.
    std::map<std::pair<int, bool>, uint32_t> descs;
    for (const auto &m : s.materials)
        for (uint32_t sl = 0; sl < OXPT_N_TEX_SLOTS; sl++) {
            int ti = m.tex[sl].texture;
            if (ti >= 0 && (size_t)ti < tex_list.size() && tex_list[ti].w > 0) descs[{ti, kSlotSrgb[sl]}] = 0;
        }
    std::vector<std::pair<int, bool>> keys;
    for (auto &kv : descs) keys.push_back(kv.first);
    // Placement needs only the level sizes; the texels are decoded,
    // filtered and converted here (atlas) or by the device builder.
    std::vector<std::vector<std::pair<uint32_t, uint32_t>>> dims;
    std::vector<std::vector<Mip>> chains;
    for (auto &k : keys) {
        dims.push_back(mip_dims(tex_list[k.first].w, tex_list[k.first].h));
        if (atlas) chains.push_back(build_mips(tex_list[k.first], k.second));
    }
    uint32_t biggest = 1;
    for (auto &c : dims) biggest = std::max(biggest, std::max(c[0].first, c[0].second));
    uint32_t side = 1;
    while (side < biggest) side <<= 1;
    side = std::min(std::max(side, 1u), OXPT_MAX_LAYER);
    std::vector<Placed> placed(dims.size());
    struct Item {
        size_t ci, li;
        uint32_t w, h;
    };
    std::vector<Item> items;
    for (size_t ci = 0; ci < dims.size(); ci++) {
        uint32_t skip = 0;
        while (skip < dims[ci].size() && (dims[ci][skip].first > side || dims[ci][skip].second > side)) skip++;
        placed[ci].skip = skip;
        placed[ci].levels = (uint32_t)dims[ci].size();
        for (size_t li = skip; li < dims[ci].size(); li++)
            items.push_back({ci, li, dims[ci][li].first, dims[ci][li].second});
    }
    std::stable_sort(items.begin(), items.end(), [](const Item &a, const Item &b) {
        return a.h != b.h ? a.h > b.h : a.w > b.w;
    });
    uint32_t layer = 0, x = 0, y = 0, shelf = 0;
    std::vector<std::pair<std::pair<size_t, size_t>, Rect>> rects;
    for (const Item &it : items) {
        if (x + it.w > side) {
            x = 0;
            y += shelf;
            shelf = 0;
        }
        if (y + it.h > side) {
            layer++;
            x = 0;
            y = 0;
            shelf = 0;
        }
        rects.push_back({{it.ci, it.li}, Rect{layer, x, y, it.w, it.h}});
        x += it.w;
        shelf = std::max(shelf, it.h);
    }
    std::sort(rects.begin(), rects.end(), [](const auto &a, const auto &b) { return a.first < b.first; });
    pk.atlas_side = side;
    pk.atlas_layers = dims.empty() ? 1u : layer + 1;
    if (atlas) pk.atlas.assign((size_t)pk.atlas_layers * side * side * 4, 0);
    for (auto &pr : rects) {
        const Rect &r = pr.second;
        if (atlas) {
            const Mip &m = chains[pr.first.first][pr.first.second];
            for (uint32_t row = 0; row < r.h; row++)
                for (uint32_t col = 0; col < r.w; col++) {
                    size_t at = (((size_t)r.layer * side + r.y + row) * side + r.x + col) * 4;
                    for (int c = 0; c < 4; c++) pk.atlas[at + c] = f32_to_f16(m.t[4 * ((size_t)row * r.w + col) + c]);
                }
        }
        placed[pr.first.first].rects.push_back(r);
    }
    pk.chains.clear();
    for (size_t ci = 0; ci < keys.size(); ci++) {
        descs[keys[ci]] = (uint32_t)t.size();
        pk.chains.push_back({(uint32_t)keys[ci].first, keys[ci].second ? 1u : 0u, (uint32_t)t.size(), 0u});
        const Texture &tx = tex_list[keys[ci].first];
        t.insert(t.end(), {placed[ci].levels, sampler_flags(tx.sampler), tx.w, tx.h, placed[ci].skip,
                           (uint32_t)placed[ci].rects.size(), 0u, 0u});
        for (const Rect &r : placed[ci].rects) t.insert(t.end(), {r.layer, r.x, r.y, r.w | (r.h << 16)});
    }
    // Materials.
    pk.off_mat = (uint32_t)t.size();
    for (const auto &m : s.materials) {
        std::vector<uint32_t> w = material_words(m);
        for (uint32_t sl = 0; sl < OXPT_N_TEX_SLOTS; sl++) {
            uint32_t at = OXPT_M_SLOTS + 8 * sl;
            int ti = m.tex[sl].texture;
            auto it = descs.find({ti, kSlotSrgb[sl]});
            if (ti < 0 || it == descs.end()) {
                w[at] = OXPT_NONE;
                continue;
            }
            w[at] = it->second;
            w[at + 1] = m.tex[sl].uv_set;
            for (int k = 0; k < 6; k++) w[at + 2 + k] = fb(m.tex[sl].xform[k]);
        }
        t.insert(t.end(), w.begin(), w.end());
    }
    // Punctual lights (src/pathtrace.rs:1164-1200).
    pk.off_plights = (uint32_t)t.size();
    for (const auto &l : s.lights) {
        float base[3] = {l.color[0] * l.intensity, l.color[1] * l.intensity, l.color[2] * l.intensity};
        float cos_outer = std::cos(l.outer_cone), cos_inner = std::cos(l.inner_cone);
        float scale = 1.0f / std::max(cos_inner - cos_outer, 0.001f);
        float offset = -cos_outer * scale;
        uint32_t kind = l.kind > 2 ? 1u : l.kind;
        t.insert(t.end(), {fb(l.position[0]), fb(l.position[1]), fb(l.position[2]), kind, fb(l.direction[0]),
                           fb(l.direction[1]), fb(l.direction[2]), fb(l.range > 0.0f ? l.range : 0.0f), fb(base[0]),
                           fb(base[1]), fb(base[2]), 0u, fb(scale), fb(offset), 0u, 0u});
    }
    // Emissive triangles (src/pathtrace.rs:1201-1210).
    pk.off_elights = (uint32_t)t.size();
    for (size_t li = 0; li < elight_tris.size(); li++) {
        uint32_t g = elight_tris[li];
        t.insert(t.end(), {fb(cdf[li]), fb(pmf[li]), g, geometry ? slot_of[g] : OXPT_NONE});
    }
    pk.n_punctual = (uint32_t)s.lights.size();
    pk.n_emissive = (uint32_t)elight_tris.size();
    return pk;
}

EnvWords env_layout(uint32_t w, uint32_t h) {
    EnvWords out;
    out.w = std::max(w, 1u);
    out.h = std::max(h, 1u);
    out.marg = 0;
    out.cond = out.h + 1;
    out.mass = out.cond + out.h * (out.w + 1);
    out.rad = out.mass + out.w * out.h;
    return out;
}

std::vector<double> env_row_weights(uint32_t h) {
    const double PI = 3.14159265358979323846;
    h = std::max(h, 1u);
    std::vector<double> st(h);
    for (uint32_t y = 0; y < h; y++) st[y] = std::sin(PI * (y + 0.5) / h);
    return st;
}

void bind(const Packed &st, const EnvMap *env, const CameraFrame &cam, const oxpt_render_options &o, uint32_t w,
          uint32_t h, std::vector<uint32_t> &tables, oxpt_params &P, bool env_on_host) {
    tables = st.tables;
    uint32_t env_base = (uint32_t)tables.size();
    EnvWords ew;
    bool has_env = env != nullptr;
    if (has_env && env_on_host) {
        ew = env_words(*env);
        tables.insert(tables.end(), ew.words.begin(), ew.words.end());
    } else if (has_env) {
        ew = env_layout(env->w, env->h);  // the words follow `tables` on the device
    }
    if (tables.empty()) tables.push_back(0);
    // P_env: 0.5 with both an environment and emissive triangles, 1 with
    // only the environment, else 0 (src/pathtrace.rs:873-877).
    float p_env = has_env ? (st.n_emissive > 0 ? 0.5f : 1.0f) : 0.0f;
    std::memset(&P, 0, sizeof P);
    for (int k = 0; k < 3; k++) {
        P.eye[k] = cam.eye[k];
        P.forward[k] = cam.forward[k];
        P.side[k] = cam.side[k];
        P.up[k] = cam.up[k];
    }
    P.eye[3] = cam.ortho ? 1.0f : 0.0f;
    P.forward[3] = cam.half_w;
    P.side[3] = cam.half_h;
    P.up[3] = 2.0f * cam.half_h / (float)h;
    P.dims[0] = w;
    P.dims[1] = h;
    P.dims[2] = o.seed;
    P.dims[3] = o.max_bounces;
    P.cfg[0] = o.rr_start;
    P.cfg[1] = o.strategy <= 2 ? o.strategy : 0u;
    P.cfg[2] = st.n_punctual;
    P.cfg[3] = st.n_emissive;
    P.env[0] = has_env ? ew.w : 0;
    P.env[1] = has_env ? ew.h : 0;
    P.env[2] = has_env ? 1u : 0u;
    P.env[3] = st.node_count;
    P.offs0[0] = st.off_sobol;
    P.offs0[1] = st.off_sheen;
    P.offs0[2] = st.off_mat;
    P.offs0[3] = st.off_plights;
    P.offs1[0] = st.off_elights;
    P.offs1[2] = has_env ? env_base + ew.marg : 0;
    P.offs1[3] = has_env ? env_base + ew.cond : 0;
    P.offs2[0] = has_env ? env_base + ew.mass : 0;
    P.offs2[1] = has_env ? env_base + ew.rad : 0;
    P.offs2[2] = st.tri_base;
    P.offs2[3] = st.wide_base;
    P.fparams[0] = std::isfinite(o.ambient) ? std::max(o.ambient, 0.0f) : 0.0f;
    P.fparams[1] = o.clamp;
    P.fparams[2] = p_env;
}

EnvWords env_words(const EnvMap &e) {
    // EnvironmentMap sampling tables: texel weight = luminance(L) *
    // sin(theta at the row centre); marginal CDF over rows (h + 1 entries),
    // one conditional CDF per row (w + 1 entries), texel mass = joint
    // probability; radiance clamped >= 0 and scaled by the intensity
    // (src/pathtrace.rs:1428-1461).
    EnvWords out;
    uint32_t w = std::max(e.w, 1u), h = std::max(e.h, 1u);
    out.w = w;
    out.h = h;
    const std::vector<double> row_w = env_row_weights(h);
    std::vector<double> wt((size_t)w * h, 0.0), rows(h, 0.0);
    std::vector<float> rad((size_t)w * h * 3, 0.0f);
    for (uint32_t y = 0; y < h; y++) {
        double st = row_w[y];
        for (uint32_t x = 0; x < w; x++) {
            size_t i = (size_t)y * w + x;
            float c[3] = {0, 0, 0};
            if (i * 3 + 2 < e.rgb.size()) {
                bool ok = true;
                for (int k = 0; k < 3; k++) ok = ok && std::isfinite(e.rgb[3 * i + k]);
                if (ok)
                    for (int k = 0; k < 3; k++) c[k] = std::max(e.rgb[3 * i + k], 0.0f) * e.intensity;
            }
            for (int k = 0; k < 3; k++) rad[3 * i + k] = c[k];
            double l = (double)luminance(c) * st;
            wt[i] = l > 0.0 ? l : 0.0;
            rows[y] += wt[i];
        }
    }
    double total = 0.0;
    for (double r : rows) total += r;
    bool uniform = !(total > 0.0);
    std::vector<float> marg(h + 1), cond((size_t)h * (w + 1)), mass((size_t)w * h);
    double run = 0.0;
    marg[0] = 0.0f;
    for (uint32_t y = 0; y < h; y++) {
        run += uniform ? 1.0 / h : rows[y] / total;
        marg[y + 1] = y + 1 == h ? 1.0f : (float)run;
        float *c = &cond[(size_t)y * (w + 1)];
        c[0] = 0.0f;
        double rr = 0.0;
        bool row_uniform = uniform || !(rows[y] > 0.0);
        for (uint32_t x = 0; x < w; x++) {
            double p = row_uniform ? 1.0 / w : wt[(size_t)y * w + x] / rows[y];
            rr += p;
            c[x + 1] = x + 1 == w ? 1.0f : (float)rr;
            double rowp = uniform ? 1.0 / h : rows[y] / total;
            mass[(size_t)y * w + x] = (float)(rowp * p);
        }
    }
    auto push = [&](const std::vector<float> &v) {
        for (float f : v) out.words.push_back(fb(f));
    };
    out.marg = 0;
    push(marg);
    out.cond = (uint32_t)out.words.size();
    push(cond);
    out.mass = (uint32_t)out.words.size();
    push(mass);
    out.rad = (uint32_t)out.words.size();
    push(rad);
    return out;
}

}  // namespace oxpt
