// Device-side shading: surface, material, layered BSDF, lights and the
// environment (oxideav-render-vulkan src/shaders/pathtrace.wgsl:765-1654).
#pragma once

#include "pt_device.hpp"

namespace oxd {

struct Surface {
    F3 position, ng, sn;
};

__device__ __forceinline__ Surface surface(const Scene &S, uint32_t slot, F3 b) {
    F3 p0 = tri_pos(S, slot, 0), p1 = tri_pos(S, slot, 1), p2 = tri_pos(S, slot, 2);
    Surface s;
    s.position = p0 * b.x + p1 * b.y + p2 * b.z;
    F3 cr = cross(p1 - p0, p2 - p0);
    float len = sqrtf(dot(cr, cr));
    s.ng = len > 0.0f ? cr / len : F3{0.0f, 0.0f, 1.0f};
    s.sn = s.ng;
    if (attr_flags(S, slot) & OXPT_F_NORMALS) {
        F3 n = nrm(interp3a(S, slot, 0, b));
        if (dot(n, n) > 0.5f) s.sn = n;
    }
    return s;
}

struct Mat {
    float4 base;
    float metallic, roughness;
    F3 normal;
    bool unlit;
    F3 emission;
    float ior, specular;
    F3 specular_color;
    float transmission, thickness;
    F3 atten_color;
    float atten_dist, clearcoat, cc_rough;
    F3 cc_normal, sheen;
    float sheen_rough;
};

__device__ F3 perturb(const Scene &S, uint32_t slot, F3 b, F3 n, float4 t, float scale) {
    if ((attr_flags(S, slot) & OXPT_F_TANGENTS) == 0) return n;
    float4 tg = interp4a(S, slot, 3, b);
    float w = attr(S, slot, 3).w < 0.0f ? -1.0f : 1.0f;
    F3 t3 = xyz(tg);
    float d = dot(n, t3);
    F3 tt = nrm(t3 - n * d);
    if (dot(tt, tt) < 0.5f) return n;
    F3 bt = cross(n, tt) * w;
    F3 ts = {(t.x * 2.0f - 1.0f) * scale, (t.y * 2.0f - 1.0f) * scale, t.z * 2.0f - 1.0f};
    F3 nn = nrm(tt * ts.x + bt * ts.y + n * ts.z);
    return dot(nn, nn) > 0.5f ? nn : n;
}

// Material at a hit (wgsl:842-902). Texture slots are sampled only when
// present; absent slots are the white no-op factor.
__device__ Mat material(const Scene &S, uint32_t slot, F3 b, F3 sn_in, const Lod &lod, bool flip) {
    F3 sn = flip ? -sn_in : sn_in;
    uint32_t m = tri_material(S, slot), o = mat_off(S, m);
    float4 tex[13];
    uint32_t present = 0;
    const float4 white = make_float4(1.0f, 1.0f, 1.0f, 1.0f);
    uint32_t uvflags = attr_flags(S, slot) & (OXPT_F_UV0 | OXPT_F_UV1);
#pragma unroll
    for (uint32_t si = 0; si < 13; si++) tex[si] = white;
    if (uvflags) {
        for (uint32_t si = 0; si < 13; si++) {
            if (slot_present(S, m, si, slot)) {
                present |= 1u << si;
                Lod l = si == OXPT_S_EMISSIVE ? base_lod() : lod;
                float4 v = slot_sample(S, m, si, slot, b, l);
                // Dynamic indexing into a register array: select per slot.
#pragma unroll
                for (uint32_t k = 0; k < 13; k++)
                    if (k == si) tex[k] = v;
            }
        }
    }
    gu32_ptr mw = S.tables + o;
    Mat r;
    float4 col = mul4(make_float4(bitsf(mw[0]), bitsf(mw[1]), bitsf(mw[2]), bitsf(mw[3])), tex[OXPT_S_BASE]);
    if (attr_flags(S, slot) & OXPT_F_COLORS) col = mul4(col, interp4a(S, slot, 9, b));
    r.base = col;
    r.metallic = clampf(bitsf(mw[8]) * tex[OXPT_S_MR].z, 0.0f, 1.0f);
    r.roughness = clampf(bitsf(mw[9]) * tex[OXPT_S_MR].y, 0.0f, 1.0f);
    r.normal = sn;
    if (present & (1u << OXPT_S_NORMAL)) r.normal = perturb(S, slot, b, sn, tex[OXPT_S_NORMAL], bitsf(mw[10]));
    r.unlit = (mw[OXPT_M_FLAGS] & 8u) != 0;
    r.emission = f3(0.0f);
    F3 e = {bitsf(mw[4]), bitsf(mw[5]), bitsf(mw[6])};
    if (!r.unlit && !is_black(e)) r.emission = e * xyz(tex[OXPT_S_EMISSIVE]);
    r.ior = bitsf(mw[12]);
    r.specular = bitsf(mw[13]) * tex[OXPT_S_SPEC].w;
    r.specular_color = F3{bitsf(mw[14]), bitsf(mw[15]), bitsf(mw[16])} * xyz(tex[OXPT_S_SPEC_COLOR]);
    r.transmission = bitsf(mw[17]) * tex[OXPT_S_TRANSMISSION].x;
    r.thickness = bitsf(mw[18]) * tex[OXPT_S_THICKNESS].y;
    r.atten_color = {bitsf(mw[19]), bitsf(mw[20]), bitsf(mw[21])};
    r.atten_dist = bitsf(mw[22]);
    r.clearcoat = bitsf(mw[23]) * tex[OXPT_S_CC].x;
    r.cc_rough = bitsf(mw[24]) * tex[OXPT_S_CC_ROUGH].y;
    r.cc_normal = sn;
    if (present & (1u << OXPT_S_CC_NORMAL)) r.cc_normal = perturb(S, slot, b, sn, tex[OXPT_S_CC_NORMAL], bitsf(mw[25]));
    r.sheen = F3{bitsf(mw[26]), bitsf(mw[27]), bitsf(mw[28])} * xyz(tex[OXPT_S_SHEEN_COLOR]);
    r.sheen_rough = bitsf(mw[29]) * tex[OXPT_S_SHEEN_ROUGH].w;
    return r;
}

// Emitted radiance of a hit at level 0 (wgsl:905-916).
__device__ F3 emission(const Scene &S, uint32_t slot, F3 b) {
    uint32_t m = tri_material(S, slot), o = mat_off(S, m);
    F3 e = tv3(S, o + 4);
    if ((S.tables[o + OXPT_M_FLAGS] & 8u) != 0 || is_black(e)) return f3(0.0f);
    if (slot_present(S, m, OXPT_S_EMISSIVE, slot)) return e * xyz(slot_sample(S, m, OXPT_S_EMISSIVE, slot, b, base_lod()));
    return e;
}

// ---- BSDF (wgsl:922-1380) ----------------------------------------------

__device__ __forceinline__ float d_ggx(float n_dot_h, float a) {
    if (n_dot_h <= 0.0f) return 0.0f;
    float a2 = a * a;
    float d = n_dot_h * n_dot_h * (a2 - 1.0f) + 1.0f;
    return a2 / (PI * d * d);
}
__device__ __forceinline__ float lambda(float c_in, float a) {
    float c = fmaxf(fabsf(c_in), 1.0e-7f);
    float a2 = a * a;
    return (sqrtf(a2 + (1.0f - a2) * c * c) / c - 1.0f) * 0.5f;
}
__device__ __forceinline__ float g1(float c, float a) { return 1.0f / (1.0f + lambda(c, a)); }
__device__ __forceinline__ float g2(float cv, float cl, float a) { return 1.0f / (1.0f + lambda(cv, a) + lambda(cl, a)); }
__device__ __forceinline__ F3 schlick(F3 f0, float c) {
    float x = 1.0f - clampf(c, 0.0f, 1.0f);
    float k = x * x * x * x * x;
    return f0 + (f3(1.0f) - f0) * k;
}
__device__ F3 sample_vndf(F3 v, float a, float u0, float u1) {
    F3 vh = nrm({a * v.x, a * v.y, v.z});
    float lensq = vh.x * vh.x + vh.y * vh.y;
    F3 t1 = {1.0f, 0.0f, 0.0f};
    if (lensq > 0.0f) {
        float il = 1.0f / sqrtf(lensq);
        t1 = {-vh.y * il, vh.x * il, 0.0f};
    }
    F3 t2 = cross(vh, t1);
    float r = sqrtf(u0);
    float phi = 2.0f * PI * u1;
    float p1 = r * cosf(phi);
    float p2 = r * sinf(phi);
    float s = 0.5f * (1.0f + vh.z);
    p2 = (1.0f - s) * sqrtf(fmaxf(1.0f - p1 * p1, 0.0f)) + s * p2;
    float pz = sqrtf(fmaxf(1.0f - p1 * p1 - p2 * p2, 0.0f));
    F3 nh = t1 * p1 + t2 * p2 + vh * pz;
    return nrm({a * nh.x, a * nh.y, fmaxf(nh.z, 1.0e-6f)});
}
__device__ __forceinline__ float vndf_reflect_pdf(F3 n, F3 v, F3 h, float a) {
    float nv = dot(n, v), nh = dot(n, h);
    if (nv <= 0.0f || nh <= 0.0f || dot(v, h) <= 0.0f) return 0.0f;
    return g1(nv, a) * d_ggx(nh, a) / (4.0f * nv);
}
__device__ __forceinline__ F3 reflect_h(F3 v, F3 h) { return 2.0f * dot(v, h) * h - v; }
__device__ __forceinline__ float d_charlie(float n_dot_h, float a) {
    float inv = 1.0f / fmaxf(a, 1.0e-3f);
    float sin2 = fmaxf(1.0f - n_dot_h * n_dot_h, 0.0f);
    if (sin2 <= 0.0f) return 0.0f;
    return (2.0f + inv) * powf(sin2, inv * 0.5f) / (2.0f * PI);
}
__device__ __forceinline__ float v_neubelt(float nl, float nv) { return 1.0f / fmaxf(4.0f * (nl + nv - nl * nv), 1.0e-6f); }

__device__ __forceinline__ float sheen_albedo(const float *lut, float c, float rough) {
    const uint32_t SC = OXPT_SHEEN_COS, SR = OXPT_SHEEN_ROUGH;
    float x = clampf(clampf(c, 0.0f, 1.0f) * (float)SC - 0.5f, 0.0f, (float)(SC - 1));
    float y = clampf(clampf(rough, 0.0f, 1.0f) * (float)(SR - 1), 0.0f, (float)(SR - 1));
    uint32_t x0 = (uint32_t)floorf(x), y0 = (uint32_t)floorf(y);
    uint32_t x1 = min(x0 + 1u, SC - 1u), y1 = min(y0 + 1u, SR - 1u);
    float fx = x - (float)x0, fy = y - (float)y0;
    float a00 = lut[y0 * SC + x0], a10 = lut[y0 * SC + x1], a01 = lut[y1 * SC + x0], a11 = lut[y1 * SC + x1];
    float a = a00 + (a10 - a00) * fx;
    float bb = a01 + (a11 - a01) * fx;
    return a + (bb - a) * fy;
}
__device__ __forceinline__ F3 basis_t(F3 n) {
    float sign = sign_of(n.z);
    float a = -1.0f / (sign + n.z);
    float b = n.x * n.y * a;
    return {1.0f + sign * n.x * n.x * a, sign * b, -sign * n.x};
}
__device__ __forceinline__ F3 basis_b(F3 n) {
    float sign = sign_of(n.z);
    float a = -1.0f / (sign + n.z);
    float b = n.x * n.y * a;
    return {b, sign + n.y * n.y * a, -n.y};
}
__device__ __forceinline__ F3 cosine_hemisphere(float u0, float u1) {
    float a = 2.0f * u0 - 1.0f, b = 2.0f * u1 - 1.0f;
    float r = 0.0f, phi = 0.0f;
    if (a == 0.0f && b == 0.0f) {
        r = 0.0f;
        phi = 0.0f;
    } else if (fabsf(a) > fabsf(b)) {
        r = a;
        phi = (PI / 4.0f) * (b / a);
    } else {
        r = b;
        phi = PI / 2.0f - (PI / 4.0f) * (a / b);
    }
    float x = r * cosf(phi), y = r * sinf(phi);
    return {x, y, sqrtf(fmaxf(1.0f - x * x - y * y, 0.0f))};
}

struct Bsdf {
    F3 v, n, ng, ncc, base;
    float metallic, alpha, specular;
    F3 f0d;
    float transmission;
    bool thin;
    float eta_i, eta_t, clearcoat, cc_alpha;
    F3 sheen;
    float sheen_rough, sheen_alpha, w_cc;
    float p0, p1, p2, p3, p4;  // diffuse, specular, transmission, clearcoat, sheen
};

__device__ __forceinline__ F3 fresnel_d(const Bsdf &bs, float c) {
    if (bs.eta_i > bs.eta_t) {
        float eta = bs.eta_i / bs.eta_t;
        float s2 = eta * eta * fmaxf(1.0f - c * c, 0.0f);
        if (s2 >= 1.0f) return f3(1.0f);
        return schlick(bs.f0d, sqrtf(1.0f - s2));
    }
    return schlick(bs.f0d, c);
}

__device__ Bsdf make_bsdf(const float *lut, const Mat &m, F3 v, F3 ng_in, bool front) {
    float fl = front ? 1.0f : -1.0f;
    F3 ng = ng_in * fl;
    F3 n = m.normal * fl;
    if (dot(n, v) <= 1.0e-4f) n = ng;
    F3 ncc = m.cc_normal * fl;
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
    b.alpha = fmaxf(m.roughness * m.roughness, MIN_ALPHA);
    b.specular = clampf(m.specular, 0.0f, 1.0f);
    b.f0d = vmin(r0 * m.specular_color, f3(1.0f));
    b.transmission = m.transmission;
    b.thin = thin;
    b.eta_i = eta_i;
    b.eta_t = eta_t;
    b.clearcoat = m.clearcoat;
    b.sheen = m.sheen;
    if (inside) {
        b.clearcoat = 0.0f;
        b.sheen = f3(0.0f);
    }
    b.cc_alpha = fmaxf(m.cc_rough * m.cc_rough, MIN_ALPHA);
    b.sheen_rough = m.sheen_rough;
    b.sheen_alpha = fmaxf(m.sheen_rough * m.sheen_rough, 1.0e-3f);
    float cv = fmaxf(dot(n, v), 1.0e-4f);
    b.w_cc = 0.0f;
    if (b.clearcoat > 0.0f) b.w_cc = b.clearcoat * schlick(f3(0.04f), dot(ncc, v)).x;
    float sheen_max = max3(b.sheen);
    float e_v = 0.0f;
    if (sheen_max > 0.0f) e_v = sheen_albedo(lut, cv, b.sheen_rough);
    float s_scale = (1.0f - b.w_cc) * (1.0f - sheen_max * e_v);
    float fd = b.specular * max3(fresnel_d(b, cv));
    float fm = mean3(schlick(b.base, cv));
    float mm = b.metallic;
    float base_mean = fmaxf(mean3(b.base), 0.0f);
    float p0 = s_scale * (1.0f - mm) * (1.0f - b.transmission) * (1.0f - fd) * base_mean;
    float p1 = s_scale * ((1.0f - mm) * fd + mm * fm);
    float p2 = s_scale * (1.0f - mm) * b.transmission * (1.0f - fd) * base_mean;
    float p3 = b.w_cc;
    float p4 = (1.0f - b.w_cc) * sheen_max * e_v;
    float sum = p0 + p1 + p2 + p3 + p4;
    if (sum > 0.0f && finite(sum)) {
        p0 = p0 / sum;
        p1 = p1 / sum;
        p2 = p2 / sum;
        p3 = p3 / sum;
        p4 = p4 / sum;
    } else {
        p0 = p1 = p2 = p3 = p4 = 0.0f;
    }
    b.p0 = p0;
    b.p1 = p1;
    b.p2 = p2;
    b.p3 = p3;
    b.p4 = p4;
    return b;
}

__device__ __forceinline__ bool bsdf_black(const Bsdf &b) {
    return b.p0 <= 0.0f && b.p1 <= 0.0f && b.p2 <= 0.0f && b.p3 <= 0.0f && b.p4 <= 0.0f;
}
__device__ __forceinline__ float lobe_p(const Bsdf &b, uint32_t i) {
    return i == 0 ? b.p0 : (i == 1 ? b.p1 : (i == 2 ? b.p2 : (i == 3 ? b.p3 : b.p4)));
}

struct Eval {
    F3 f;
    float pdf;
};

// f(v, l) |n.l| and the mixture pdf of l (wgsl:1190-1284).
__device__ Eval bsdf_eval(const float *lut, const Bsdf &b, F3 l) {
    F3 v = b.v, n = b.n;
    float nv = fmaxf(dot(n, v), 1.0e-4f);
    if (dot(b.ng, l) > 0.0f) {
        float nl = dot(n, l);
        F3 h = nrm(v + l);
        float vh = fmaxf(dot(v, h), 0.0f);
        float nh = dot(n, h);
        F3 f = f3(0.0f);
        float pdf = 0.0f;
        if (nl > 0.0f) {
            F3 fd = fresnel_d(b, vh);
            F3 fm = schlick(b.base, vh);
            float spec_c = d_ggx(nh, b.alpha) * g2(nv, nl, b.alpha) / (4.0f * nv * nl);
            float m = b.metallic;
            float diff_k = (1.0f - m) * (1.0f - b.transmission) * (1.0f - b.specular * max3(fd)) / PI;
            F3 layer = spec_c * ((1.0f - m) * b.specular * fd + m * fm) + diff_k * b.base;
            float smax = max3(b.sheen);
            if (smax > 0.0f) {
                float scale = fminf(1.0f - smax * sheen_albedo(lut, nv, b.sheen_rough),
                                    1.0f - smax * sheen_albedo(lut, nl, b.sheen_rough));
                float sh = d_charlie(nh, b.sheen_alpha) * v_neubelt(nl, nv);
                layer = layer * scale + b.sheen * sh;
            }
            f = layer * (1.0f - b.w_cc);
            pdf += (b.p0 + b.p4) * nl / PI;
            pdf += b.p1 * vndf_reflect_pdf(n, v, h, b.alpha);
        }
        if (b.clearcoat > 0.0f) {
            float cnv = fmaxf(dot(b.ncc, v), 1.0e-4f);
            float cnl = dot(b.ncc, l);
            if (cnl > 0.0f) {
                float cnh = dot(b.ncc, h);
                float c = b.w_cc * d_ggx(cnh, b.cc_alpha) * g2(cnv, cnl, b.cc_alpha) / (4.0f * cnv * cnl);
                f = f + f3(c);
                pdf += b.p3 * vndf_reflect_pdf(b.ncc, v, h, b.cc_alpha);
            }
        }
        return {f * fmaxf(nl, 0.0f), pdf};
    }
    if (b.transmission <= 0.0f || b.metallic >= 1.0f) return {f3(0.0f), 0.0f};
    float nl = -dot(n, l);
    if (nl <= 0.0f) return {f3(0.0f), 0.0f};
    float smax = max3(b.sheen);
    float sheen_scale = 1.0f;
    if (smax > 0.0f) sheen_scale = 1.0f - smax * sheen_albedo(lut, nv, b.sheen_rough);
    float w = (1.0f - b.metallic) * b.transmission * (1.0f - b.w_cc) * sheen_scale;
    float val = 0.0f, pdf_t = 0.0f;
    if (b.thin) {
        F3 lm = l + n * (2.0f * nl);
        F3 h = nrm(v + lm);
        float vh = dot(v, h), nh = dot(n, h);
        if (vh <= 0.0f || nh <= 0.0f) return {f3(0.0f), 0.0f};
        float fd = b.specular * max3(fresnel_d(b, vh));
        val = d_ggx(nh, b.alpha) * g2(nv, nl, b.alpha) / (4.0f * nv * nl) * (1.0f - fd);
        pdf_t = vndf_reflect_pdf(n, v, h, b.alpha);
    } else {
        float ei = b.eta_i, et = b.eta_t;
        F3 h = nrm(-(ei * v + et * l));
        if (dot(h, n) < 0.0f) h = -h;
        float vh = dot(v, h), lh = dot(l, h), nh = dot(n, h);
        if (vh <= 0.0f || lh >= 0.0f || nh <= 0.0f) return {f3(0.0f), 0.0f};
        float denom = ei * vh + et * lh;
        if (fabsf(denom) < 1.0e-6f) return {f3(0.0f), 0.0f};
        float d = d_ggx(nh, b.alpha);
        float fd = b.specular * max3(fresnel_d(b, vh));
        float jac = et * et * fabsf(lh) / (denom * denom);
        val = vh * fabsf(lh) * d * g2(nv, nl, b.alpha) * et * et * (1.0f - fd) / (nv * nl * denom * denom);
        pdf_t = g1(nv, b.alpha) * vh * d / nv * jac;
    }
    return {b.base * (w * val * nl), b.p2 * pdf_t};
}

struct BSample {
    bool ok;
    F3 l, f;
    float pdf;
};

__device__ __forceinline__ bool refract_h(F3 v, F3 h, float eta, F3 *out) {
    float ci = dot(v, h);
    float s2 = eta * eta * fmaxf(1.0f - ci * ci, 0.0f);
    if (s2 >= 1.0f) return false;
    float ct = sqrtf(1.0f - s2);
    float k = eta * ci - ct;
    *out = nrm(-eta * v + k * h);
    return true;
}

__device__ BSample bsdf_sample(const float *lut, const Bsdf &b, F3 u) {
    BSample res;
    res.ok = false;
    res.l = f3(0.0f);
    res.f = f3(0.0f);
    res.pdf = 0.0f;
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
    F3 v = b.v, l;
    if (lobe == 0 || lobe == 4) {
        F3 t = basis_t(b.n), bb = basis_b(b.n);
        F3 c = cosine_hemisphere(u.x, u.y);
        l = t * c.x + bb * c.y + b.n * c.z;
    } else {
        F3 nn = lobe == 3 ? b.ncc : b.n;
        float a = lobe == 3 ? b.cc_alpha : b.alpha;
        F3 t = basis_t(nn), bb = basis_b(nn);
        F3 vl = {dot(v, t), dot(v, bb), dot(v, nn)};
        if (vl.z <= 0.0f) return res;
        F3 hl = sample_vndf(vl, a, u.x, u.y);
        F3 h = t * hl.x + bb * hl.y + nn * hl.z;
        if (lobe != 2) {
            l = reflect_h(v, h);
        } else if (b.thin) {
            F3 r = reflect_h(v, h);
            float d = dot(r, b.n);
            l = r + b.n * (-2.0f * d);
        } else {
            F3 rl;
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

// ---- Lights (wgsl:1386-1654) -------------------------------------------

struct LightSample {
    bool ok;
    F3 l;
    float distance;
    F3 radiance;
};

__device__ LightSample punctual_sample(const Scene &S, uint32_t j, F3 p) {
    gu32_ptr lw = S.tables + S.off_plights + 16u * j;
    uint4 w0 = ldgu4(lw), w1 = ldgu4(lw + 4), w2 = ldgu4(lw + 8), w3 = ldgu4(lw + 12);
    LightSample r;
    r.ok = false;
    r.l = f3(0.0f);
    r.distance = 0.0f;
    r.radiance = f3(0.0f);
    uint32_t kind = w0.w;
    F3 base = {bitsf(w2.x), bitsf(w2.y), bitsf(w2.z)};
    F3 dir = {bitsf(w1.x), bitsf(w1.y), bitsf(w1.z)};
    if (kind == 0) {
        r.ok = true;
        r.l = -dir;
        r.distance = -1.0f;
        r.radiance = base;
        return r;
    }
    F3 to = F3{bitsf(w0.x), bitsf(w0.y), bitsf(w0.z)} - p;
    float d2 = dot(to, to);
    if (!(d2 > 1.0e-12f)) return r;
    float d = sqrtf(d2);
    F3 l = to / d;
    float att = 1.0f / d2;
    float range = bitsf(w1.w);
    if (range > 0.0f) {
        float x = d / range;
        att *= clampf(1.0f - x * x * x * x, 0.0f, 1.0f);
    }
    if (kind == 2) {
        float cd = dot(dir, -l);
        float a = clampf(cd * bitsf(w3.x) + bitsf(w3.y), 0.0f, 1.0f);
        att *= a * a;
    }
    if (att <= 0.0f) return r;
    r.ok = true;
    r.l = l;
    r.distance = d;
    r.radiance = base * att;
    return r;
}

// partition_point(c <= u) over n ascending f32 words at o (wgsl:1440-1455).
__device__ __forceinline__ uint32_t partition_point(const Scene &S, uint32_t o, uint32_t n, float u) {
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        uint32_t mid = (lo + hi) / 2u;
        if (tf(S, o + mid) <= u) lo = mid + 1u;
        else hi = mid;
    }
    return lo;
}
__device__ __forceinline__ float elight_pmf(const Scene &S, uint32_t li) { return tf(S, S.off_elights + 4u * li + 1u); }
__device__ __forceinline__ uint32_t elight_slot(const Scene &S, uint32_t li) { return S.tables[S.off_elights + 4u * li + 3u]; }
// Emissive-light CDF index (wgsl:1466-1484): first record whose cdf > u.
__device__ __forceinline__ uint32_t pick_elight(const Scene &S, float u) {
    uint32_t n = S.n_emissive, lo = 0, hi = n;
    while (lo < hi) {
        uint32_t mid = (lo + hi) / 2u;
        const float c = S.lds_ecdf ? S.lds_ecdf[mid] : tf(S, S.off_elights + 4u * mid);
        if (c <= u) lo = mid + 1u;
        else hi = mid;
    }
    return min(lo, n - 1u);
}
// Marginal CDF word y (the LDS copy when staged).
__device__ __forceinline__ float env_marg_at(const Scene &S, uint32_t y) {
    return S.lds_marg ? S.lds_marg[y] : tf(S, S.env_marg + y);
}
__device__ __forceinline__ uint32_t env_row(const Scene &S, float u0) {
    uint32_t lo = 0, hi = S.env_h + 1u;  // partition_point(c <= u0), as above
    while (lo < hi) {
        uint32_t mid = (lo + hi) / 2u;
        if (env_marg_at(S, mid) <= u0) lo = mid + 1u;
        else hi = mid;
    }
    return min(max(lo, 1u), S.env_h) - 1u;
}
__device__ __forceinline__ uint32_t env_col(const Scene &S, uint32_t y, float u1) {
    uint32_t row = S.env_cond + y * (S.env_w + 1u);
    return min(max(partition_point(S, row, S.env_w + 1u, u1), 1u), S.env_w) - 1u;
}

__device__ float triangle_solid_angle(F3 p, F3 a0, F3 b0, F3 c0) {
    F3 a = nrm_any(a0 - p), b = nrm_any(b0 - p), c = nrm_any(c0 - p);
    float num = fabsf(dot(a, cross(b, c)));
    float den = 1.0f + dot(a, b) + dot(b, c) + dot(c, a);
    float o = 2.0f * atan2f(num, den);
    return finite(o) ? fmaxf(o, 0.0f) : 0.0f;
}

// Arvo 1995 uniform spherical-triangle sampling (wgsl:1510-1556).
__device__ F3 sample_spherical_triangle(F3 p, F3 a0, F3 b0, F3 c0, float u0, float u1) {
    F3 zero = f3(0.0f);
    F3 a = nrm_any(a0 - p), b = nrm_any(b0 - p), c = nrm_any(c0 - p);
    F3 nab = nrm_any(cross(a, b)), nbc = nrm_any(cross(b, c)), nca = nrm_any(cross(c, a));
    if (dot(a, a) == 0.0f || dot(b, b) == 0.0f || dot(c, c) == 0.0f || dot(nab, nab) == 0.0f ||
        dot(nbc, nbc) == 0.0f || dot(nca, nca) == 0.0f)
        return zero;
    float alpha = acosf(clampf(-dot(nab, nca), -1.0f, 1.0f));
    float beta = acosf(clampf(-dot(nbc, nab), -1.0f, 1.0f));
    float gamma = acosf(clampf(-dot(nca, nbc), -1.0f, 1.0f));
    float area = alpha + beta + gamma - PI;
    if (!(area > 0.0f)) return zero;
    float ap = u0 * area;
    float s = sinf(ap - alpha), t = cosf(ap - alpha);
    float sa = sinf(alpha), ca = cosf(alpha);
    float cos_c = dot(a, b);
    float uu = t - ca, vv = s + sa * cos_c;
    float den = (vv * s + uu * t) * sa;
    if (fabsf(den) < 1.0e-30f) return zero;
    float q = clampf(((vv * t - uu * s) * ca - vv) / den, -1.0f, 1.0f);
    F3 ca_perp = nrm_any(c - dot(c, a) * a);
    if (dot(ca_perp, ca_perp) == 0.0f) return zero;
    float r = sqrtf(fmaxf(1.0f - q * q, 0.0f));
    F3 cp = q * a + r * ca_perp;
    float z = 1.0f - u1 * (1.0f - dot(cp, b));
    F3 cb_perp = nrm_any(cp - dot(cp, b) * b);
    float w = sqrtf(fmaxf(1.0f - z * z, 0.0f));
    F3 d = b;
    if (dot(cb_perp, cb_perp) > 0.0f) d = z * b + w * cb_perp;
    return nrm_any(d);
}

// tri_pdf given the light's vertices and their solid angle from p (the NEE
// sampler has both already).
__device__ float tri_pdf_omega(const Scene &S, uint32_t li, F3 a, F3 b, F3 c, float omega, F3 p, F3 q) {
    float sel = (1.0f - S.p_env) * elight_pmf(S, li);
    if (omega >= MIN_SPHERICAL_SOLID_ANGLE) return sel / omega;
    F3 cr = cross(b - a, c - a);
    float len = sqrtf(dot(cr, cr));
    F3 to = q - p;
    float d2 = dot(to, to);
    if (len <= 0.0f || d2 <= 0.0f) return 0.0f;
    float cs = fabsf(dot(cr, to) / (len * sqrtf(d2)));
    if (cs <= 0.0f) return 0.0f;
    return sel * d2 / (0.5f * len * cs);
}
__device__ float tri_pdf(const Scene &S, uint32_t li, uint32_t slot, F3 p, F3 q) {
    F3 a = tri_pos(S, slot, 0), b = tri_pos(S, slot, 1), c = tri_pos(S, slot, 2);
    return tri_pdf_omega(S, li, a, b, c, triangle_solid_angle(p, a, b, c), p, q);
}

struct Texel {
    uint32_t x, y;
    float st;
};
__device__ __forceinline__ Texel env_texel(const Scene &S, F3 d) {
    uint32_t w = S.env_w, h = S.env_h;
    float u = 0.5f + atan2f(d.x, -d.z) / (2.0f * PI);
    float th = acosf(clampf(d.y, -1.0f, 1.0f));
    float v = th / PI;
    uint32_t x = min((uint32_t)fmaxf(u * (float)w, 0.0f), w - 1u);
    uint32_t y = min((uint32_t)fmaxf(v * (float)h, 0.0f), h - 1u);
    return {x, y, sinf(th)};
}
__device__ __forceinline__ F3 env_radiance_t(const Scene &S, const Texel &t) {
    return tv3(S, S.env_rad + 3u * (t.y * S.env_w + t.x));
}
__device__ __forceinline__ float env_pdf_t(const Scene &S, const Texel &t) {
    if (t.st <= 1.0e-6f) return 0.0f;
    uint32_t w = S.env_w, h = S.env_h;
    return tf(S, S.env_mass + t.y * w + t.x) * (float)(w * h) / (2.0f * PI * PI * t.st);
}

struct EnvSample {
    bool ok;
    F3 d, le;
    float pdf;
};
__device__ EnvSample env_sample(const Scene &S, float u0, float u1) {
    EnvSample r;
    r.ok = false;
    r.d = f3(0.0f);
    r.le = f3(0.0f);
    r.pdf = 0.0f;
    uint32_t w = S.env_w, h = S.env_h;
    uint32_t y = env_row(S, u0);
    uint32_t row = S.env_cond + y * (w + 1u);
    uint32_t x = env_col(S, y, u1);
    float my = (u0 - env_marg_at(S, y)) / fmaxf(env_marg_at(S, y + 1u) - env_marg_at(S, y), 1.0e-12f);
    float mx = (u1 - tf(S, row + x)) / fmaxf(tf(S, row + x + 1u) - tf(S, row + x), 1.0e-12f);
    float u = ((float)x + clampf(mx, 0.0f, 0.99999f)) / (float)w;
    float v = ((float)y + clampf(my, 0.0f, 0.99999f)) / (float)h;
    float phi = 2.0f * PI * (u - 0.5f);
    float th = PI * v;
    float st = sinf(th), ct = cosf(th);
    F3 d = {st * sinf(phi), ct, -st * cosf(phi)};
    Texel t = env_texel(S, d);
    float pdf = env_pdf_t(S, t);
    if (pdf <= 0.0f || !finite(pdf)) return r;
    r.ok = true;
    r.d = d;
    r.le = env_radiance_t(S, t);
    r.pdf = pdf;
    return r;
}

__device__ __forceinline__ float power_heuristic(float a, float b) {
    float a2 = a * a, b2 = b * b;
    if (a2 + b2 > 0.0f && finite(a2 + b2)) return a2 / (a2 + b2);
    if (!finite(a) && a > 0.0f) return 1.0f;
    return 0.0f;
}
__device__ __forceinline__ F3 spawn(F3 p, F3 ng, F3 dir) {
    return dot(dir, ng) >= 0.0f ? offset_ray_origin(p, ng) : offset_ray_origin(p, -ng);
}

}  // namespace oxd
