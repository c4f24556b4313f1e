// The megakernel templates of the refine step (trace_persistent) and of
// the batched launch (trace_batch), with the path state machine they share
// (oxideav-render-vulkan src/shaders/pathtrace.wgsl:1714-2073 trace_sample +
// trace_main). Their forms (occupancy x mode x walk) are instantiated in
// trace_forms.hip, one object per (kernel, mode), so the builds of the
// forms compile in parallel; pathtrace.hip holds the other kernels and the
// launchers.
#pragma once

#include "pt_kernels.hpp"
#include "pt_shade.hpp"

namespace oxd {

__device__ __forceinline__ Scene make_scene(const oxpt_params &P, const float4 *geom, const float4 *attrs,
                                            const uint32_t *tables, const uint16_t *atlas, uint32_t atlas_side,
                                            const float *lds_sheen) {
    Scene S;
    S.geom = (gf4_ptr)geom;
    S.attrs = (gf4_ptr)attrs;
    S.tables = (gu32_ptr)tables;
    S.atlas = (gu16_ptr)atlas;
    S.atlas_side = atlas_side;
    S.tri_base = P.offs2[2];
    S.wide_base = P.offs2[3];
    S.tris = S.geom + S.tri_base;
    S.wide = S.geom + S.wide_base;
    S.node_count = P.env[3];
    S.off_mat = P.offs0[2];
    S.off_plights = P.offs0[3];
    S.off_elights = P.offs1[0];
    S.n_emissive = P.cfg[3];
    S.env_w = P.env[0];
    S.env_h = P.env[1];
    S.env_marg = P.offs1[2];
    S.env_cond = P.offs1[3];
    S.env_mass = P.offs2[0];
    S.env_rad = P.offs2[1];
    S.p_env = P.fparams[2];
    S.lds_sheen = lds_sheen;
    S.lds_marg = nullptr;
    S.lds_ecdf = nullptr;
    return S;
}
__device__ __forceinline__ Scene make_scene(const TraceArgs &A, const float *lds_sheen) {
    return make_scene(A.P, A.geom, A.attrs, A.tables, A.atlas, A.atlas_side, lds_sheen);
}

// LDS copies of the light-selection CDFs for single-scene kernels: the
// environment marginal (up to 1024 words) and the emissive CDF (up to 64
// records); larger tables are read from global memory. Every workgroup
// stages them once, after which the binary searches run on LDS.
constexpr uint32_t LDS_MARG_WORDS = 1024, LDS_ECDF_WORDS = 64;
__device__ __forceinline__ void stage_light_cdfs(const uint32_t *tables, const oxpt_params &P, float *lds_cdf,
                                                 Scene &S) {
    const bool marg = P.env[2] != 0 && P.env[1] + 1u <= LDS_MARG_WORDS;
    const bool ecdf = P.cfg[3] != 0 && P.cfg[3] <= LDS_ECDF_WORDS;
    if (marg)
        for (uint32_t i = threadIdx.x; i <= P.env[1]; i += blockDim.x) lds_cdf[i] = bitsf(tables[P.offs1[2] + i]);
    if (ecdf)
        for (uint32_t i = threadIdx.x; i < P.cfg[3]; i += blockDim.x)
            lds_cdf[LDS_MARG_WORDS + i] = bitsf(tables[P.offs1[0] + 4u * i]);
    __syncthreads();
    S.lds_marg = marg ? lds_cdf : nullptr;
    S.lds_ecdf = ecdf ? lds_cdf + LDS_MARG_WORDS : nullptr;
}

__device__ __forceinline__ F3 clamp_c(float cl, F3 c) {
    float m = max3(c);
    if (cl > 0.0f && m > cl) return c * (cl / m);
    return c;
}

// Per-lane path state of the one-ray-per-iteration state machine
// (wgsl:1733-1764).
struct PathState {
    uint32_t pixel_seed, path_seed, index;
    F3 d, beta, radiance, prev_pos, sigma;
    float prev_pdf;
    bool in_medium;
    uint32_t k;
    bool shadow;
    F3 r_o, r_d;
    float r_tmax;
    uint32_t r_id, r_skip;
    F3 pending;
    Bsdf bsdf;
    F3 p, ng;
    float thickness, transmission;
    F3 atten_color;
    float atten_dist;
    bool front;
    float4 pl_u;
    uint32_t job;
    // Early forms: the continuation ray's origin, sampled at the hit, and
    // whether a continuation follows the NEE shadow rays; with punctual
    // lights also the NEE jobs still to trace (bit j: job j) and the
    // vertex's bounce (the rays' ids).
    F3 c_o;
    bool cont;
    uint32_t mask, nee_k;
};

// Per-lane NEE queue of the early form with punctual lights, word-major
// across the launch's lanes: punctual job j's pending contribution at words
// 3j..3j+2, then the area / environment job's shadow ray (origin, direction,
// t_max, ray id, skip) and its pending contribution.
struct NeeQueue {
    float *q;
    uint32_t stride, lane;
    __device__ __forceinline__ float &operator[](uint32_t w) const { return q[(size_t)w * stride + lane]; }
};

// Camera ray of sample `index` of pixel (x, y) (wgsl:1715-1764).
__device__ __forceinline__ void init_path(const oxpt_params &P, const uint32_t *lds_sobol, uint32_t x, uint32_t y,
                                          uint32_t index, PathState &st) {
    st.pixel_seed = hash_mix(hash_mix(P.dims[2], y), x);
    st.index = index;
    st.path_seed = hash_mix(st.pixel_seed, index);
    float4 p0 = pattern(lds_sobol, st.pixel_seed, index, 0);
    float width = (float)P.dims[0], height = (float)P.dims[1];
    float ndc_x = (((float)x + p0.x) / fmaxf(width, 1.0f)) * 2.0f - 1.0f;
    float ndc_y = 1.0f - (((float)y + p0.y) / fmaxf(height, 1.0f)) * 2.0f;
    F3 eye = {P.eye[0], P.eye[1], P.eye[2]};
    F3 fwd = {P.forward[0], P.forward[1], P.forward[2]};
    F3 side = {P.side[0], P.side[1], P.side[2]};
    F3 up = {P.up[0], P.up[1], P.up[2]};
    F3 o, d;
    if (P.eye[3] > 0.5f) {
        o = eye + (side * (ndc_x * P.forward[3]) + up * (ndc_y * P.side[3]));
        d = fwd;
    } else {
        o = eye;
        d = nrm(side * (ndc_x * P.forward[3]) + up * (ndc_y * P.side[3]) + fwd);
    }
    st.d = d;
    st.beta = f3(1.0f);
    st.radiance = f3(0.0f);
    st.prev_pdf = 0.0f;
    st.prev_pos = f3(0.0f);
    st.in_medium = false;
    st.sigma = f3(0.0f);
    st.k = 0;
    st.shadow = false;
    st.r_o = o;
    st.r_d = d;
    st.r_tmax = INF;
    st.r_id = 0;
    st.r_skip = NONE;
    st.pending = f3(0.0f);
    st.p = f3(0.0f);
    st.ng = f3(0.0f);
    st.thickness = 0.0f;
    st.transmission = 0.0f;
    st.atten_color = f3(1.0f);
    st.atten_dist = -1.0f;
    st.front = true;
    st.pl_u = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    st.job = 0;
    st.c_o = f3(0.0f);
    st.cont = false;
    st.mask = 0;
    st.nee_k = 0;
}

// The shadow ray of NEE job `job` at a surface vertex: punctual light
// job < n_punct, then the emissive / environment set (wgsl:1863-1997).
// ok with its ray and the contribution it adds when unoccluded.
struct NeeRay {
    bool ok;
    F3 o, d, pending;
    float tmax;
    uint32_t id, skip;
};

__device__ __forceinline__ NeeRay nee_job(const oxpt_params &P, const Scene &S, const Bsdf &bsdf, F3 p, F3 ng,
                                          float4 pl_u, F3 beta, uint32_t k, uint32_t job) {
    const uint32_t strategy = P.cfg[1];
    const bool has_env = P.env[2] != 0;
    const uint32_t n_elights = P.cfg[3], n_punct = P.cfg[2];
    const float clampv = P.fparams[1];
    const uint32_t ray_id = k << 16u;
    NeeRay out;
    out.ok = false;
    bool ok = false;
    F3 l = f3(0.0f);
    float tmax = INF;
    uint32_t rid = ray_id | 0xffffu;
    uint32_t skip = NONE;
    F3 le = f3(0.0f);
    float pl = -1.0f;
    if (job < n_punct) {
        LightSample ls = punctual_sample(S, job, p);
        if (ls.ok) {
            ok = true;
            l = ls.l;
            if (ls.distance >= 0.0f) tmax = ls.distance * (1.0f - 1.0e-4f);
            rid = ray_id | 0x8000u | (job & 0x7fffu);
            le = ls.radiance;
        }
    } else if (strategy != 2 && (has_env || n_elights > 0)) {
        float pe = S.p_env;
        if (pl_u.z < pe) {
            if (has_env) {
                EnvSample es = env_sample(S, pl_u.x, pl_u.y);
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
            uint32_t li = pick_elight(S, us);
            uint32_t lslot = elight_slot(S, li);
            F3 a = tri_pos(S, lslot, 0), b = tri_pos(S, lslot, 1), c = tri_pos(S, lslot, 2);
            bool has_target = false;
            F3 q = f3(0.0f);
            const float omega = triangle_solid_angle(p, a, b, c);
            if (omega >= MIN_SPHERICAL_SOLID_ANGLE) {
                F3 sl = sample_spherical_triangle(p, a, b, c, pl_u.x, pl_u.y);
                if (dot(sl, sl) > 0.0f) {
                    F3 n = cross(b - a, c - a);
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
                float su = sqrtf(pl_u.x);
                F3 w3 = {1.0f - su, su * (1.0f - pl_u.y), su * pl_u.y};
                q = a * w3.x + b * w3.y + c * w3.z;
                has_target = true;
            }
            if (has_target) {
                F3 e1 = b - a, e2 = c - a, rr = q - a;
                float d11 = dot(e1, e1), d12 = dot(e1, e2), d22 = dot(e2, e2);
                float r1 = dot(rr, e1), r2 = dot(rr, e2);
                float det = d11 * d22 - d12 * d12;
                F3 to = q - p;
                float dist2 = dot(to, to);
                if (finite(det) && fabsf(det) > 1.17549435e-38f && dist2 > 1.0e-12f) {
                    float bu = (d22 * r1 - d12 * r2) / det;
                    float bv = (d11 * r2 - d12 * r1) / det;
                    F3 lb = vmin(vmax(F3{1.0f - bu - bv, bu, bv}, f3(0.0f)), f3(1.0f));
                    F3 lng = surface(S, lslot, lb).ng;
                    bool lds = (mat_flags(S, tri_material(S, lslot)) & 4u) != 0;
                    float dist = sqrtf(dist2);
                    F3 ld = to / dist;
                    float cos_l = -dot(lng, ld);
                    float tp = tri_pdf_omega(S, li, a, b, c, omega, p, q);
                    if ((cos_l > 0.0f || (lds && cos_l < 0.0f)) && tp > 0.0f && finite(tp)) {
                        ok = true;
                        l = ld;
                        le = emission(S, lslot, lb);
                        pl = tp;
                        tmax = fmaxf(dist * (1.0f - 1.0e-4f), 0.0f);
                        skip = tri_global(S, lslot);
                    }
                }
            }
        }
    }
    if (ok) {
        Eval e = bsdf_eval(S.lds_sheen, bsdf, l);
        if (!is_black(e.f) && !is_black(le)) {
            if (pl < 0.0f) {
                out.pending = clamp_c(clampv, beta * e.f * le);
            } else {
                float w = 1.0f;
                if (strategy == 0) w = power_heuristic(pl, e.pdf);
                out.pending = clamp_c(clampv, beta * e.f * le * (w / pl));
            }
            out.ok = true;
            out.o = spawn(p, ng, l);
            out.d = l;
            out.tmax = tmax;
            out.id = rid;
            out.skip = skip;
        }
    }
    return out;
}

// BSDF sampling at vertex k (wgsl:1999-2039): beta, the previous pdf and
// position, the medium on entering a thick transmissive surface, roulette
// from rr_start, then k + 1 and the new direction. False when the path ends;
// otherwise *origin is the continuation ray's origin.
__device__ __forceinline__ bool bsdf_continue(const oxpt_params &P, const uint32_t *lds_sobol, const float *lds_sheen,
                                              PathState &st, const Bsdf &bsdf, F3 p, F3 ng, float thickness,
                                              float transmission, F3 atten_color, float atten_dist, bool front,
                                              F3 *origin) {
    float4 pb_u = pattern(lds_sobol, st.pixel_seed, st.index, 1u + 2u * st.k);
    BSample bs = bsdf_sample(lds_sheen, bsdf, F3{pb_u.x, pb_u.y, pb_u.z});
    if (!bs.ok) return false;
    st.beta = st.beta * (bs.f * (1.0f / bs.pdf));
    if (!finite3(st.beta)) return false;
    st.prev_pdf = bs.pdf;
    st.prev_pos = p;
    if (dot(bs.l, ng) < 0.0f && thickness > 0.0f && transmission > 0.0f) {
        if (front && atten_dist >= 0.0f) {
            F3 lc = vmax(atten_color, f3(1.0e-6f));
            st.sigma = F3{-logf(lc.x), -logf(lc.y), -logf(lc.z)} / atten_dist;
            st.in_medium = true;
        } else {
            st.in_medium = false;
        }
    }
    if (st.k + 1u >= P.cfg[0]) {
        float q = clampf(max3(st.beta), 0.05f, 1.0f);
        if (q < 1.0f) {
            if (pb_u.w >= q) return false;
            st.beta = st.beta * (1.0f / q);
        }
    }
    st.k += 1;
    st.d = bs.l;
    *origin = spawn(p, ng, bs.l);
    return true;
}

// Early form with punctual lights: the next queued NEE shadow ray (job
// order: punctual lights, then the area / environment job), else the
// continuation. Returns true when the path ended.
__device__ __forceinline__ bool next_queued(const Scene &S, PathState &st, const NeeQueue &Q, uint32_t n_punct) {
    if (st.mask != 0) {
        const uint32_t j = (uint32_t)__ffs(st.mask) - 1u;
        st.mask &= st.mask - 1u;
        if (j < n_punct) {
            // The light's ray as nee_job made it (punctual_sample is exact).
            const LightSample ls = punctual_sample(S, j, st.p);
            st.r_o = spawn(st.p, st.ng, ls.l);
            st.r_d = ls.l;
            st.r_tmax = ls.distance >= 0.0f ? ls.distance * (1.0f - 1.0e-4f) : INF;
            st.r_id = (st.nee_k << 16u) | 0x8000u | (j & 0x7fffu);
            st.r_skip = NONE;
            st.pending = F3{Q[3u * j], Q[3u * j + 1u], Q[3u * j + 2u]};
        } else {
            const uint32_t a = 3u * n_punct;
            st.r_o = F3{Q[a], Q[a + 1u], Q[a + 2u]};
            st.r_d = F3{Q[a + 3u], Q[a + 4u], Q[a + 5u]};
            st.r_tmax = Q[a + 6u];
            st.r_id = fbits(Q[a + 7u]);
            st.r_skip = fbits(Q[a + 8u]);
            st.pending = F3{Q[a + 9u], Q[a + 10u], Q[a + 11u]};
        }
        st.shadow = true;
        return false;
    }
    if (!st.cont) return true;
    st.shadow = false;
    st.r_o = st.c_o;
    st.r_d = st.d;
    st.r_tmax = INF;
    st.r_id = st.k << 16u;
    st.r_skip = NONE;
    return false;
}

// Advances the state machine past the ray just traced, given its hit
// (wgsl:1766-2040). Returns true when the path ended; *covered says
// whether the camera ray hit anything. Otherwise st.r_* holds the next ray.
//
// Mode 1 is the form for scenes without punctual lights (one NEE job per
// vertex): the NEE shadow ray and the BSDF continuation are both made at
// the hit, so the BSDF and the vertex need not live across a traversal; the
// shadow ray is traced first, then the continuation. Mode 2 does the same
// with punctual lights: every NEE job is evaluated at the hit, its pending
// contribution (and the area job's whole ray) goes to the lane's NEE queue,
// and each punctual shadow ray is rebuilt from its light when traced. Every
// random number and every sum is the same, in the same order, so the images
// are bit-identical to the general form (mode 0).
template <int Mode>
__device__ bool advance(const oxpt_params &P, const Scene &S, const uint32_t *lds_sobol, PathState &st,
                        const Hit &hit, bool *covered, const NeeQueue &Q) {
    const uint32_t strategy = P.cfg[1];
    const bool has_env = P.env[2] != 0;
    const uint32_t n_punct = P.cfg[2];
    const float clampv = P.fparams[1];
    *covered = true;
    if (st.shadow) {
        if (hit.slot == NONE) st.radiance = st.radiance + st.pending;
        if constexpr (Mode == 2) {
            return next_queued(S, st, Q, n_punct);
        } else if constexpr (Mode == 1) {
            // The continuation was sampled at the hit; trace it now.
            if (!st.cont) return true;
            st.shadow = false;
            st.r_o = st.c_o;
            st.r_d = st.d;
            st.r_tmax = INF;
            st.r_id = st.k << 16u;
            st.r_skip = NONE;
            return false;
        }
        st.job += 1;
    } else {
        if (st.in_medium) {
            if (hit.slot == NONE) return true;
            F3 e = -st.sigma * hit.t;
            st.beta = st.beta * F3{expf(e.x), expf(e.y), expf(e.z)};
        }
        if (hit.slot == NONE) {
            if (st.k == 0) {
                *covered = false;
                return true;
            }
            F3 le = f3(P.fparams[0]);
            float w = 1.0f;
            if (has_env) {
                Texel t = env_texel(S, st.d);
                le = env_radiance_t(S, t);
                if (strategy == 0) w = power_heuristic(st.prev_pdf, S.p_env * env_pdf_t(S, t));
                else if (strategy == 1) w = 0.0f;
            }
            if (w > 0.0f) st.radiance = st.radiance + clamp_c(clampv, st.beta * le * w);
            return true;
        }
        Surface surf = surface(S, hit.slot, hit.b);
        uint32_t mflags = mat_flags(S, tri_material(S, hit.slot));
        bool double_sided = (mflags & 4u) != 0;
        Lod lod = base_lod();
        if (st.k == 0) {
            float cw = P.up[3];
            if (!(P.eye[3] > 0.5f)) cw = P.up[3] * hit.t;
            lod = {cw, st.d};
        }
        bool flip = !hit.front && double_sided;
        Mat mat = material(S, hit.slot, hit.b, surf.sn, lod, flip);
        if (flip) {
            mat.normal = -mat.normal;
            mat.cc_normal = -mat.cc_normal;
        }
        if ((hit.front || double_sided) && !is_black(mat.emission)) {
            F3 le = mat.emission;
            if (st.k == 0) {
                st.radiance = st.radiance + st.beta * le;
            } else {
                uint32_t li = tri_light(S, hit.slot);
                float pl = li == NONE ? 0.0f : tri_pdf(S, li, hit.slot, st.prev_pos, surf.position);
                float w = 1.0f;
                if (pl > 0.0f) {
                    if (strategy == 0) w = power_heuristic(st.prev_pdf, pl);
                    else if (strategy == 1) w = 0.0f;
                }
                if (w > 0.0f) st.radiance = st.radiance + clamp_c(clampv, st.beta * le * w);
            }
        }
        if (mat.unlit) {
            F3 contrib = st.beta * xyz(mat.base);
            st.radiance = st.radiance + (st.k == 0 ? contrib : clamp_c(clampv, contrib));
            return true;
        }
        if (st.k >= P.dims[3]) return true;
        if constexpr (Mode == 2) {
            const Bsdf bsdf = make_bsdf(S.lds_sheen, mat, -st.d, surf.ng, hit.front);
            if (bsdf_black(bsdf)) return true;
            const F3 p = surf.position, ng = bsdf.ng;
            const float4 pl_u = pattern(lds_sobol, st.pixel_seed, st.index, 2u + 2u * st.k);
            uint32_t mask = 0;
            for (uint32_t j = 0; j <= n_punct; j++) {
                const NeeRay nr = nee_job(P, S, bsdf, p, ng, pl_u, st.beta, st.k, j);
                if (!nr.ok) continue;
                mask |= 1u << j;
                if (j < n_punct) {
                    Q[3u * j] = nr.pending.x;
                    Q[3u * j + 1u] = nr.pending.y;
                    Q[3u * j + 2u] = nr.pending.z;
                } else {
                    const uint32_t a = 3u * n_punct;
                    Q[a] = nr.o.x;
                    Q[a + 1u] = nr.o.y;
                    Q[a + 2u] = nr.o.z;
                    Q[a + 3u] = nr.d.x;
                    Q[a + 4u] = nr.d.y;
                    Q[a + 5u] = nr.d.z;
                    Q[a + 6u] = nr.tmax;
                    Q[a + 7u] = bitsf(nr.id);
                    Q[a + 8u] = bitsf(nr.skip);
                    Q[a + 9u] = nr.pending.x;
                    Q[a + 10u] = nr.pending.y;
                    Q[a + 11u] = nr.pending.z;
                }
            }
            st.nee_k = st.k;
            F3 co = f3(0.0f);
            st.cont = bsdf_continue(P, lds_sobol, S.lds_sheen, st, bsdf, p, ng, mat.thickness, mat.transmission,
                                    mat.atten_color, mat.atten_dist, hit.front, &co);
            st.c_o = co;
            st.p = p;
            st.ng = ng;
            st.mask = mask;
            return next_queued(S, st, Q, n_punct);
        }
        if constexpr (Mode == 1) {
            const Bsdf bsdf = make_bsdf(S.lds_sheen, mat, -st.d, surf.ng, hit.front);
            if (bsdf_black(bsdf)) return true;
            const F3 p = surf.position, ng = bsdf.ng;
            const NeeRay nr = nee_job(P, S, bsdf, p, ng, pattern(lds_sobol, st.pixel_seed, st.index, 2u + 2u * st.k),
                                      st.beta, st.k, n_punct);
            F3 co = f3(0.0f);
            const bool cont = bsdf_continue(P, lds_sobol, S.lds_sheen, st, bsdf, p, ng, mat.thickness, mat.transmission,
                                            mat.atten_color, mat.atten_dist, hit.front, &co);
            if (nr.ok) {
                st.pending = nr.pending;
                st.r_o = nr.o;
                st.r_d = nr.d;
                st.r_tmax = nr.tmax;
                st.r_id = nr.id;
                st.r_skip = nr.skip;
                st.shadow = true;
                st.cont = cont;
                st.c_o = co;
                return false;
            }
            if (!cont) return true;
            st.shadow = false;
            st.r_o = co;
            st.r_d = st.d;
            st.r_tmax = INF;
            st.r_id = st.k << 16u;
            st.r_skip = NONE;
            return false;
        }
        st.bsdf = make_bsdf(S.lds_sheen, mat, -st.d, surf.ng, hit.front);
        if (bsdf_black(st.bsdf)) return true;
        st.p = surf.position;
        st.ng = st.bsdf.ng;
        st.thickness = mat.thickness;
        st.transmission = mat.transmission;
        st.atten_color = mat.atten_color;
        st.atten_dist = mat.atten_dist;
        st.front = hit.front;
        st.pl_u = pattern(lds_sobol, st.pixel_seed, st.index, 2u + 2u * st.k);
        st.job = 0;
    }

    // ---- Next NEE shadow ray: punctual j < n, then the area set (wgsl:1863-1997).
    while (st.job <= n_punct) {
        const NeeRay nr = nee_job(P, S, st.bsdf, st.p, st.ng, st.pl_u, st.beta, st.k, st.job);
        if (nr.ok) {
            st.pending = nr.pending;
            st.r_o = nr.o;
            st.r_d = nr.d;
            st.r_tmax = nr.tmax;
            st.r_id = nr.id;
            st.r_skip = nr.skip;
            st.shadow = true;
            return false;
        }
        st.job += 1;
    }

    // ---- BSDF sampling -> next path ray (wgsl:1999-2039).
    F3 co = f3(0.0f);
    if (!bsdf_continue(P, lds_sobol, S.lds_sheen, st, st.bsdf, st.p, st.ng, st.thickness, st.transmission,
                       st.atten_color, st.atten_dist, st.front, &co))
        return true;
    st.shadow = false;
    st.r_o = co;
    st.r_d = st.d;
    st.r_tmax = INF;
    st.r_id = st.k << 16u;
    st.r_skip = NONE;
    return false;
}

// One iteration: trace the current ray and advance the state machine.
// `lds_stack` / `lds_depth`: this wave's LDS share of the traversal stack
// (depth 0: the whole stack in private memory). Opaque: the walk built for
// scenes whose materials are all alpha mode OPAQUE (accept_hit).
template <int Mode = 0, bool Opaque = false, bool Defer = false>
__device__ __forceinline__ bool step_path(const oxpt_params &P, const Scene &S, const uint32_t *lds_sobol,
                                          PathState &st, bool *covered, lds_u64 *lds_stack = nullptr,
                                          uint32_t lds_depth = 0, const NeeQueue &Q = NeeQueue{nullptr, 0, 0}) {
    // The NEE-queue form (punctual lights: most rays are shadow rays) walks
    // any-hit queries without ordering the children.
    Hit hit = traverse<-1, Mode == 2, Opaque, Defer>(S, make_ray(st.r_o, st.r_d), st.r_tmax, st.path_seed, st.r_id,
                                      !st.shadow && st.k == 0, st.shadow, st.r_skip, lds_stack, lds_depth);
    return advance<Mode>(P, S, lds_sobol, st, hit, covered, Q);
}

__device__ __forceinline__ void stage_tables(const uint32_t *tables, const oxpt_params &P, uint32_t *lds_sobol,
                                             float *lds_sheen) {
    for (uint32_t i = threadIdx.x; i < OXPT_SOBOL_WORDS; i += blockDim.x) lds_sobol[i] = tables[P.offs0[0] + i];
    for (uint32_t i = threadIdx.x; i < OXPT_SHEEN_WORDS; i += blockDim.x) lds_sheen[i] = bitsf(tables[P.offs0[1] + i]);
    __syncthreads();
}

// Work item -> pixel: items run through 8x8 tiles, tiles row-major.
__device__ __forceinline__ void item_pixel(uint32_t tiles_x, uint32_t item, uint32_t &x, uint32_t &y) {
    uint32_t tile = item >> 6, in = item & 63u;
    uint32_t tx = tile % tiles_x, ty = tile / tiles_x;
    x = tx * 8u + (in & 7u);
    y = ty * 8u + (in >> 3);
}

// A copy of a launch-constant record (kernel arguments, a batched job)
// read through the constant address space: the loads are invariant scalar
// loads, which the compiler may repeat where a value is needed instead of
// keeping it live across the walk (the 1080p frame: 962 -> 970 Mpaths/s;
// the job table read through a plain global pointer cost 2 %). The record
// must not change while the kernel runs.
template <class T>
__device__ __forceinline__ T ld_const(const T *p) {
    static_assert(sizeof(T) % 4 == 0, "records are whole words");
    T out;
    uint32_t *d = reinterpret_cast<uint32_t *>(&out);
    const __attribute__((address_space(4))) uint32_t *q = (const __attribute__((address_space(4))) uint32_t *)p;
#pragma unroll
    for (uint32_t i = 0; i < sizeof(T) / 4; i++) d[i] = q[i];
    return out;
}

// MinWaves > 0 asks the compiler for that many waves per SIMD (fewer
// VGPRs, more latency hiding, possible spills): an autotuner knob.
// lds_depth: traversal-stack entries per lane kept in dynamic LDS
// (blockDim / 64 * lds_depth * 512 bytes), the rest in private memory.
template <int MinWaves, int Mode = 0, bool Opaque = false>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(MinWaves > 0 ? MinWaves : 1)))
trace_persistent(TraceArgs args, uint32_t lds_depth) {
    const TraceArgs A = ld_const(&args);
    __shared__ uint32_t lds_sobol[OXPT_SOBOL_WORDS];
    __shared__ float lds_sheen[OXPT_SHEEN_WORDS];
    extern __shared__ uint64_t lds_dyn[];
    lds_u64 *lds_stack = to_lds(lds_dyn) + (threadIdx.x >> 6) * lds_depth * 64u;
    __shared__ float lds_cdf[LDS_MARG_WORDS + LDS_ECDF_WORDS];
    stage_tables(A.tables, A.P, lds_sobol, lds_sheen);
    Scene S = make_scene(A, lds_sheen);
    stage_light_cdfs(A.tables, A.P, lds_cdf, S);
    const uint32_t lane = __lane_id();
    const uint64_t lt_mask = lane == 0 ? 0ull : (~0ull >> (64u - lane));
    const uint32_t w = A.P.dims[0], h = A.P.dims[1];
    // A lane holds samples [done, run_end) of pixel (px, py): a pixel's
    // whole run added into `acc`, or (to_buffer) one tail path whose sample
    // goes to the sample buffer at slot_base + 64 * sample.
    bool has_pixel = false, exhausted = false, need_init = false, to_buffer = false;
    uint32_t px = 0, py = 0, done = 0, run_end = 0, slot_base = 0;
    // Wave-uniform: the pixel queue is spent, claims go to the tail paths.
    bool pixels_spent = A.split == 0;
    float4 acc = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    PathState st;
    const NeeQueue Q{A.nee_q, gridDim.x * blockDim.x, blockIdx.x * blockDim.x + threadIdx.x};
    for (;;) {
        // Regenerate: idle lanes claim work with one atomic per wave — whole
        // pixels while items before the split remain, then single tail
        // paths (a wave's lanes take one tile at one sample when they claim
        // together).
        bool want = !has_pixel && !exhausted;
        uint64_t m = __ballot(want);
        if (m) {
            const uint32_t leader = __ffsll((unsigned long long)m) - 1u;
            const uint32_t n = (uint32_t)__popcll(m);
            uint32_t base = 0;
            if (lane == leader) base = atomicAdd(A.work + (pixels_spent ? 1u : 0u), n);
            base = __builtin_amdgcn_readfirstlane(__shfl(base, leader));
            const uint32_t rank = (uint32_t)__popcll(m & lt_mask);
            if (!pixels_spent) {
                if (want) {
                    const uint32_t item = base + rank;
                    // Items outside the image (ragged edge tiles) have no
                    // pixel: the lane claims again on the next ballot.
                    if (item < A.split) {
                        item_pixel(A.tiles_x, item, px, py);
                        if (px < w && py < h) {
                            has_pixel = true;
                            to_buffer = false;
                            done = 0;
                            run_end = A.count;
                            need_init = true;
                            acc = A.accum[(size_t)py * w + px];
                        }
                    }
                }
                if (base + n >= A.split) pixels_spent = true;
            } else if (want) {
                const uint32_t q = base + rank;
                if (q >= A.tail_work) {
                    exhausted = true;
                } else {
                    const uint32_t tile = q / (64u * A.count), rem = q - tile * 64u * A.count;
                    const uint32_t in = rem & 63u;
                    item_pixel(A.tiles_x, A.split + tile * 64u + in, px, py);
                    if (px < w && py < h) {
                        has_pixel = true;
                        to_buffer = true;
                        done = rem >> 6;
                        run_end = done + 1u;
                        slot_base = tile * 64u * A.count + in;
                        need_init = true;
                    }
                }
            }
        }
        if (__ballot(has_pixel || (!exhausted)) == 0) break;
        if (!has_pixel) continue;
        if (need_init) {
            init_path(A.P, lds_sobol, px, py, A.first + done, st);
            need_init = false;
        }
        // A hit's shadow rays are traced before the wave returns to the claim
        // step, so the lanes of a wave stay in the same phase of their
        // bounces (less divergent shading): with punctual lights (many
        // shadow rays per hit) and in the early form without them (lanes
        // whose hit made no shadow ray would otherwise shade a step ahead;
        // the 1080p frame 1042 -> 1055 Mpaths/s).
        bool covered, ended;
        while (!(ended = step_path<Mode, Opaque>(A.P, S, lds_sobol, st, &covered, lds_stack, lds_depth, Q)) && st.shadow &&
               (Mode == 1 || A.P.cfg[2] != 0)) {
        }
        if (ended) {
            if (to_buffer) {
                F3 r = covered && finite3(st.radiance) ? st.radiance : f3(0.0f);
                A.samples[slot_base + 64u * done] = make_float4(r.x, r.y, r.z, covered ? 1.0f : 0.0f);
            } else if (covered) {
                F3 r = finite3(st.radiance) ? st.radiance : f3(0.0f);
                acc = make_float4(acc.x + r.x, acc.y + r.y, acc.z + r.z, acc.w + 1.0f);
            }
            done += 1;
            if (done == run_end) {
                if (!to_buffer) A.accum[(size_t)py * w + px] = acc;
                has_pixel = false;
            } else {
                need_init = true;
            }
        }
    }
}

// Batched form: many (scene, camera, sample run) jobs in one persistent
// launch, each job worked as trace_persistent works a frame. A wave works
// on one job at a time (so the job's Params and scene pointers stay in
// scalar registers). Its lanes claim whole pixels of the job first — items
// [0, split) — and trace each pixel's samples in order into the pixel's
// accumulator; once the job's pixels are spent, lanes claim single paths of
// the job's tail items [split, n_items) — (tile, sample, pixel), so lanes
// that claim together take one tile at one sample — whose (rgb, covered)
// go to the launch's sample buffer, and batch_reduce adds them into the
// accumulator in sample order. So lanes out of pixels take tail paths
// instead of idling while the rest of their wave finishes whole runs, the
// wave leaves a job after at most one tail path per lane, and every job's
// image equals its image alone. The host orders jobs largest first and
// gives the last ones (whose pixel runs would outlast the launch) no pixel
// runs at all.
template <int MinWaves, int Mode = 0, bool Opaque = false>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(MinWaves)))
trace_batch(BatchArgs B, const Job *__restrict__ jobs, uint32_t lds_depth) {
    __shared__ uint32_t lds_sobol[OXPT_SOBOL_WORDS];
    __shared__ float lds_sheen[OXPT_SHEEN_WORDS];
    extern __shared__ uint64_t lds_dyn[];
    lds_u64 *lds_stack = to_lds(lds_dyn) + (threadIdx.x >> 6) * lds_depth * 64u;
    // Every job carries the same estimator constants (checked on the host).
    const uint32_t n_jobs = min(*B.n_jobs, B.max_jobs);
    if (n_jobs == 0) return;
    stage_tables(jobs[0].tables, jobs[0].P, lds_sobol, lds_sheen);
    const uint32_t lane = __lane_id();
    const uint64_t lt_mask = lane == 0 ? 0ull : (~0ull >> (64u - lane));
    PathState st;
    const NeeQueue Q{B.nee_q, gridDim.x * blockDim.x, blockIdx.x * blockDim.x + threadIdx.x};
    uint32_t j = 0;
    for (;;) {
        j = __builtin_amdgcn_readfirstlane(j);
        const Job J = ld_const(jobs + j);
        const Scene S = make_scene(J.P, J.geom, J.attrs, J.tables, J.atlas, J.atlas_side, lds_sheen);
        const uint32_t w = J.P.dims[0], h = J.P.dims[1], tiles_x = J.tiles_x();
        uint32_t *const work = B.counters + 2u * j;  // [0] pixels, [1] tail paths
        // A lane holds samples [done, run_end) of pixel (px, py): a pixel's
        // whole run added into `acc`, or (to_buffer) one tail path whose
        // sample goes to the sample buffer at slot + 64 * sample.
        bool has_pixel = false, exhausted = false, need_init = false, to_buffer = false;
        uint32_t px = 0, py = 0, done = 0, run_end = 0, slot = 0;
        // Wave-uniform: the job's pixel queue is spent, claims go to its tail.
        bool pixels_spent = J.split == 0;
        float4 acc = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        for (;;) {
            const bool want = !has_pixel && !exhausted;
            const uint64_t m = __ballot(want);
            if (m) {
                const uint32_t leader = __ffsll((unsigned long long)m) - 1u;
                const uint32_t n = (uint32_t)__popcll(m);
                uint32_t base = 0;
                if (lane == leader) base = atomicAdd(work + (pixels_spent ? 1u : 0u), n);
                base = __builtin_amdgcn_readfirstlane(__shfl(base, leader));
                const uint32_t rank = (uint32_t)__popcll(m & lt_mask);
                if (!pixels_spent) {
                    if (want) {
                        const uint32_t item = base + rank;
                        // Items outside the image (ragged edge tiles) have no
                        // pixel: the lane claims again on the next ballot.
                        if (item < J.split) {
                            const uint32_t tile = item >> 6, in = item & 63u;
                            const uint32_t ty = udiv(tile, J.tx_m, J.tx_l()), tx = tile - ty * tiles_x;
                            px = tx * 8u + (in & 7u);
                            py = ty * 8u + (in >> 3);
                            if (px < w && py < h) {
                                has_pixel = true;
                                to_buffer = false;
                                done = 0;
                                run_end = J.count;
                                need_init = true;
                                acc = J.accum[(size_t)py * w + px];
                            }
                        }
                    }
                    if (base + n >= J.split) pixels_spent = true;
                } else if (want) {
                    const uint32_t q = base + rank;
                    if (q >= J.tail_work) {
                        exhausted = true;
                    } else {
                        const uint32_t per_tile = 64u * J.count;
                        const uint32_t tt = udiv(q, J.tile_m, J.tile_l()), rem = q - tt * per_tile;
                        const uint32_t in = rem & 63u, tile = (J.split >> 6) + tt;
                        const uint32_t ty = udiv(tile, J.tx_m, J.tx_l()), tx = tile - ty * tiles_x;
                        px = tx * 8u + (in & 7u);
                        py = ty * 8u + (in >> 3);
                        if (px < w && py < h) {
                            has_pixel = true;
                            to_buffer = true;
                            done = rem >> 6;
                            run_end = done + 1u;
                            slot = J.sample_base + tt * per_tile + in;
                            need_init = true;
                        }
                    }
                }
            }
            if (__ballot(has_pixel || !exhausted) == 0) break;  // job j is drained for this wave
            if (!has_pixel) continue;
            if (need_init) {
                init_path(J.P, lds_sobol, px, py, J.first + done, st);
                need_init = false;
            }
            // A hit's shadow rays are traced before the wave returns to the
            // claim step, so its lanes shade their next hits together: with
            // punctual lights (many shadow rays per hit) and in the early
            // form (lanes whose hit made no shadow ray would otherwise shade
            // a step ahead of the rest; on the uneven batch of 64 terrain
            // scenes 874 -> 952 Mpaths/s, on a 256x256 room at 16 spp 688 ->
            // 905).
            bool covered, ended;
            while (!(ended = step_path<Mode, Opaque, true>(J.P, S, lds_sobol, st, &covered, lds_stack, lds_depth, Q)) &&
                   st.shadow && (Mode == 1 || J.P.cfg[2] != 0)) {
            }
            if (ended) {
                if (to_buffer) {
                    const F3 r = covered && finite3(st.radiance) ? st.radiance : f3(0.0f);
                    B.samples[slot + 64u * done] = make_float4(r.x, r.y, r.z, covered ? 1.0f : 0.0f);
                } else if (covered) {
                    const F3 r = finite3(st.radiance) ? st.radiance : f3(0.0f);
                    acc = make_float4(acc.x + r.x, acc.y + r.y, acc.z + r.z, acc.w + 1.0f);
                }
                done += 1;
                if (done == run_end) {
                    if (!to_buffer) J.accum[(size_t)py * w + px] = acc;
                    has_pixel = false;
                } else {
                    need_init = true;
                }
            }
        }
        // Next job with unclaimed work.
        uint32_t nj = j + 1;
        while (nj < n_jobs &&
               __hip_atomic_load(B.counters + 2u * nj, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= jobs[nj].split &&
               __hip_atomic_load(B.counters + 2u * nj + 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >=
                   jobs[nj].tail_work)
            nj++;
        if (nj >= n_jobs) break;
        j = nj;
    }
}

}  // namespace oxd

namespace oxd {

// Launchers of the megakernel forms, one per (kernel, mode), each defined
// in its own build of trace_forms.hip. `opaque` selects the opaque walk
// (modes 1 and 2; the general mode has only the general walk).
void launch_persistent_m0(const TraceArgs &a, uint32_t kernel, uint32_t block, uint32_t grid, uint32_t lds_depth,
                          size_t shm, bool opaque, hipStream_t s);
void launch_persistent_m1(const TraceArgs &a, uint32_t kernel, uint32_t block, uint32_t grid, uint32_t lds_depth,
                          size_t shm, bool opaque, hipStream_t s);
void launch_persistent_m2(const TraceArgs &a, uint32_t kernel, uint32_t block, uint32_t grid, uint32_t lds_depth,
                          size_t shm, bool opaque, hipStream_t s);
void launch_batch_m0(const BatchArgs &a, uint32_t kernel, uint32_t block, uint32_t grid, uint32_t lds_depth,
                     size_t shm, bool opaque, hipStream_t s);
void launch_batch_m1(const BatchArgs &a, uint32_t kernel, uint32_t block, uint32_t grid, uint32_t lds_depth,
                     size_t shm, bool opaque, hipStream_t s);
void launch_batch_m2(const BatchArgs &a, uint32_t kernel, uint32_t block, uint32_t grid, uint32_t lds_depth,
                     size_t shm, bool opaque, hipStream_t s);

}  // namespace oxd
