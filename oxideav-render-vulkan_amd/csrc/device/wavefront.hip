// Wavefront path tracing for gfx950: the estimator of the megakernel
// (pathtrace.hip step_path, i.e. oxideav-render-vulkan
// src/shaders/pathtrace.wgsl:1714-2045) split into separate kernels so the
// BVH walk runs without the path state in registers:
//
//   wf_shade   — per path slot: add the previous vertex's unoccluded NEE
//                contributions in light order, shade the path ray's hit
//                (emission, material, layered BSDF), emit every NEE shadow
//                ray of the vertex, sample the BSDF for the next path ray;
//                a finished path writes its sample and the slot claims the
//                next (job, pixel, sample) item (ballot + one atomic per wave)
//   wf_extend  — closest hit of each slot's path ray
//   wf_occlude — any hit of each shadow ray
//   wf_reduce  — per pixel, add the pass's samples into the accumulator in
//                sample order
//
// Every contribution is added to a path's radiance in the same order as the
// megakernel, and every pixel's samples are added to its accumulator in
// sample order, so the accumulators are bit-identical to the megakernel's.
#include "pt_kernels.hpp"
#include "pt_shade.hpp"
#include "wavefront.hpp"

namespace oxd {

namespace {

constexpr uint32_t F_PATH_RAY = 1u;   // a path ray was emitted; its hit is in hits[slot]
constexpr uint32_t F_ENDED = 2u;      // path finished; waiting for its last shadow rays
constexpr uint32_t F_IN_MEDIUM = 4u;
constexpr uint32_t F_COVERED = 8u;    // cleared when the camera ray misses

__device__ __forceinline__ F3 clamp_c(float cl, F3 c) {
    float m = max3(c);
    if (cl > 0.0f && m > cl) return c * (cl / m);
    return c;
}

__device__ __forceinline__ Scene job_scene(const Job &J, const float *lds_sheen) {
    Scene S;
    S.geom = (gf4_ptr)J.geom;
    S.attrs = (gf4_ptr)J.attrs;
    S.tables = (gu32_ptr)J.tables;
    S.atlas = (gu16_ptr)J.atlas;
    S.atlas_side = J.atlas_side;
    S.tri_base = J.P.offs2[2];
    S.wide_base = J.P.offs2[3];
    S.tris = S.geom + S.tri_base;
    S.wide = S.geom + S.wide_base;
    S.node_count = J.P.env[3];
    S.off_mat = J.P.offs0[2];
    S.off_plights = J.P.offs0[3];
    S.off_elights = J.P.offs1[0];
    S.n_emissive = J.P.cfg[3];
    S.env_w = J.P.env[0];
    S.env_h = J.P.env[1];
    S.env_marg = J.P.offs1[2];
    S.env_cond = J.P.offs1[3];
    S.env_mass = J.P.offs2[0];
    S.env_rad = J.P.offs2[1];
    S.p_env = J.P.fparams[2];
    S.lds_sheen = lds_sheen;
    S.lds_marg = nullptr;
    S.lds_ecdf = nullptr;
    return S;
}

// Item -> (job, pixel, sample) of the current pass. Items run job-major,
// then 8x8 tile, then sample, then pixel within the tile, so the 64 lanes of
// a fresh wave take one tile at one sample index.
__device__ __forceinline__ bool decode_item(const WfArgs &A, uint32_t item, uint32_t &job, uint32_t &px, uint32_t &py,
                                            uint32_t &s) {
    uint32_t lo = 0, hi = A.n_jobs;  // last job with base <= item
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) / 2u;
        if (A.job_base[mid] <= item) lo = mid;
        else hi = mid;
    }
    job = lo;
    const Job &J = A.jobs[job];
    uint32_t local = item - A.job_base[job];
    uint32_t per_tile = 64u * J.count;
    uint32_t tile = local / per_tile, rem = local - tile * per_tile;
    s = rem >> 6;
    uint32_t in = rem & 63u;
    px = (tile % J.tiles_x()) * 8u + (in & 7u);
    py = (tile / J.tiles_x()) * 8u + (in >> 3);
    return px < J.P.dims[0] && py < J.P.dims[1];
}

}  // namespace

template <int Waves>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(Waves))) wf_shade(WfArgs A) {
    __shared__ uint32_t lds_sobol[OXPT_SOBOL_WORDS];
    __shared__ float lds_sheen[OXPT_SHEEN_WORDS];
    for (uint32_t i = threadIdx.x; i < OXPT_SOBOL_WORDS; i += blockDim.x)
        lds_sobol[i] = A.jobs[0].tables[A.jobs[0].P.offs0[0] + i];
    for (uint32_t i = threadIdx.x; i < OXPT_SHEEN_WORDS; i += blockDim.x)
        lds_sheen[i] = bitsf(A.jobs[0].tables[A.jobs[0].P.offs0[1] + i]);
    __syncthreads();
    if (blockIdx.x == 0 && threadIdx.x == 0) A.live[(A.iter + 1u) & (WF_LIVE_RING - 1u)] = 0;
    const uint32_t slot = blockIdx.x * blockDim.x + threadIdx.x;
    const bool in_range = slot < A.n_slots;
    PathSlot st;
    if (in_range) st = A.slots[slot];
    else st.item = NONE;
    bool busy = in_range && st.item != NONE;
    if (busy) {
        const Job &J = A.jobs[st.job];
        const oxpt_params &P = J.P;
        const Scene S = job_scene(J, lds_sheen);
        F3 radiance = {st.rad_x, st.rad_y, st.rad_z};
        // 1. The previous vertex's shadow rays, in light order.
        uint32_t nsh = st.n_shadow;
        for (uint32_t j = 0; j < nsh; j++) {
            if (!A.occluded[(size_t)j * A.n_slots + slot]) {
                float4 pr = A.shadows[(size_t)j * A.n_slots + slot].pending_rid;
                radiance = radiance + F3{pr.x, pr.y, pr.z};
            }
        }
        st.n_shadow = 0;
        bool ended = (st.flags & F_ENDED) != 0;
        if (!ended && (st.flags & F_PATH_RAY)) {
            // 2. The path ray's hit (wgsl:1773-1861).
            st.flags &= ~F_PATH_RAY;
            const uint32_t strategy = P.cfg[1];
            const bool has_env = P.env[2] != 0;
            const uint32_t n_elights = P.cfg[3], n_punct = P.cfg[2];
            const float clampv = P.fparams[1];
            HitRec hr = A.hits[slot];
            uint32_t hslot = hr.slot & 0x7fffffffu;
            bool hit_front = (hr.slot >> 31) != 0;
            bool miss = hr.slot == NONE;
            F3 beta = {st.beta_x, st.beta_y, st.beta_z};
            F3 d = {st.d_x, st.d_y, st.d_z};
            if (st.flags & F_IN_MEDIUM) {
                if (miss) {
                    ended = true;
                } else {
                    F3 sg = {st.sigma_x, st.sigma_y, st.sigma_z};
                    F3 e = -sg * hr.t;
                    beta = beta * F3{expf(e.x), expf(e.y), expf(e.z)};
                }
            }
            if (!ended && miss) {
                if (st.k == 0) {
                    st.flags &= ~F_COVERED;
                } else {
                    F3 le = f3(P.fparams[0]);
                    float w = 1.0f;
                    if (has_env) {
                        Texel t = env_texel(S, d);
                        le = env_radiance_t(S, t);
                        if (strategy == 0) w = power_heuristic(st.prev_pdf, S.p_env * env_pdf_t(S, t));
                        else if (strategy == 1) w = 0.0f;
                    }
                    if (w > 0.0f) radiance = radiance + clamp_c(clampv, beta * le * w);
                }
                ended = true;
            }
            if (!ended) {
                F3 b = {1.0f - hr.b1 - hr.b2, hr.b1, hr.b2};
                // Barycentrics are stored as (u, v); the first weight is
                // recomputed exactly as intersect_tri forms it.
                Surface surf = surface(S, hslot, b);
                uint32_t mflags = mat_flags(S, tri_material(S, hslot));
                bool double_sided = (mflags & 4u) != 0;
                Lod lod = base_lod();
                if (st.k == 0) {
                    float cw = P.up[3];
                    if (!(P.eye[3] > 0.5f)) cw = P.up[3] * hr.t;
                    lod = {cw, d};
                }
                bool flip = !hit_front && double_sided;
                Mat mat = material(S, hslot, b, surf.sn, lod, flip);
                if (flip) {
                    mat.normal = -mat.normal;
                    mat.cc_normal = -mat.cc_normal;
                }
                if ((hit_front || double_sided) && !is_black(mat.emission)) {
                    F3 le = mat.emission;
                    if (st.k == 0) {
                        radiance = radiance + beta * le;
                    } else {
                        uint32_t li = tri_light(S, hslot);
                        F3 prev_pos = {st.prev_x, st.prev_y, st.prev_z};
                        float pl = li == NONE ? 0.0f : tri_pdf(S, li, hslot, prev_pos, surf.position);
                        float w = 1.0f;
                        if (pl > 0.0f) {
                            if (strategy == 0) w = power_heuristic(st.prev_pdf, pl);
                            else if (strategy == 1) w = 0.0f;
                        }
                        if (w > 0.0f) radiance = radiance + clamp_c(clampv, beta * le * w);
                    }
                }
                if (mat.unlit) {
                    F3 contrib = beta * xyz(mat.base);
                    radiance = radiance + (st.k == 0 ? contrib : clamp_c(clampv, contrib));
                    ended = true;
                } else if (st.k >= P.dims[3]) {
                    ended = true;
                }
                Bsdf bsdf;
                if (!ended) {
                    bsdf = make_bsdf(S.lds_sheen, mat, -d, surf.ng, hit_front);
                    if (bsdf_black(bsdf)) ended = true;
                }
                if (!ended) {
                    const F3 p = surf.position, ng = bsdf.ng;
                    const float4 pl_u = pattern(lds_sobol, st.pixel_seed, st.index, 2u + 2u * st.k);
                    const uint32_t ray_id = st.k << 16u;
                    // 3. Every NEE job of the vertex, in order (wgsl:1863-1997).
                    for (uint32_t job = 0; job <= n_punct; job++) {
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
                                F3 a = tri_pos(S, lslot, 0), bb = tri_pos(S, lslot, 1), c = tri_pos(S, lslot, 2);
                                bool has_target = false;
                                F3 q = f3(0.0f);
                                const float omega = triangle_solid_angle(p, a, bb, c);
                                if (omega >= MIN_SPHERICAL_SOLID_ANGLE) {
                                    F3 sl = sample_spherical_triangle(p, a, bb, c, pl_u.x, pl_u.y);
                                    if (dot(sl, sl) > 0.0f) {
                                        F3 n = cross(bb - a, c - a);
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
                                    q = a * w3.x + bb * w3.y + c * w3.z;
                                    has_target = true;
                                }
                                if (has_target) {
                                    F3 e1 = bb - a, e2 = c - a, rr = q - a;
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
                                        float tp = tri_pdf_omega(S, li, a, bb, c, omega, p, q);
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
                                F3 pending;
                                if (pl < 0.0f) {
                                    pending = clamp_c(clampv, beta * e.f * le);
                                } else {
                                    float w = 1.0f;
                                    if (strategy == 0) w = power_heuristic(pl, e.pdf);
                                    pending = clamp_c(clampv, beta * e.f * le * (w / pl));
                                }
                                F3 o = spawn(p, ng, l);
                                ShadowRec rec;
                                rec.o_tmax = make_float4(o.x, o.y, o.z, tmax);
                                rec.d_skip = make_float4(l.x, l.y, l.z, bitsf(skip));
                                rec.pending_rid = make_float4(pending.x, pending.y, pending.z, bitsf(rid));
                                if (st.n_shadow < A.max_shadow) {
                                    A.shadows[(size_t)st.n_shadow * A.n_slots + slot] = rec;
                                    st.n_shadow++;
                                }
                            }
                        }
                    }
                    // 4. BSDF sample -> next path ray (wgsl:1999-2039).
                    float4 pb_u = pattern(lds_sobol, st.pixel_seed, st.index, 1u + 2u * st.k);
                    BSample bs = bsdf_sample(S.lds_sheen, bsdf, F3{pb_u.x, pb_u.y, pb_u.z});
                    if (!bs.ok) {
                        ended = true;
                    } else {
                        beta = beta * (bs.f * (1.0f / bs.pdf));
                        if (!finite3(beta)) {
                            ended = true;
                        } else {
                            st.prev_pdf = bs.pdf;
                            st.prev_x = p.x;
                            st.prev_y = p.y;
                            st.prev_z = p.z;
                            if (dot(bs.l, ng) < 0.0f && mat.thickness > 0.0f && mat.transmission > 0.0f) {
                                if (hit_front && mat.atten_dist >= 0.0f) {
                                    F3 lc = vmax(mat.atten_color, f3(1.0e-6f));
                                    F3 sg = F3{-logf(lc.x), -logf(lc.y), -logf(lc.z)} / mat.atten_dist;
                                    st.sigma_x = sg.x;
                                    st.sigma_y = sg.y;
                                    st.sigma_z = sg.z;
                                    st.flags |= F_IN_MEDIUM;
                                } else {
                                    st.flags &= ~F_IN_MEDIUM;
                                }
                            }
                            if (st.k + 1u >= P.cfg[0]) {
                                float q = clampf(max3(beta), 0.05f, 1.0f);
                                if (q < 1.0f) {
                                    if (pb_u.w >= q) ended = true;
                                    else beta = beta * (1.0f / q);
                                }
                            }
                            if (!ended) {
                                st.k += 1;
                                F3 o = spawn(p, ng, bs.l);
                                st.d_x = bs.l.x;
                                st.d_y = bs.l.y;
                                st.d_z = bs.l.z;
                                st.o_x = o.x;
                                st.o_y = o.y;
                                st.o_z = o.z;
                                st.flags |= F_PATH_RAY;
                            }
                        }
                    }
                }
            }
            st.beta_x = beta.x;
            st.beta_y = beta.y;
            st.beta_z = beta.z;
            if (ended) st.flags |= F_ENDED;
        }
        st.rad_x = radiance.x;
        st.rad_y = radiance.y;
        st.rad_z = radiance.z;
        // 5. A finished path with no shadow rays left writes its sample.
        if ((st.flags & F_ENDED) && st.n_shadow == 0) {
            bool cov = (st.flags & F_COVERED) != 0;
            F3 r = cov && finite3(radiance) ? radiance : f3(0.0f);
            A.samples[st.item] = make_float4(r.x, r.y, r.z, cov ? 1.0f : 0.0f);
            st.item = NONE;
            busy = false;
        }
    }
    // 6. Empty slots claim the next items: one atomic per wave.
    const uint32_t lane = __lane_id();
    bool want = in_range && !busy;
    uint64_t m = __ballot(want);
    if (m) {
        uint32_t leader = __ffsll((unsigned long long)m) - 1u;
        uint32_t base = 0;
        if (lane == leader) base = atomicAdd(A.item_counter, (uint32_t)__popcll(m));
        base = __shfl(base, leader);
        const uint64_t lt_mask = lane == 0 ? 0ull : (~0ull >> (64u - lane));
        if (want) {
            uint32_t item = base + (uint32_t)__popcll(m & lt_mask);
            if (item < A.n_items) {
                uint32_t job, px, py, s;
                if (decode_item(A, item, job, px, py, s)) {
                    const Job &J = A.jobs[job];
                    const oxpt_params &P = J.P;
                    st.item = item;
                    st.job = job;
                    st.index = J.first + s;
                    st.pixel_seed = hash_mix(hash_mix(P.dims[2], py), px);
                    st.path_seed = hash_mix(st.pixel_seed, st.index);
                    float4 p0 = pattern(lds_sobol, st.pixel_seed, st.index, 0);
                    float width = (float)P.dims[0], height = (float)P.dims[1];
                    float ndc_x = (((float)px + p0.x) / fmaxf(width, 1.0f)) * 2.0f - 1.0f;
                    float ndc_y = 1.0f - (((float)py + p0.y) / fmaxf(height, 1.0f)) * 2.0f;
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
                    st.k = 0;
                    st.flags = F_PATH_RAY | F_COVERED;
                    st.n_shadow = 0;
                    st.prev_pdf = 0.0f;
                    st.beta_x = st.beta_y = st.beta_z = 1.0f;
                    st.rad_x = st.rad_y = st.rad_z = 0.0f;
                    st.prev_x = st.prev_y = st.prev_z = 0.0f;
                    st.sigma_x = st.sigma_y = st.sigma_z = 0.0f;
                    st.d_x = d.x;
                    st.d_y = d.y;
                    st.d_z = d.z;
                    st.o_x = o.x;
                    st.o_y = o.y;
                    st.o_z = o.z;
                    busy = true;
                } else {
                    // A ragged-edge item outside the image: nothing to trace.
                    A.samples[item] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                }
            }
        }
    }
    if (in_range) A.slots[slot] = st;
    uint64_t live = __ballot(busy);
    if (lane == 0 && live) atomicAdd(A.live + (A.iter & (WF_LIVE_RING - 1u)), (uint32_t)__popcll(live));
}

// Closest hit of every slot's path ray. Waves = occupancy asked of the
// compiler, LdsDepth = stack entries per lane kept in LDS.
template <int Waves, int LdsDepth>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(Waves))) wf_extend(WfArgs A) {
    __shared__ uint64_t lds_stack[LdsDepth > 0 ? 4 * 64 * LdsDepth : 1];
    const uint32_t slot = blockIdx.x * blockDim.x + threadIdx.x;
    if (slot >= A.n_slots) return;
    const PathSlot &st = A.slots[slot];
    uint32_t flags = st.flags;
    if (st.item == NONE || !(flags & F_PATH_RAY)) return;
    const Job &J = A.jobs[st.job];
    Scene S = job_scene(J, nullptr);
    Ray r = make_ray(F3{st.o_x, st.o_y, st.o_z}, F3{st.d_x, st.d_y, st.d_z});
    Hit h = traverse<LdsDepth>(S, r, INF, st.path_seed, st.k << 16u, st.k == 0, false, NONE,
                               to_lds(lds_stack) + (threadIdx.x >> 6) * 64u * (LdsDepth > 0 ? LdsDepth : 0));
    HitRec hr;
    hr.slot = h.slot == NONE ? NONE : (h.slot | (h.front ? 0x80000000u : 0u));
    hr.t = h.t;
    hr.b1 = h.b.y;
    hr.b2 = h.b.z;
    A.hits[slot] = hr;
}

// Any hit of every shadow ray.
template <int Waves, int LdsDepth>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(Waves))) wf_occlude(WfArgs A) {
    __shared__ uint64_t lds_stack[LdsDepth > 0 ? 4 * 64 * LdsDepth : 1];
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t slot = i % A.n_slots, j = i / A.n_slots;
    if (j >= A.max_shadow) return;
    const PathSlot &st = A.slots[slot];
    if (st.item == NONE || j >= st.n_shadow) return;
    const Job &J = A.jobs[st.job];
    Scene S = job_scene(J, nullptr);
    ShadowRec rec = A.shadows[(size_t)j * A.n_slots + slot];
    Ray r = make_ray(F3{rec.o_tmax.x, rec.o_tmax.y, rec.o_tmax.z}, F3{rec.d_skip.x, rec.d_skip.y, rec.d_skip.z});
    Hit h = traverse<LdsDepth>(S, r, rec.o_tmax.w, st.path_seed, fbits(rec.pending_rid.w), false, true,
                               fbits(rec.d_skip.w), to_lds(lds_stack) + (threadIdx.x >> 6) * 64u * (LdsDepth > 0 ? LdsDepth : 0));
    A.occluded[(size_t)j * A.n_slots + slot] = h.slot != NONE ? 1u : 0u;
}

// Per pixel: the pass's samples, in sample order, into the accumulator.
__global__ void __launch_bounds__(256) wf_reduce(WfArgs A, uint32_t job) {
    const Job &J = A.jobs[job];
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t n_pix = J.n_items;  // 64 * tiles
    if (i >= n_pix) return;
    uint32_t tile = i >> 6, in = i & 63u;
    uint32_t px = (tile % J.tiles_x()) * 8u + (in & 7u), py = (tile / J.tiles_x()) * 8u + (in >> 3);
    if (px >= J.P.dims[0] || py >= J.P.dims[1]) return;
    size_t at = (size_t)py * J.P.dims[0] + px;
    float4 acc = J.accum[at];
    const float4 *smp = A.samples + A.job_base[job] + (size_t)tile * 64u * J.count + in;
    for (uint32_t s = 0; s < J.count; s++) {
        float4 r = smp[(size_t)s * 64u];
        if (r.w > 0.0f) acc = make_float4(acc.x + r.x, acc.y + r.y, acc.z + r.z, acc.w + 1.0f);
    }
    J.accum[at] = acc;
}

__global__ void wf_clear_slots(PathSlot *slots, uint32_t n, uint32_t *item_counter, uint32_t *live) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        PathSlot st;
        st.item = NONE;
        st.n_shadow = 0;
        st.flags = 0;
        slots[i] = st;
    }
    if (i == 0) *item_counter = 0;
    if (i < WF_LIVE_RING) live[i] = 0;
}

hipError_t wf_launch_clear(const WfArgs &A, hipStream_t s) {
    uint32_t n = std::max(A.n_slots, WF_LIVE_RING);
    hipLaunchKernelGGL(wf_clear_slots, dim3((n + 255u) / 256u), dim3(256), 0, s, A.slots, A.n_slots, A.item_counter,
                       A.live);
    return hipGetLastError();
}

// variant: bits 0-3 trace form (0: 6 waves/8 stack entries in LDS, 1: 5
// waves/private stack, 2: 8 waves/8 in LDS, 3: 5 waves/16 in LDS, 4: 6
// waves/private), bits 4-7 shade occupancy (0: 3 waves, 1: 2, 2: 4).
hipError_t wf_launch_iteration(const WfArgs &A, uint32_t variant, hipStream_t s) {
    uint32_t blocks = (A.n_slots + 255u) / 256u;
    switch ((variant >> 4) & 15u) {
        case 1: hipLaunchKernelGGL(wf_shade<2>, dim3(blocks), dim3(256), 0, s, A); break;
        case 2: hipLaunchKernelGGL(wf_shade<4>, dim3(blocks), dim3(256), 0, s, A); break;
        default: hipLaunchKernelGGL(wf_shade<3>, dim3(blocks), dim3(256), 0, s, A); break;
    }
    uint64_t sh = (uint64_t)A.n_slots * A.max_shadow;
    dim3 sb((uint32_t)((sh + 255u) / 256u));
    switch (variant & 15u) {
        case 1:
            hipLaunchKernelGGL((wf_extend<5, 0>), dim3(blocks), dim3(256), 0, s, A);
            if (sh) hipLaunchKernelGGL((wf_occlude<5, 0>), sb, dim3(256), 0, s, A);
            break;
        case 2:
            hipLaunchKernelGGL((wf_extend<8, 8>), dim3(blocks), dim3(256), 0, s, A);
            if (sh) hipLaunchKernelGGL((wf_occlude<8, 8>), sb, dim3(256), 0, s, A);
            break;
        case 3:
            hipLaunchKernelGGL((wf_extend<5, 16>), dim3(blocks), dim3(256), 0, s, A);
            if (sh) hipLaunchKernelGGL((wf_occlude<5, 16>), sb, dim3(256), 0, s, A);
            break;
        case 4:
            hipLaunchKernelGGL((wf_extend<6, 0>), dim3(blocks), dim3(256), 0, s, A);
            if (sh) hipLaunchKernelGGL((wf_occlude<6, 0>), sb, dim3(256), 0, s, A);
            break;
        default:
            hipLaunchKernelGGL((wf_extend<6, 8>), dim3(blocks), dim3(256), 0, s, A);
            if (sh) hipLaunchKernelGGL((wf_occlude<6, 8>), sb, dim3(256), 0, s, A);
            break;
    }
    return hipGetLastError();
}

hipError_t wf_launch_reduce(const WfArgs &A, uint32_t job, uint32_t n_items, hipStream_t s) {
    hipLaunchKernelGGL(wf_reduce, dim3((n_items + 255u) / 256u), dim3(256), 0, s, A, job);
    return hipGetLastError();
}

}  // namespace oxd
