// Path-trace kernels for gfx950 (CDNA4, wave64).
//
// trace_persistent — the refine step (oxideav-render-vulkan
//   src/shaders/pathtrace.wgsl:1714-2073 trace_sample + trace_main). Each
//   lane owns one pixel at a time and traces that pixel's samples
//   [first, first + count) in order, adding them into the pixel's float4
//   accumulator exactly as the reference does, so results do not depend on
//   scheduling. The path is a state machine that traces exactly one ray per
//   loop iteration (the continuation ray or one NEE shadow ray), and a lane
//   whose pixel is done takes the next one at once: idle lanes of a wave are
//   counted with a ballot, one lane claims that many pixels with a single
//   atomic on the work counter, and each idle lane takes its rank among
//   them. Work items are ordered in 8x8 pixel tiles so a fresh wave starts
//   on one coherent tile. The last tiles of a launch (the tail) are claimed
//   one path at a time instead, each path's (rgb, covered) going to a sample
//   buffer that tail_reduce adds into the accumulator in sample order: lanes
//   out of pixels take tail paths rather than idling while the rest of the
//   wave finishes whole pixels, and the sums are the same additions.
// trace_tiles — the plain form: one thread per pixel, 8x8 tiles per wave,
//   no regeneration (the reference's dispatch shape).
// resolve_pt — accumulator -> RGBA8 / RGBA32F (src/shaders/resolve.wgsl:151-188).
// debug_* — integer sampler words, light-CDF picks and closest hits for the
//   bit-exact parity probes.
#include "pt_trace.hpp"

namespace oxd {

// The persistent kernel's tail: per tail pixel, its samples in sample order
// into the accumulator (the same additions the per-pixel run makes).
__global__ void __launch_bounds__(256) tail_reduce(TraceArgs A) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= A.n_items - A.split) return;
    uint32_t px, py;
    item_pixel(A.tiles_x, A.split + i, px, py);
    if (px >= A.P.dims[0] || py >= A.P.dims[1]) return;
    const size_t at = (size_t)py * A.P.dims[0] + px;
    float4 acc = A.accum[at];
    const float4 *smp = A.samples + (size_t)(i >> 6) * 64u * A.count + (i & 63u);
    for (uint32_t s = 0; s < A.count; s++) {
        const float4 r = smp[(size_t)s * 64u];
        if (r.w > 0.0f) acc = make_float4(acc.x + r.x, acc.y + r.y, acc.z + r.z, acc.w + 1.0f);
    }
    A.accum[at] = acc;
}

// Plain form: one thread per pixel of an 8x8 tile, samples in order (built
// for 4 waves per SIMD, like the default persistent form).
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4))) trace_tiles(TraceArgs A) {
    __shared__ uint32_t lds_sobol[OXPT_SOBOL_WORDS];
    __shared__ float lds_sheen[OXPT_SHEEN_WORDS];
    __shared__ float lds_cdf[LDS_MARG_WORDS + LDS_ECDF_WORDS];
    stage_tables(A.tables, A.P, lds_sobol, lds_sheen);
    Scene S = make_scene(A, lds_sheen);
    stage_light_cdfs(A.tables, A.P, lds_cdf, S);
    uint32_t px, py;
    item_pixel(A.tiles_x, blockIdx.x * 64u + threadIdx.x, px, py);
    const uint32_t w = A.P.dims[0], h = A.P.dims[1];
    if (px >= w || py >= h) return;
    float4 acc = A.accum[(size_t)py * w + px];
    PathState st;
    for (uint32_t i = 0; i < A.count; i++) {
        init_path(A.P, lds_sobol, px, py, A.first + i, st);
        bool covered;
        while (!step_path(A.P, S, lds_sobol, st, &covered)) {
        }
        if (covered) {
            F3 r = finite3(st.radiance) ? st.radiance : f3(0.0f);
            acc = make_float4(acc.x + r.x, acc.y + r.y, acc.z + r.z, acc.w + 1.0f);
        }
    }
    A.accum[(size_t)py * w + px] = acc;
}

// Per tail pixel of every job (blockIdx.y = job): the launch's samples, in
// sample order, into the accumulator (pixel runs were added by the trace
// kernel). The grid is fixed (grid-stride over pixels, jobs past the
// device-side count exit), so the step stays capturable.
__global__ void __launch_bounds__(256) batch_reduce(BatchArgs B, const Job *__restrict__ jobs) {
    const uint32_t jb = blockIdx.y;
    if (jb >= min(*B.n_jobs, B.max_jobs)) return;
    const Job &J = jobs[jb];
    const uint32_t tiles_x = J.tiles_x();
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < J.n_items - J.split; i += gridDim.x * blockDim.x) {
        const uint32_t tt = i >> 6, in = i & 63u, tile = (J.split >> 6) + tt;
        const uint32_t px = (tile % tiles_x) * 8u + (in & 7u), py = (tile / tiles_x) * 8u + (in >> 3);
        if (px >= J.P.dims[0] || py >= J.P.dims[1]) continue;
        const size_t at = (size_t)py * J.P.dims[0] + px;
        float4 acc = J.accum[at];
        const float4 *smp = B.samples + J.sample_base + (size_t)tt * 64u * J.count + in;
        for (uint32_t s = 0; s < J.count; s++) {
            const float4 r = smp[(size_t)s * 64u];
            if (r.w > 0.0f) acc = make_float4(acc.x + r.x, acc.y + r.y, acc.z + r.z, acc.w + 1.0f);
        }
        J.accum[at] = acc;
    }
}

// The persistent kernel's two queue counters (pixels, tail paths).
__global__ void reset_counter(uint32_t *c) {
    if (threadIdx.x < 2) c[threadIdx.x] = 0;
}

// The batched launch's two counters per job (pixels, tail paths).
__global__ void reset_counters(uint32_t *c, const uint32_t *n, uint32_t cap) {
    uint32_t m = 2u * min(*n, cap);
    for (uint32_t i = threadIdx.x; i < m; i += blockDim.x) c[i] = 0;
}

// ---- Resolve (resolve.wgsl:38-75,151-188) ------------------------------

__device__ __forceinline__ float srgb_to_linear(float c) {
    if (c <= 0.04045f) return c / 12.92f;
    return powf((c + 0.055f) / 1.055f, 2.4f);
}
__device__ __forceinline__ float linear_to_srgb(float c) {
    float x = clampf(c, 0.0f, 1.0f);
    if (x <= 0.0031308f) return 12.92f * x;
    return 1.055f * powf(x, 1.0f / 2.4f) - 0.055f;
}
__device__ __forceinline__ float aces(float x) {
    float y = x * 0.6f;
    return clampf((y * (2.51f * y + 0.03f)) / (y * (2.43f * y + 0.59f) + 0.14f), 0.0f, 1.0f);
}
__device__ __forceinline__ F3 tone_map(uint32_t op, F3 rgb_in) {
    F3 rgb = vmax(rgb_in, f3(0.0f));
    if (op == 1) {
        float l = dot(rgb, F3{0.2126f, 0.7152f, 0.0722f});
        if (l <= 0.0f) return f3(0.0f);
        return vmin(rgb / (1.0f + l), f3(1.0f));
    }
    if (op == 2) return {aces(rgb.x), aces(rgb.y), aces(rgb.z)};
    return vmin(rgb, f3(1.0f));
}
__device__ __forceinline__ uint32_t unorm8(float v) { return (uint32_t)floorf(clampf(v, 0.0f, 1.0f) * 255.0f + 0.5f); }

__global__ void resolve_pt(const float4 *__restrict__ accum, uint32_t n, oxpt_resolve_params rp, void *out) {
    uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n) return;
    bool hdr = rp.cfg[3] == 1;
    F3 bg_lin = {srgb_to_linear(rp.background[0]), srgb_to_linear(rp.background[1]), srgb_to_linear(rp.background[2])};
    float ba = rp.background[3];
    float ns = rp.exposure[1];
    float4 a = accum[idx];
    float4 o;
    if (a.w <= 0.0f || ns <= 0.0f) {
        o = make_float4(bg_lin.x, bg_lin.y, bg_lin.z, ba);
    } else {
        F3 m = F3{a.x, a.y, a.z} / a.w;
        float f = clampf(a.w / ns, 0.0f, 1.0f);
        F3 c = m;
        if (!hdr) c = tone_map(rp.cfg[2], m * rp.exposure[0]);
        if (f >= 1.0f) {
            o = make_float4(c.x, c.y, c.z, 1.0f);
        } else {
            float a_sum = f + (1.0f - f) * ba;
            if (a_sum > 0.0f) {
                F3 r = (f * c + (1.0f - f) * bg_lin * ba) / a_sum;
                o = make_float4(r.x, r.y, r.z, a_sum);
            } else {
                F3 r = f * c + (1.0f - f) * bg_lin;
                o = make_float4(r.x, r.y, r.z, 0.0f);
            }
        }
    }
    if (hdr) {
        reinterpret_cast<float4 *>(out)[idx] = o;
    } else {
        uint32_t px = unorm8(linear_to_srgb(o.x)) | unorm8(linear_to_srgb(o.y)) << 8 |
                      unorm8(linear_to_srgb(o.z)) << 16 | unorm8(o.w) << 24;
        reinterpret_cast<uint32_t *>(out)[idx] = px;
    }
}

// ---- Parity probes -------------------------------------------------------

__global__ void debug_pattern_words(const uint32_t *__restrict__ tables, uint32_t off_sobol, uint32_t seed, uint32_t n,
                                    const uint32_t *x, const uint32_t *y, const uint32_t *index, const uint32_t *p,
                                    uint4 *out) {
    __shared__ uint32_t lds_sobol[OXPT_SOBOL_WORDS];
    for (uint32_t i = threadIdx.x; i < OXPT_SOBOL_WORDS; i += blockDim.x) lds_sobol[i] = tables[off_sobol + i];
    __syncthreads();
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t pixel_seed = hash_mix(hash_mix(seed, y[i]), x[i]);
    uint32_t w[4];
    sobol_owen_words(lds_sobol, index[i], hash_mix(pixel_seed, p[i]), w);
    out[i] = make_uint4(w[0], w[1], w[2], w[3]);
}

// The claim step's exact divisions (udiv with the host's udiv_magic pairs).
__global__ void debug_udiv(uint32_t n, const uint32_t *num, const uint32_t *m, const uint32_t *l, uint32_t *out) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = udiv(num[i], m[i], l[i]);
}

// Texel indices of one coordinate as filter_level forms them: the nearest
// tap, and the bilinear footprint's two taps, wrapped by the sampler mode.
__global__ void debug_texel_index(uint32_t n, const float *u, const int32_t *size, const uint32_t *mode, int32_t *out3) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float v = u[i];
    if (!finite(v)) v = 0.0f;
    v = clampf(v, -1.0e6f, 1.0e6f);
    const int32_t w = size[i];
    const float x = v * (float)w;
    const int32_t x0 = sat_i32(floorf(x - 0.5f));
    out3[3 * i] = wrap_index(sat_i32(floorf(x)), w, mode[i]);
    out3[3 * i + 1] = wrap_index(x0, w, mode[i]);
    out3[3 * i + 2] = wrap_index(next_i32(x0), w, mode[i]);
}

__global__ void debug_light_picks(TraceArgs A, uint32_t n, const float *u, uint32_t *out3) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const Scene S = make_scene(A, nullptr);
    float v = u[i];
    out3[3 * i] = S.n_emissive > 0 ? pick_elight(S, v) : NONE;
    if (A.P.env[2]) {
        uint32_t yy = env_row(S, v);
        out3[3 * i + 1] = yy;
        out3[3 * i + 2] = env_col(S, yy, v);
    } else {
        out3[3 * i + 1] = NONE;
        out3[3 * i + 2] = NONE;
    }
}

__global__ void debug_closest_hit(TraceArgs A, uint32_t n, const float *o3, const float *d3, uint32_t *out2) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const Scene S = make_scene(A, nullptr);
    F3 o = {o3[3 * i], o3[3 * i + 1], o3[3 * i + 2]}, d = {d3[3 * i], d3[3 * i + 1], d3[3 * i + 2]};
    Hit h = traverse(S, make_ray(o, d), INF, 0, 0, false, false, NONE);
    out2[2 * i] = h.slot;
    out2[2 * i + 1] = h.slot == NONE ? 0u : fbits(h.t);
}

}  // namespace oxd

// ---- Host launchers (called by the runtime) ------------------------------

namespace oxd {

// Kernel forms (oxpt_tuning.kernel): 0 the default (4 waves per SIMD),
// 1 plain per-pixel, 2..6 persistent built for 3, 4, 5, 6, 8 waves per SIMD,
// 8 persistent built for 2 waves per SIMD (256 VGPRs).
// mode 1 (no punctual lights) and mode 2 (1..31 punctual lights, NEE queue
// in args.nee_q) select the early-continuation builds of forms 0 and 2..5
// (3, 4, 5 and 6 waves per SIMD; form 0 is the 4-waves build), and in
// those modes `opaque` the walk built for all-opaque scenes; same images.
static bool early_form(uint32_t kernel, uint32_t mode) {
    return (mode == 1 || mode == 2) && (kernel == 0 || (kernel >= 2 && kernel <= 5));
}

hipError_t launch_trace(const TraceArgs &args, uint32_t kernel, uint32_t block, uint32_t grid, uint32_t lds_depth,
                        uint32_t mode, bool opaque, hipStream_t s) {
    if (kernel == 1) {
        uint32_t blocks = args.n_items / 64u;
        hipLaunchKernelGGL(trace_tiles, dim3(blocks), dim3(64), 0, s, args);
        return hipGetLastError();
    }
    if (early_form(kernel, mode) && mode == 2 && args.nee_q == nullptr) return hipErrorInvalidValue;
    hipLaunchKernelGGL(reset_counter, dim3(1), dim3(64), 0, s, args.work);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const size_t shm = (size_t)(block / 64u) * lds_depth * 64u * sizeof(uint64_t);
    if (!early_form(kernel, mode)) launch_persistent_m0(args, kernel, block, grid, lds_depth, shm, false, s);
    else if (mode == 2) launch_persistent_m2(args, kernel, block, grid, lds_depth, shm, opaque, s);
    else launch_persistent_m1(args, kernel, block, grid, lds_depth, shm, opaque, s);
    if (args.split < args.n_items) {
        e = hipGetLastError();
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(tail_reduce, dim3((args.n_items - args.split + 255u) / 256u), dim3(256), 0, s, args);
    }
    return hipGetLastError();
}

// Batched forms: the same numbering (default: the 4-waves build).
hipError_t launch_trace_batch(const BatchArgs &args, uint32_t kernel, uint32_t block, uint32_t grid,
                              uint32_t lds_depth, uint32_t mode, bool opaque, hipStream_t s) {
    if (early_form(kernel, mode) && mode == 2 && args.nee_q == nullptr) return hipErrorInvalidValue;
    hipLaunchKernelGGL(reset_counters, dim3(1), dim3(256), 0, s, args.counters, args.n_jobs, args.max_jobs);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const size_t shm = (size_t)(block / 64u) * lds_depth * 64u * sizeof(uint64_t);
    if (!early_form(kernel, mode)) launch_batch_m0(args, kernel, block, grid, lds_depth, shm, false, s);
    else if (mode == 2) launch_batch_m2(args, kernel, block, grid, lds_depth, shm, opaque, s);
    else launch_batch_m1(args, kernel, block, grid, lds_depth, shm, opaque, s);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(batch_reduce, dim3(64, args.max_jobs), dim3(256), 0, s, args, args.jobs);
    return hipGetLastError();
}

hipError_t launch_resolve(const float4 *accum, uint32_t n, const oxpt_resolve_params &rp, void *out, hipStream_t s) {
    uint32_t blocks = (n + 255u) / 256u;
    if (blocks == 0) return hipSuccess;
    hipLaunchKernelGGL(resolve_pt, dim3(blocks), dim3(256), 0, s, accum, n, rp, out);
    return hipGetLastError();
}

hipError_t launch_debug_pattern(const uint32_t *tables, uint32_t off_sobol, uint32_t seed, uint32_t n, const uint32_t *x,
                                const uint32_t *y, const uint32_t *index, const uint32_t *p, uint32_t *out4,
                                hipStream_t s) {
    uint32_t blocks = (n + 255u) / 256u;
    if (blocks == 0) return hipSuccess;
    hipLaunchKernelGGL(debug_pattern_words, dim3(blocks), dim3(256), 0, s, tables, off_sobol, seed, n, x, y, index, p,
                       reinterpret_cast<uint4 *>(out4));
    return hipGetLastError();
}

hipError_t launch_debug_udiv(uint32_t n, const uint32_t *num, const uint32_t *m, const uint32_t *l, uint32_t *out,
                             hipStream_t s) {
    uint32_t blocks = (n + 255u) / 256u;
    if (blocks == 0) return hipSuccess;
    hipLaunchKernelGGL(debug_udiv, dim3(blocks), dim3(256), 0, s, n, num, m, l, out);
    return hipGetLastError();
}

hipError_t launch_debug_texel_index(uint32_t n, const float *u, const int32_t *size, const uint32_t *mode,
                                    int32_t *out3, hipStream_t s) {
    uint32_t blocks = (n + 255u) / 256u;
    if (blocks == 0) return hipSuccess;
    hipLaunchKernelGGL(debug_texel_index, dim3(blocks), dim3(256), 0, s, n, u, size, mode, out3);
    return hipGetLastError();
}

hipError_t launch_debug_picks(const TraceArgs &args, uint32_t n, const float *u, uint32_t *out3, hipStream_t s) {
    uint32_t blocks = (n + 255u) / 256u;
    if (blocks == 0) return hipSuccess;
    hipLaunchKernelGGL(debug_light_picks, dim3(blocks), dim3(256), 0, s, args, n, u, out3);
    return hipGetLastError();
}

hipError_t launch_debug_hits(const TraceArgs &args, uint32_t n, const float *o3, const float *d3, uint32_t *out2,
                             hipStream_t s) {
    uint32_t blocks = (n + 255u) / 256u;
    if (blocks == 0) return hipSuccess;
    hipLaunchKernelGGL(debug_closest_hit, dim3(blocks), dim3(256), 0, s, args, n, o3, d3, out2);
    return hipGetLastError();
}

}  // namespace oxd
