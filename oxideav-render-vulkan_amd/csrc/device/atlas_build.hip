// Device-built tables for gfx950, the host packer's words bit for bit:
//  - the texture atlas: level-0 decode, 2x2 box-filtered mips and binary16
//    conversion into the atlas layers (csrc/host/pack.cpp build_mips,
//    f32_to_f16; the reference's AtlasBuilder, src/pathtrace.rs:1482-1627,
//    and f32_to_f16, src/scene.rs:244-272);
//  - the environment sampling tables (pack.cpp env_words; the reference's
//    env_words, src/pathtrace.rs:1428-1461): texel weights in parallel,
//    every sum in the host's order (one thread per row, one for the
//    marginal), in f64.
//
// Levels are produced one at a time into two float4 scratch buffers; each
// kernel is one thread per texel, reads are coalesced row-major and the
// atlas store is one 8-byte write per texel.
#include "bvh_build.hpp"

#include <algorithm>

#include "oxpt/oxpt.h"

namespace oxd {
namespace {

// binary32 -> binary16, round to nearest even, overflow to infinity,
// subnormals kept, NaN stays NaN (the host packer's f32_to_f16).
__device__ inline uint32_t to_f16(float v) {
    const uint32_t x = __float_as_uint(v);
    const uint32_t sign = (x >> 16) & 0x8000u;
    const uint32_t e8 = (x >> 23) & 0xffu, frac = x & 0x7fffffu;
    if (e8 == 0xffu) return sign | 0x7c00u | (frac ? 0x200u : 0u);
    const int e = (int)e8 - 112;
    if (e >= 31) return sign | 0x7c00u;
    if (e <= 0) {
        if (e < -10) return sign;
        const uint32_t full = frac | 0x800000u, sh = (uint32_t)(14 - e);
        uint32_t q = full >> sh;
        const uint32_t r = full & ((1u << sh) - 1u), halfway = 1u << (sh - 1);
        q += (r > halfway || (r == halfway && (q & 1u))) ? 1u : 0u;
        return sign | q;
    }
    uint32_t q = ((uint32_t)e << 10) | (frac >> 13);
    const uint32_t r = frac & 0x1fffu;
    q += (r > 0x1000u || (r == 0x1000u && (q & 1u))) ? 1u : 0u;
    return sign | q;
}

__global__ void k_decode(const uint8_t *rgba8, uint32_t n, const float *lut, uint32_t srgb, float4 *out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uchar4 c = reinterpret_cast<const uchar4 *>(rgba8)[i];
    const float *col = lut + (srgb ? 0 : 256);
    out[i] = make_float4(col[c.x], col[c.y], col[c.z], lut[256 + c.w]);
}

// Level (w, h) from level (pw, ph): the four taps (2x + dx, 2y + dy),
// clamped to the edge, summed in the order (0,0) (1,0) (0,1) (1,1), x 0.25.
__global__ void k_down(const float4 *p, uint32_t pw, uint32_t ph, float4 *m, uint32_t w, uint32_t h) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= w * h) return;
    const uint32_t x = i % w, y = i / w;
    float4 s = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    for (uint32_t dy = 0; dy < 2; dy++)
        for (uint32_t dx = 0; dx < 2; dx++) {
            const uint32_t sx = min(2 * x + dx, pw - 1), sy = min(2 * y + dy, ph - 1);
            const float4 t = p[(size_t)sy * pw + sx];
            s.x += t.x;
            s.y += t.y;
            s.z += t.z;
            s.w += t.w;
        }
    m[i] = make_float4(0.25f * s.x, 0.25f * s.y, 0.25f * s.z, 0.25f * s.w);
}

__global__ void k_store(const float4 *lv, uint32_t w, uint32_t h, uint2 *atlas, uint32_t side, uint32_t layer,
                        uint32_t x0, uint32_t y0) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= w * h) return;
    const uint32_t x = i % w, y = i / w;
    const float4 t = lv[i];
    atlas[((size_t)layer * side + y0 + y) * side + x0 + x] =
        make_uint2(to_f16(t.x) | to_f16(t.y) << 16, to_f16(t.z) | to_f16(t.w) << 16);
}

// ---- environment sampling tables (env_words) ----------------------------------

__device__ inline float lum(float r, float g, float b) { return 0.2126f * r + 0.7152f * g + 0.0722f * b; }

// Per texel: radiance (max(rgb, 0) x intensity, 0 when any channel is not
// finite) and weight luminance x row weight (f64, negatives to 0).
__global__ void k_env_texel(const float *rgb, uint32_t w, uint32_t h, float intensity, const double *row_w,
                            double *wt, uint32_t *rad) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= w * h) return;
    const float *q = rgb + 3 * (size_t)i;
    float c[3] = {0.0f, 0.0f, 0.0f};
    if (isfinite(q[0]) && isfinite(q[1]) && isfinite(q[2]))
        for (int k = 0; k < 3; k++) c[k] = (q[k] < 0.0f ? 0.0f : q[k]) * intensity;  // std::max(q, 0.0f)
    for (int k = 0; k < 3; k++) rad[3 * (size_t)i + k] = __float_as_uint(c[k]);
    const double l = (double)lum(c[0], c[1], c[2]) * row_w[i / w];
    wt[i] = l > 0.0 ? l : 0.0;
}

// One thread per row: the row's weight sum in column order.
__global__ void k_env_rows(const double *wt, uint32_t w, uint32_t h, double *rows) {
    const uint32_t y = blockIdx.x * blockDim.x + threadIdx.x;
    if (y >= h) return;
    double r = 0.0;
    for (uint32_t x = 0; x < w; x++) r += wt[(size_t)y * w + x];
    rows[y] = r;
}

// One thread: the total in row order, the marginal CDF and each row's
// probability (uniform rows when the map has no positive weight).
__global__ void k_env_marginal(const double *rows, uint32_t h, uint32_t *marg, double *rowp, uint32_t *uniform) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    double total = 0.0;
    for (uint32_t y = 0; y < h; y++) total += rows[y];
    const bool uni = !(total > 0.0);
    double run = 0.0;
    marg[0] = __float_as_uint(0.0f);
    for (uint32_t y = 0; y < h; y++) {
        const double p = uni ? 1.0 / h : rows[y] / total;
        run += p;
        rowp[y] = p;
        marg[y + 1] = __float_as_uint(y + 1 == h ? 1.0f : (float)run);
    }
    *uniform = uni ? 1u : 0u;
}

// One thread per row: the conditional CDF in column order and the texel
// masses (row probability x column probability).
__global__ void k_env_conditional(const double *wt, const double *rows, const double *rowp, const uint32_t *uniform,
                                  uint32_t w, uint32_t h, uint32_t *cond, uint32_t *mass) {
    const uint32_t y = blockIdx.x * blockDim.x + threadIdx.x;
    if (y >= h) return;
    const bool row_uniform = *uniform || !(rows[y] > 0.0);
    uint32_t *c = cond + (size_t)y * (w + 1);
    c[0] = __float_as_uint(0.0f);
    double rr = 0.0;
    for (uint32_t x = 0; x < w; x++) {
        const double p = row_uniform ? 1.0 / w : wt[(size_t)y * w + x] / rows[y];
        rr += p;
        c[x + 1] = __float_as_uint(x + 1 == w ? 1.0f : (float)rr);
        mass[(size_t)y * w + x] = __float_as_uint((float)(rowp[y] * p));
    }
}

inline uint32_t blocks(size_t n) { return (uint32_t)((n + 255) / 256); }

}  // namespace

int build_env(const float *rgb, uint32_t w, uint32_t h, float intensity, const double *row_w, uint32_t *out,
              hipStream_t s, const DevAlloc &al, std::string &err) {
    const size_t n = (size_t)w * h;
    const size_t bytes = (2 * n + 2 * h + 1) * 8;
    void *scratch = nullptr;
    if (int rc = al.alloc(bytes, &scratch, "environment build: weights")) return rc;
    double *wt = static_cast<double *>(scratch), *rows = wt + n, *rowp = rows + h;
    uint32_t *uniform = reinterpret_cast<uint32_t *>(rowp + h);
    uint32_t *marg = out, *cond = marg + (h + 1), *mass = cond + (size_t)h * (w + 1), *rad = mass + n;
    k_env_texel<<<blocks(n), 256, 0, s>>>(rgb, w, h, intensity, row_w, wt, rad);
    k_env_rows<<<blocks(h), 256, 0, s>>>(wt, w, h, rows);
    k_env_marginal<<<1, 64, 0, s>>>(rows, h, marg, rowp, uniform);
    k_env_conditional<<<blocks(h), 256, 0, s>>>(wt, rows, rowp, uniform, w, h, cond, mass);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    al.release(scratch, bytes);
    if (e != hipSuccess) {
        err = std::string("hip: device environment build: ") + hipGetErrorString(e);
        return OXPT_E_BACKEND;
    }
    return 0;
}

int build_atlas(const std::vector<AtlasChainDev> &chains, const float *lut, uint16_t *atlas, uint32_t side,
                uint32_t layers, hipStream_t s, const DevAlloc &al, std::string &err) {
    size_t most = 1;
    for (const auto &c : chains) most = std::max(most, (size_t)c.w * c.h);
    void *a = nullptr, *b = nullptr;
    if (int rc = al.alloc(most * 16, &a, "atlas build: level buffer")) return rc;
    if (int rc = al.alloc(most * 16, &b, "atlas build: level buffer")) {
        al.release(a, most * 16);
        return rc;
    }
    hipError_t e = hipMemsetAsync(atlas, 0, (size_t)side * side * layers * 8, s);
    for (const auto &c : chains) {
        if (e != hipSuccess) break;
        float4 *cur = static_cast<float4 *>(a), *nxt = static_cast<float4 *>(b);
        uint32_t w = c.w, h = c.h;
        k_decode<<<blocks((size_t)w * h), 256, 0, s>>>(c.rgba8, w * h, lut, c.srgb, cur);
        for (uint32_t li = 0; li < c.levels; li++) {
            if (li >= c.skip) {
                const uint32_t *r = &c.rects[4 * (li - c.skip)];
                k_store<<<blocks((size_t)w * h), 256, 0, s>>>(cur, w, h, reinterpret_cast<uint2 *>(atlas), side, r[0],
                                                              r[1], r[2]);
            }
            if (li + 1 < c.levels) {
                const uint32_t nw = std::max(1u, w / 2), nh = std::max(1u, h / 2);
                k_down<<<blocks((size_t)nw * nh), 256, 0, s>>>(cur, w, h, nxt, nw, nh);
                std::swap(cur, nxt);
                w = nw;
                h = nh;
            }
        }
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    al.release(a, most * 16);
    al.release(b, most * 16);
    if (e != hipSuccess) {
        err = std::string("hip: device atlas build: ") + hipGetErrorString(e);
        return OXPT_E_BACKEND;
    }
    return 0;
}

}  // namespace oxd
