// Device texture atlas for gfx950: level-0 decode, 2x2 box-filtered mips
// and binary16 conversion into the atlas layers, the host packer's texels
// bit for bit (csrc/host/pack.cpp build_mips, f32_to_f16). The reference
// builds its Rgba16Float atlas on the CPU (AtlasBuilder,
// src/pathtrace.rs:1482-1627; f32_to_f16, src/scene.rs:244-272).
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

inline uint32_t blocks(size_t n) { return (uint32_t)((n + 255) / 256); }

}  // namespace

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
