// Device scene preparation: the binned-SAH BVH, the leaf-ordered triangle
// and attribute words and the 4-wide nodes, built on the GPU from the
// triangle soup. The words are the host packer's (csrc/host/bvh.cpp,
// csrc/host/pack.cpp) bit for bit; the reference does this on the host in
// GpuPathTracer::upload (src/pathtrace.rs:970-1239).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <functional>
#include <string>
#include <vector>

namespace oxd {

// The soup in device memory, one record per triangle in input order. An
// attribute array may be null when no triangle's flags select it.
struct SoupDev {
    uint32_t n = 0;
    const float *pos = nullptr;       // 9 per triangle
    const float *normals = nullptr;   // 9
    const float *tangents = nullptr;  // 12
    const float *uv0 = nullptr;       // 6
    const float *uv1 = nullptr;       // 6
    const float *colors = nullptr;    // 12
    const uint32_t *material = nullptr;
    const uint8_t *flags = nullptr;   // OXPT_F_* per triangle
    const uint32_t *light_of = nullptr;  // emissive-light index or OXPT_NONE
    // Lights found on the device instead: per-material emission luminance.
    const float *mat_lum = nullptr;
    uint32_t n_mat = 0;
};

// Device allocation through the caller (the context's budget accounting).
struct DevAlloc {
    std::function<int(size_t, void **, const char *)> alloc;
    std::function<void(void *, size_t)> release;
};

struct BuiltGeometry {
    float4 *geom = nullptr;  // binary nodes | triangles | pad | 4-wide nodes
    size_t geom_vec4 = 0;
    float4 *attrs = nullptr;
    size_t attrs_vec4 = 0;
    uint32_t node_count = 0, tri_base = 0, wide_base = 0;
    uint32_t bvh_depth = 0, wide_depth = 0;
    double ms = 0.0;  // wall time of the build, upload of the soup excluded
    // Lights found on the device: their records {cdf, pmf, triangle, slot}.
    uint32_t n_emissive = 0;
    std::vector<uint32_t> emissive_records;
};

// Subtrees of at most this many triangles are finished by one wave in LDS.
constexpr uint32_t kBuildSmall = 512;
// Largest soup the device builder takes (its scans cover 2^24 slots).
constexpr uint32_t kBuildMaxTris = 1u << 24;

// Builds everything on stream `s`. `emissive_slots` (n_emissive words,
// device memory, may be null) receives the leaf slot of every emissive
// triangle, indexed by its light index. With soup.mat_lum set the lights
// are found here instead (n_emissive, emissive_slots ignored) and their
// records returned in out->emissive_records. Returns 0 or a negative OXPT_E_*
// code with `err` set; temporaries are released either way.
int build_geometry(const SoupDev &soup, uint32_t n_emissive, uint32_t *emissive_slots, hipStream_t s,
                   const DevAlloc &a, BuiltGeometry *out, std::string &err);

// One mip chain of the device-built texture atlas: level 0 RGBA8 in device
// memory, its level count, the levels dropped before the first kept one,
// and the rectangles (layer, x, y, w | h << 16) of the kept levels.
struct AtlasChainDev {
    const uint8_t *rgba8 = nullptr;
    uint32_t w = 0, h = 0, srgb = 0, levels = 0, skip = 0;
    std::vector<uint32_t> rects;  // 4 words per kept level
};

// Decodes, filters and converts every chain into `atlas` (side x side x
// layers RGBA binary16, cleared first) on stream `s`; lut: 512 device
// floats (the host texel_decode_lut). Rectangles must lie inside the
// atlas and match the level sizes (checked by the caller).
int build_atlas(const std::vector<AtlasChainDev> &chains, const float *lut, uint16_t *atlas, uint32_t side,
                uint32_t layers, hipStream_t s, const DevAlloc &a, std::string &err);

// Environment section (marginal h + 1, conditional h x (w + 1), mass w h,
// radiance 3 w h words) into `out` on stream `s`; rgb (w h x 3) and row_w
// (h row weights sin(pi (y + 0.5) / h)) in device memory.
int build_env(const float *rgb, uint32_t w, uint32_t h, float intensity, const double *row_w, uint32_t *out,
              hipStream_t s, const DevAlloc &a, std::string &err);

}  // namespace oxd
