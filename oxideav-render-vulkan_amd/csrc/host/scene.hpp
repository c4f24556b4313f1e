// Host scene model: the triangle soup, materials, textures, lights and
// camera that the packer turns into the kernels' words. It stands in for
// the reference's prepared trace scene (oxideav-render `TraceScene`,
// consumed at oxideav-render-vulkan src/pathtrace.rs:970-1239).
#pragma once

#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "oxpt/oxpt.h"

namespace oxpt {

struct Texture {
    uint32_t w = 0, h = 0;       // 0 x 0: absent (its material slots sample nothing)
    std::vector<uint8_t> rgba8;  // level 0
    oxpt_sampler sampler{2, 2, 1, 5};
    std::string uri;             // named texture: pixels come from a texture resolver
};

// Per-triangle record: positions plus optional per-corner attributes.
struct Soup {
    std::vector<float> pos;       // 9 per tri
    std::vector<float> normals;   // 9 per tri (or empty)
    std::vector<float> tangents;  // 12 per tri
    std::vector<float> uv0;       // 6 per tri
    std::vector<float> uv1;       // 6 per tri
    std::vector<float> colors;    // 12 per tri
    std::vector<uint32_t> material;
    std::vector<uint8_t> flags;   // OXPT_F_* per tri
    size_t count() const { return material.size(); }
};

struct HostScene {
    Soup soup;
    std::vector<oxpt_material> materials;
    std::vector<Texture> textures;
    std::vector<oxpt_light> lights;
    bool has_camera = false;
    oxpt_camera camera{};
    // Content generation: bumped on every edit so cached packs notice.
    uint64_t generation = 0;
};

struct EnvMap {
    uint32_t w = 0, h = 0;
    std::vector<float> rgb;  // w*h*3
    float intensity = 1.0f;
};

// Resolved camera (Camera::resolve): orthonormal basis + half extents.
struct CameraFrame {
    float eye[3], forward[3], side[3], up[3];
    float half_w, half_h;
    bool ortho;
};

// What Camera::resolve needs from an uploaded scene: its camera (if any)
// and its bounds, for framing scenes without one.
struct CameraSource {
    bool has_camera = false;
    oxpt_camera camera{};
    bool empty = true;
    float lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
};
CameraSource camera_source(const HostScene &s);
CameraFrame resolve_camera(const CameraSource &s, uint32_t w, uint32_t h);

// Packed scene: the exact words of include/oxpt/layout.h.
struct Packed {
    std::vector<float> geom;      // float4 words
    std::vector<float> attrs;     // float4 words
    std::vector<uint32_t> tables; // scene part (no environment)
    std::vector<uint16_t> atlas;  // binary16 RGBA
    uint32_t atlas_side = 1, atlas_layers = 1;
    uint32_t node_count = 0, tri_base = 0;
    uint32_t n_punctual = 0, n_emissive = 0;
    uint32_t off_sobol = 0, off_sheen = 0, off_mat = 0, off_plights = 0, off_elights = 0;
    uint32_t bvh_depth = 0;
    uint32_t wide_base = 0;   // first float4 of the 4-wide nodes in geom (0: none)
    uint32_t wide_depth = 0;
    // One mip chain per (texture, colour space) a material slot binds:
    // {texture, srgb, table word of its descriptor, 0} (oxpt_atlas_chain).
    std::vector<oxpt_atlas_chain> chains;
};

struct EnvWords {
    std::vector<uint32_t> words;
    uint32_t w = 0, h = 0, marg = 0, cond = 0, mass = 0, rad = 0;
};

// Build the BVH (binned SAH, 32-byte nodes) over the soup: node words (8
// per node) and the leaf-ordered triangle ids. Returns the tree depth.
uint32_t build_bvh(const Soup &soup, std::vector<uint32_t> &node_words, std::vector<uint32_t> &order);

// Traversal-stack entries of the 4-wide walk (the kernels' capacity).
constexpr uint32_t kWideStack = 96;
// Collapse the binary tree into 4-wide nodes (32 words each, see bvh.cpp).
// Returns the wide depth, or 0 (and no nodes) when the tree does not fit
// the wide walk's limits (leaves of at most 30 slots, 2^27 slots or nodes,
// 3 * depth <= kWideStack); the kernels then walk the binary tree.
uint32_t build_wide(const std::vector<uint32_t> &node_words, std::vector<float> &wide);

// Pack everything but the environment (GpuPathTracer::upload). With
// geometry = false the BVH, triangle and attribute words are left to the
// device builder (oxpt_scene_build): geom / attrs stay empty and the
// emissive records' slot words hold OXPT_NONE. light_of_out receives each
// triangle's emissive-light index (OXPT_NONE for none).
// With atlas = false the texels are left to the device builder
// (oxpt_atlas_desc): descriptors and placement are written, pk.atlas stays
// empty.
// With emissive = false no emissive records are written (n_emissive = 0):
// the device builder finds the lights from material_emission() and appends
// the records at the end of the tables.
// textures (when given) replaces s.textures: the resolved texture list of a
// scene whose textures are named by uri (oxpt_tracer_set_texture_resolver).
Packed pack_scene(const HostScene &s, bool geometry = true, std::vector<uint32_t> *light_of_out = nullptr,
                  bool atlas = true, bool emissive = true, const std::vector<Texture> *textures = nullptr);
// Per material: luminance of the emissive factor for emitting materials
// (not unlit, some component > 0), else 0 (EmissiveLights::build's weight
// per unit area).
std::vector<float> material_emission(const HostScene &s);
// Environment section (env_words, src/pathtrace.rs:1428-1461).
EnvWords env_words(const EnvMap &e);

// Tables with the environment appended and the 176-byte Params
// (ensure_binding, src/pathtrace.rs:847-926).
// env_on_host = false: Params get the environment layout but its words
// are not appended (oxpt_scene_set_tables_env builds them on the device).
void bind(const Packed &pk, const EnvMap *env, const CameraFrame &cam, const oxpt_render_options &o, uint32_t w,
          uint32_t h, std::vector<uint32_t> &tables, oxpt_params &P, bool env_on_host = true);
// Offsets of the environment section (no words), and the per-row weights
// sin(pi (y + 0.5) / h) its texel weights use.
EnvWords env_layout(uint32_t w, uint32_t h);
std::vector<double> env_row_weights(uint32_t h);

// Device scene preparation (oxpt_scene_build) of a host scene: pk receives
// the host-packed tables with the device-found emissive records and the
// geometry counts Params need; geom / attrs / atlas stay on the device.
int device_prepare(oxpt_ctx *ctx, const HostScene &s, Packed &pk, oxpt_scene **out, oxpt_build_info *info,
                   const std::vector<Texture> *textures = nullptr);
// Upload the scene tables, with the environment section built on the
// device after them when env is a full map (else the tables alone).
int upload_tables(oxpt_ctx *ctx, oxpt_scene *dev, const std::vector<uint32_t> &scene_tables, const EnvMap *env);

// Stream-ordered helpers of the serving loop (csrc/device/runtime.cpp).
// ctx_mark records a point in ctx's stream; mark_wait blocks the calling
// thread until the stream has passed it, then releases the mark.
int ctx_mark(oxpt_ctx *c, void **mark);
int mark_wait(oxpt_ctx *c, void *mark);
// Device memory counted in the context's budget (staging for readbacks).
int device_alloc(oxpt_ctx *c, size_t bytes, void **p);
void device_free(oxpt_ctx *c, void *p, size_t bytes);
// Device -> host copy on ctx's stream; returns when it is done.
int copy_to_host(oxpt_ctx *c, void *host, const void *dev, size_t bytes);

void sobol_directions(uint32_t out[OXPT_SOBOL_WORDS]);
void sheen_table(float out[OXPT_SHEEN_WORDS]);
uint16_t f32_to_f16(float v);
// Level-0 texel decode: [0, 256) sRGB EOTF of i / 255, [256, 512) i / 255.
void texel_decode_lut(float out[512]);

}  // namespace oxpt

struct oxpt_host_scene {
    oxpt::HostScene s;
};
struct oxpt_env_map {
    oxpt::EnvMap e;
};
// A packed scene on the host (no device): words + params.
struct oxpt_packed {
    oxpt::Packed pk;
    std::vector<uint32_t> tables;
    oxpt_params P;
};
