/*
 * oxpt.h — C ABI of liboxpt, the MI355X (gfx950) path tracer.
 *
 * Two layers, both `extern "C"`, plain pointers and sizes:
 *
 *  1. The device boundary (oxpt_ctx / oxpt_scene / oxpt_accum): the HIP
 *     kernels behind opaque handles. This is what the reference's GPU
 *     backend does through wgpu (oxideav-render-vulkan src/pathtrace.rs);
 *     a Rust `-sys` crate would bind exactly these declarations.
 *  2. The host tracer (oxpt_tracer): the progressive `GpuPathTracer` API
 *     with the reference's reset rules (src/pathtrace.rs:309-725), built
 *     in C++ on top of layer 1, plus the scene builders it consumes.
 *
 * Conventions (src/pathtrace.rs, src/gpu.rs):
 *  - every call returns an int status: OXPT_OK (0) or a negative OXPT_E_*;
 *    the message is available from oxpt_last_error / oxpt_tracer_last_error
 *    (the reference's Error::Backend(String) / Error::InvalidOptions text);
 *  - the caller owns host buffers, the library owns device memory;
 *  - one handle per thread, or external locking (the reference's `&mut self`);
 *    device opening is serialised process-wide (src/gpu.rs:978-992);
 *  - refine is asynchronous (stream-ordered); resolve / image / hdr / finish
 *    block (src/pathtrace.rs:578-583, src/gpu.rs:1095-1107).
 */
#ifndef OXPT_OXPT_H
#define OXPT_OXPT_H

#include <stddef.h>
#include <stdint.h>

#include "oxpt/layout.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes. */
#define OXPT_OK 0
#define OXPT_E_BACKEND -1  /* Error::Backend: device, limits, allocation, launch */
#define OXPT_E_INVALID -2  /* Error::InvalidOptions: e.g. oversize dimensions */
#define OXPT_E_ARG -3      /* bad argument (null handle, short buffer) */

/* =====================================================================
 * Layer 1: device boundary.
 * ===================================================================== */

typedef struct oxpt_ctx oxpt_ctx;
typedef struct oxpt_scene oxpt_scene;
typedef struct oxpt_accum oxpt_accum;
typedef struct oxpt_image oxpt_image;

/* Packed scene words (include/oxpt/layout.h), copied at upload. Replaces the
 * buffer creation in GpuPathTracer::upload / ensure_binding
 * (src/pathtrace.rs:970-1239, 847-966). */
typedef struct oxpt_scene_desc {
    const float *geom;       /* float4 words */
    size_t geom_vec4;
    const float *attrs;      /* float4 words */
    size_t attrs_vec4;
    const uint32_t *tables;
    size_t table_words;
    const uint16_t *atlas;   /* binary16 RGBA texels, may be NULL */
    uint32_t atlas_side;
    uint32_t atlas_layers;
} oxpt_scene_desc;

/* Opens device `device_ordinal` (open_device_blocking, src/gpu.rs:987-1056).
 * OXPT_E_BACKEND when no such HIP device exists. */
int oxpt_ctx_create(int device_ordinal, oxpt_ctx **out);
/* A second context on base's device with its own stream (and its own launch
 * scratch), sharing base's objects and memory budget: scenes and
 * accumulators made through either may be used with the other, so uploads
 * and readbacks on one overlap launches on the other. Work on the two
 * streams is not ordered; the caller orders it (finish). Objects belong to
 * base: destroy peers and objects before base. No counterpart in the
 * reference (wgpu has one queue per device). */
int oxpt_ctx_create_peer(oxpt_ctx *base, oxpt_ctx **out);
void oxpt_ctx_destroy(oxpt_ctx *ctx);
/* Text of the last failure on this context ("" if none). */
const char *oxpt_last_error(const oxpt_ctx *ctx);
/* "<device name> (gfx arch, HIP)" — adapter_summary (src/pathtrace.rs:464-467). */
const char *oxpt_ctx_summary(const oxpt_ctx *ctx);
/* Name of the trace kernel the last refine / refine_batch launched (for
 * profiling and reports; no counterpart in the reference). */
const char *oxpt_ctx_last_kernel(const oxpt_ctx *ctx);
/* Device memory budget in bytes that uploads and accumulators may use
 * (default: the device's total memory); allocation past it fails with
 * OXPT_E_BACKEND, as check_size does (src/pathtrace.rs:567-576). */
int oxpt_ctx_set_budget(oxpt_ctx *ctx, uint64_t bytes);
uint64_t oxpt_ctx_bytes_in_use(const oxpt_ctx *ctx);

int oxpt_scene_upload(oxpt_ctx *ctx, const oxpt_scene_desc *desc, oxpt_scene **out);
/* Replace the table words only (environment change, ensure_binding,
 * src/pathtrace.rs:847-869); geometry and atlas stay resident. */
int oxpt_scene_set_tables(oxpt_ctx *ctx, oxpt_scene *scene, const uint32_t *tables, size_t words);
/* Equirectangular environment for the device-built sampling tables. */
typedef struct oxpt_env_desc {
    uint32_t w, h;
    const float *rgb;          /* w * h * 3 floats */
    float intensity;
    uint32_t pad;
    const double *row_weight;  /* h values: sin(pi (y + 0.5) / h) */
} oxpt_env_desc;
/* As oxpt_scene_set_tables, then the environment section built on the
 * device right after the `words` table words: marginal CDF (h + 1),
 * conditional CDFs (h x (w + 1)), texel mass (w h), radiance (3 w h);
 * texel weight = luminance(max(rgb, 0) x intensity) x row_weight[y] in
 * f64, sums in row order (env_words, src/pathtrace.rs:1428-1461; used by
 * env_sample / env_pdf, src/shaders/pathtrace.wgsl:1594-1654). */
int oxpt_scene_set_tables_env(oxpt_ctx *ctx, oxpt_scene *scene, const uint32_t *tables, size_t words,
                              const oxpt_env_desc *env);
void oxpt_scene_destroy(oxpt_scene *scene);
/* Device bytes held by a scene. */
uint64_t oxpt_scene_bytes(const oxpt_scene *scene);

/* Texture atlas built on the device: each chain decodes one texture's
 * RGBA8 level 0 (colour channels through decode_lut[0..255] when srgb,
 * else decode_lut[256..511]; alpha through decode_lut[256..511]), box-
 * filters its mips (2x2, edge texels repeated) and writes every kept level
 * as binary16 RGBA (round to nearest even) at the rectangles its texture
 * descriptor in the tables lists ([levels, flags, w, h, skip, kept, 0, 0]
 * + kept x {layer, x, y, w | h << 16}). Replaces AtlasBuilder's texel work
 * and f32_to_f16 (src/pathtrace.rs:1482-1627, src/scene.rs:244-272). */
typedef struct oxpt_texture_desc {
    uint32_t w, h;
    const uint8_t *rgba8;  /* w * h * 4 bytes */
} oxpt_texture_desc;
typedef struct oxpt_atlas_chain {
    uint32_t texture;      /* index into textures */
    uint32_t srgb;         /* 1: colour channels are sRGB-encoded */
    uint32_t desc_offset;  /* table word of the chain's texture descriptor */
    uint32_t pad;
} oxpt_atlas_chain;
typedef struct oxpt_atlas_desc {
    size_t n_textures;
    const oxpt_texture_desc *textures;
    size_t n_chains;
    const oxpt_atlas_chain *chains;
    const float *decode_lut;  /* 512 floats */
} oxpt_atlas_desc;

/* Triangle soup for the device build, one record per triangle in input
 * order (the prepared scene's triangles that GpuPathTracer::upload walks,
 * src/pathtrace.rs:971-1118), plus the table and atlas words the caller
 * packed (tables with OXPT_NONE in the emissive records' slot words). An
 * attribute array may be NULL when no triangle's flags select it. */
typedef struct oxpt_soup_desc {
    size_t n_tris;                /* 1 .. 2^24 */
    const float *pos;             /* 9 floats per triangle */
    const float *normals;         /* 9, for OXPT_F_NORMALS */
    const float *tangents;        /* 12, for OXPT_F_TANGENTS */
    const float *uv0;             /* 6, for OXPT_F_UV0 */
    const float *uv1;             /* 6, for OXPT_F_UV1 */
    const float *colors;          /* 12, for OXPT_F_COLORS */
    const uint32_t *material;     /* material index */
    const uint8_t *flags;         /* OXPT_F_* */
    const uint32_t *light_of;     /* emissive-light index or OXPT_NONE; NULL: none */
    uint32_t n_emissive;
    uint32_t elight_offset;       /* table word of emissive record 0 */
    const uint32_t *tables;
    size_t table_words;
    const uint16_t *atlas;        /* binary16 RGBA texels, may be NULL */
    uint32_t atlas_side;
    uint32_t atlas_layers;
    const oxpt_atlas_desc *atlas_build;  /* instead of atlas: build it on the device (or NULL) */
    /* Emissive lights found on the device (EmissiveLights::build,
     * src/pathtrace.rs:1201-1210): per material the luminance of its
     * emissive factor (0: does not emit). Every triangle with area x that
     * luminance > 0 (area in f64) is a light, indexed in input order; the
     * CDF over those weights and the records {cdf, pmf, triangle, slot} are
     * appended after the tables (light_of must be NULL, n_emissive 0). */
    const float *material_emission;
    uint32_t n_materials;
} oxpt_soup_desc;

/* What the device build produced: the Params words that depend on the
 * geometry (node count: env[3], triangle base: offs2[2], 4-wide root:
 * offs2[3]), the tree depths and timings. */
typedef struct oxpt_build_info {
    uint32_t node_count, tri_base, wide_base, bvh_depth, wide_depth;
    uint32_t pad;
    double upload_ms;  /* soup, tables and atlas host -> device */
    double build_ms;   /* BVH, triangle / attribute gather, 4-wide collapse */
    double atlas_ms;   /* device atlas (textures uploaded, mips, f16), 0 if none */
    uint32_t n_emissive;     /* emissive lights (found on the device or given) */
    uint32_t elight_offset;  /* table word of emissive record 0 */
} oxpt_build_info;

/* Builds the scene on the device: the binned-SAH BVH, the triangles and
 * attributes in leaf-slot order and the 4-wide nodes, the same words the
 * host packer writes (the tree depends only on the triangle set: leaves
 * list ascending triangle ids). Fills the emissive records' slot words of
 * the uploaded tables; emissive_slots (n_emissive words, may be NULL)
 * receives them for the caller's own table copy. Replaces the host BVH
 * and geometry packing of GpuPathTracer::upload (src/pathtrace.rs:970-1118);
 * the reference takes its tree from oxideav-mesh3d's Bvh on the CPU. */
int oxpt_scene_build(oxpt_ctx *ctx, const oxpt_soup_desc *desc, uint32_t *emissive_slots, oxpt_scene **out,
                     oxpt_build_info *info);
/* Copy a resident scene's words to host memory, blocking: which = 0 geom
 * (float4 words), 1 attrs (float4 words), 2 tables (u32), 3 atlas (u16);
 * *words (may be NULL) receives the element count, host_out may be NULL
 * to ask for it. No counterpart in the reference (tests and tools). */
int oxpt_scene_read(oxpt_ctx *ctx, const oxpt_scene *scene, int which, void *host_out, size_t bytes,
                    size_t *words);

/* Accumulator of w*h float4 (src/pathtrace.rs:551-561), zeroed. */
int oxpt_accum_create(oxpt_ctx *ctx, uint32_t w, uint32_t h, oxpt_accum **out);
void oxpt_accum_destroy(oxpt_accum *acc);
/* reset(): clear to zero (src/pathtrace.rs:495-502). */
int oxpt_accum_clear(oxpt_ctx *ctx, oxpt_accum *acc);
/* Copy the raw accumulator (w*h*16 bytes) to host, blocking. */
int oxpt_accum_read(oxpt_ctx *ctx, const oxpt_accum *acc, float *host_out, size_t host_bytes);
/* Overwrite the raw accumulator from host (w*h*16 bytes). */
int oxpt_accum_write(oxpt_ctx *ctx, oxpt_accum *acc, const float *host_in, size_t host_bytes);

/* Trace samples [first, first + count) of every pixel and add them to the
 * accumulator, in sample order per pixel; asynchronous. Replaces the
 * refine dispatch loop (src/pathtrace.rs:584-671) and trace_main
 * (src/shaders/pathtrace.wgsl:2047-2073). `P` is the 176-byte Params. */
int oxpt_refine(oxpt_ctx *ctx, const oxpt_scene *scene, oxpt_accum *acc, const oxpt_params *P,
                uint32_t first, uint32_t count);
/* Block until every queued launch finished (finish, src/pathtrace.rs:673-679). */
int oxpt_finish(oxpt_ctx *ctx);
/* Resolve into host memory, blocking: RGBA8 when rp->cfg[3] == 0 (w*h*4
 * bytes), RGBA32F when rp->cfg[3] == 1 (w*h*16 bytes). Replaces
 * encode_resolve + readback_texture (src/pathtrace.rs:769-844,
 * src/shaders/resolve.wgsl:151-188, src/gpu.rs:1060-1117). */
int oxpt_resolve(oxpt_ctx *ctx, const oxpt_accum *acc, const oxpt_resolve_params *rp,
                 void *host_out, size_t host_bytes);
/* Resolve into caller-owned device memory of the context's GPU, tightly
 * packed rows (RGBA8 or RGBA32F as oxpt_resolve), stream-ordered and
 * without blocking or readback: draw_texture, which encodes the resolve
 * pass into the caller's texture view (src/pathtrace.rs:687-691). A
 * viewer maps that memory for display (HIP / graphics interop); oxpt_finish
 * orders host reads after it. Host pointers are refused (OXPT_E_ARG). */
int oxpt_resolve_device(oxpt_ctx *ctx, const oxpt_accum *acc, const oxpt_resolve_params *rp,
                        void *device_out, size_t bytes);

/* Kernel timing: HIP events around every trace launch on the context's
 * stream. Returns launches and total milliseconds since the last reset. */
/* A display image the library owns (the reference's tracer-owned output
 * texture, src/pathtrace.rs:687-691): w x h RGBA8 (hdr = 0) or RGBA32F rows
 * in device memory. oxpt_image_resolve fills it from an accumulator like
 * oxpt_resolve_device (asynchronous; finish before another stream or
 * process reads it). oxpt_image_export writes a 64-byte HIP IPC handle
 * (hipIpcMemHandle_t) a viewer process opens with hipIpcOpenMemHandle to
 * read the same memory without a copy; images of 2 GiB or more are not
 * exported. */
int oxpt_image_create(oxpt_ctx *ctx, uint32_t w, uint32_t h, int hdr, oxpt_image **out);
void oxpt_image_destroy(oxpt_image *image);
void *oxpt_image_device_ptr(const oxpt_image *image);
size_t oxpt_image_bytes(const oxpt_image *image);
int oxpt_image_resolve(oxpt_ctx *ctx, const oxpt_accum *acc, const oxpt_resolve_params *rp, oxpt_image *image);
int oxpt_image_export(oxpt_ctx *ctx, const oxpt_image *image, void *handle_out, size_t handle_bytes);
int oxpt_ctx_timing(oxpt_ctx *ctx, int reset, uint64_t *launches, double *total_ms);

/* Batched refine (many scenes, cameras and sample runs of uneven size in one
 * persistent launch): every item adds samples [first, first + count) of
 * every pixel of its accumulator, in sample order per pixel, so each item's
 * accumulator equals what oxpt_refine would produce for it alone. Items must
 * belong to `ctx`; at most one item per accumulator. No counterpart in the
 * reference (its refine serves one scene at a time). */
typedef struct oxpt_batch_item {
    const oxpt_scene *scene;
    oxpt_accum *accum;
    oxpt_params params;
    uint32_t first, count;
} oxpt_batch_item;
int oxpt_refine_batch(oxpt_ctx *ctx, const oxpt_batch_item *items, size_t n);
/* oxpt_refine_batch that allocates nothing: no counterpart in the reference;
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc). Same
 * accumulators as oxpt_refine_batch. Its job table, work counters and
 * per-path sample buffer live in one fixed device array of 64 MiB + 68 KiB
 * that each GPU holds from library load to process end; it takes at most 256
 * items, and one sample of every item (8x8-tile padded pixels) at most
 * 4,194,304 paths (larger requests fail with OXPT_E_ARG; samples are split
 * into rounds that fit). Rounds with punctual lights run the general kernel
 * form (the NEE queue form would need memory). Every call on a GPU shares that
 * array: calls that overlap on one GPU (other contexts, other threads) race.
 * Blocks until its launches are done. */
int oxpt_refine_batch_noalloc(oxpt_ctx *ctx, const oxpt_batch_item *items, size_t n);
/* Replay the batched step from a HIP graph captured once per launch shape
 * (the job table and count are written to device memory before each
 * replay). Off by default. */
int oxpt_ctx_set_graph(oxpt_ctx *ctx, int enable);
uint64_t oxpt_ctx_graph_launches(const oxpt_ctx *ctx);

/* Launch shape and traversal knobs (autotuner). 0 keeps the default. */
typedef struct oxpt_tuning {
    uint32_t block;          /* threads per workgroup: 64, 128, 256 */
    uint32_t waves_per_cu;   /* persistent residency: workgroups = CUs * this / waves per WG */
    uint32_t max_paths;      /* paths per launch cap (samples are split, order kept); default 2^32 - 1 */
    uint32_t kernel;         /* 0 = default: oxpt_refine runs the persistent megakernel at 4 waves per SIMD,
                                6 for scenes without punctual lights (the early-continuation build; frames
                                with fewer pixels than resident lanes: the one-job batched kernel),
                                oxpt_refine_batch the batched (one path per work item) megakernel at 4,
                                6 when no job has punctual lights (rounds of fewer than 2^24 paths: 3, or
                                5 from 2^22 paths up when no job has punctual lights);
                                1 = plain one-pixel-per-thread, 2..6 = persistent built for 3, 4, 5, 6, 8
                                waves per SIMD, 7 = wavefront (separate shade / extend / occlude kernels),
                                8 = persistent built for 2 waves per SIMD */
    uint32_t wf_variant;     /* wavefront: bits 0-3 traversal form (0 = 6 waves / 8 stack entries in LDS,
                                1 = 5 waves / private stack, 2 = 8 / 8, 3 = 5 / 16, 4 = 6 / private),
                                bits 4-7 shade occupancy (0 = 3 waves, 1 = 2, 2 = 4) */
    uint32_t stack_lds;      /* megakernel traversal-stack entries per lane kept in LDS (the rest in private
                                memory): 0 = default for the form, 1..64, or OXPT_STACK_PRIVATE */
    uint32_t bvh;            /* 0 = walk the scene's 4-wide tree when it has one, 2 = the binary tree */
    uint32_t tail;           /* persistent megakernel: the last tail/16 of the frame's 8x8 tiles are traced
                                one path per claim (then added in sample order; at most 2^28 buffered
                                samples per launch): 0 = default (8, half the frame), 1..16, or
                                OXPT_TAIL_NONE (every pixel as one run of samples) */
    uint32_t early;          /* scenes with at most 31 punctual lights (forms 0 and 2..5): 0 = default,
                                the megakernel forms built to make the NEE shadow rays and the BSDF
                                continuation together at a hit (the BSDF does not live across
                                traversals; with punctual lights the pending contributions wait in a
                                per-lane queue in device memory, (3 * lights + 12) * 4 bytes per
                                resident lane); OXPT_EARLY_OFF = the general forms. Same images
                                either way */
    uint32_t opaque;         /* early forms: 0 = default, scenes whose materials all have alpha mode OPAQUE
                                (every job's, in a batch) run the traversal built without the alpha test
                                (no texture lookup in the walk loop); OXPT_OPAQUE_OFF = always the
                                general traversal. Same images either way */
} oxpt_tuning;
#define OXPT_EARLY_OFF 0xFFFFu
#define OXPT_OPAQUE_OFF 0xFFFFu
#define OXPT_STACK_PRIVATE 0xFFFFu
#define OXPT_TAIL_NONE 0xFFFFu
int oxpt_ctx_set_tuning(oxpt_ctx *ctx, const oxpt_tuning *t);
int oxpt_ctx_get_tuning(const oxpt_ctx *ctx, oxpt_tuning *t);
/* Autotuner (no counterpart in the reference): times the refine kernel
 * forms (occupancy x traversal-stack LDS share x tail share, and the wavefront engine)
 * on this scene and Params with spp_per_trial samples (1..64) into a
 * scratch accumulator, leaves the fastest tuning set on the context and
 * returns it (and its trial time). All forms give bit-identical images. */
int oxpt_autotune(oxpt_ctx *ctx, const oxpt_scene *scene, const oxpt_params *params, uint32_t spp_per_trial,
                  oxpt_tuning *best, double *best_ms);
/* The batched launch's autotuner: times the caller's items (at most
 * spp_per_trial samples each, into scratch accumulators) under the
 * megakernel occupancy forms, samples per dispatch (max_paths 2^32 - 1,
 * 2^28, 2^26: how many rounds the batch splits into) and traversal-stack LDS
 * shares, and keeps the fastest in ctx's tuning. Same images for every
 * candidate. No counterpart in the reference (its dispatch sizes,
 * PATHS_PER_DISPATCH / PATHS_PER_SUBMIT at src/pathtrace.rs:85-95, are
 * constants). */
int oxpt_autotune_batch(oxpt_ctx *ctx, const oxpt_batch_item *items, size_t n, uint32_t spp_per_trial,
                        oxpt_tuning *best, double *best_ms);

/* Debug / parity probes (no counterpart in the reference: they expose the
 * kernel's integer sampler and light-CDF picks for bit-exact checks).
 * Pattern words: for n tuples (x, y, index, p) the 4 scrambled Sobol' words
 * of pattern p (pathtrace.wgsl:202-224) before to_unit. */
int oxpt_debug_pattern_words(oxpt_ctx *ctx, const oxpt_scene *scene, uint32_t seed, size_t n,
                             const uint32_t *x, const uint32_t *y, const uint32_t *index,
                             const uint32_t *p, uint32_t *out4);
/* Light picks for n uniforms: {emissive index (pathtrace.wgsl:1466-1484),
 * env row, env column (pathtrace.wgsl:1633-1635)}. */
int oxpt_debug_light_picks(oxpt_ctx *ctx, const oxpt_scene *scene, const oxpt_params *P, size_t n,
                           const float *u, uint32_t *out3);
/* Closest hit of n rays (o, d float3 each): {slot, t bits}. */
int oxpt_debug_closest_hit(oxpt_ctx *ctx, const oxpt_scene *scene, const oxpt_params *P, size_t n,
                           const float *o3, const float *d3, uint32_t *out2);
/* The batched claim step's exact unsigned divisions (no counterpart in the
 * reference, which divides with WGSL's `/`): the host's multiply-shift pair
 * for divisor d (host only, no device), and the device quotient num[i] /
 * den[i] computed with it (den[i] != 0). */
void oxpt_udiv_magic(uint32_t d, uint32_t *m, uint32_t *l);
int oxpt_debug_udiv(oxpt_ctx *ctx, size_t n, const uint32_t *num, const uint32_t *den, uint32_t *out);
/* Texel indices of one texture coordinate u on a level of `size` texels
 * (1..65535) with wrap mode 0 clamp / 1 mirror / 2 repeat, as the kernels'
 * filter_level forms them (pathtrace.wgsl:335-380): {nearest, bilinear x0,
 * bilinear x0 + 1}, each wrapped. */
int oxpt_debug_texel_index(oxpt_ctx *ctx, size_t n, const float *u, const int32_t *size, const uint32_t *mode,
                           int32_t *out3);

/* =====================================================================
 * Layer 2: host tracer and scenes.
 * ===================================================================== */

typedef struct oxpt_host_scene oxpt_host_scene;
typedef struct oxpt_env_map oxpt_env_map;
typedef struct oxpt_tracer oxpt_tracer;

/* Texture binding of a material slot (KHR_texture_transform as a 2x3). */
typedef struct oxpt_tex_binding {
    int32_t texture;  /* index into the scene's textures, -1 = none */
    uint32_t uv_set;  /* 0 or 1 */
    float xform[6];   /* rows (a, b, c), (d, e, f): u' = a u + b v + c */
} oxpt_tex_binding;

/* glTF metallic-roughness material with the KHR extensions the estimator
 * reads (src/pathtrace.rs:1338-1423). */
typedef struct oxpt_material {
    float base_color[4];
    float emissive[3];
    uint32_t alpha_mode;   /* 0 opaque, 1 mask, 2 blend */
    float alpha_cutoff;
    uint32_t double_sided;
    uint32_t unlit;
    float metallic, roughness, normal_scale, ior;
    uint32_t has_specular; float specular, specular_color[3];
    uint32_t has_transmission; float transmission;
    uint32_t has_volume; float thickness, attenuation_color[3], attenuation_distance; /* <= 0: infinite */
    uint32_t has_clearcoat; float clearcoat, clearcoat_roughness, clearcoat_normal_scale;
    uint32_t has_sheen; float sheen_color[3], sheen_roughness;
    oxpt_tex_binding tex[OXPT_N_TEX_SLOTS]; /* slot order as OXPT_S_* */
} oxpt_material;

/* glTF sampler: wrap 0 clamp, 1 mirrored repeat, 2 repeat; mag 0 nearest,
 * 1 linear; min 0..5 = Nearest, Linear, NMN, LMN, NML, LML. */
typedef struct oxpt_sampler {
    uint32_t wrap_s, wrap_t, mag, min;
} oxpt_sampler;

/* Punctual light (KHR_lights_punctual). kind 0 directional, 1 point, 2 spot. */
typedef struct oxpt_light {
    uint32_t kind;
    float position[3], direction[3];
    float color[3], intensity;
    float range;  /* <= 0: none */
    float inner_cone, outer_cone;
} oxpt_light;

/* Camera: pinhole (yfov radians) or orthographic (ymag half-height). */
typedef struct oxpt_camera {
    float eye[3], target[3], up[3];
    float yfov;
    uint32_t orthographic;
    float ymag;
} oxpt_camera;

/* Render options: the fields of RenderOptions + PathTraceOptions the path
 * tracer reads (src/pathtrace.rs:885-926, 782-791). */
typedef struct oxpt_render_options {
    uint32_t width, height;
    uint32_t samples_per_pixel;
    uint32_t max_bounces;
    uint32_t rr_start;
    uint32_t strategy;   /* 0 MIS, 1 light only, 2 BSDF only */
    uint32_t seed;
    float clamp;         /* 0 = off */
    float ambient;
    uint8_t background[4];
    uint32_t tone_map;   /* 0 clamp, 1 Reinhard, 2 ACES fitted */
    float exposure;
} oxpt_render_options;

/* Defaults: 256x256, 64 spp, 8 bounces, roulette from 3, MIS, seed 0, no
 * clamp, ambient 0, background (16,16,20,255), clamp tone map, exposure 1. */
void oxpt_render_options_default(oxpt_render_options *o);

oxpt_host_scene *oxpt_host_scene_new(void);
void oxpt_host_scene_free(oxpt_host_scene *s);
int oxpt_host_scene_add_material(oxpt_host_scene *s, const oxpt_material *m);
/* Material with glTF defaults (white, metallic 1, roughness 1, ior 1.5). */
void oxpt_material_default(oxpt_material *m);
/* RGBA8 texture (level 0); mips are built at pack time. srgb_hint unused by
 * the estimator (each slot picks its colour space). Returns the index. */
int oxpt_host_scene_add_texture(oxpt_host_scene *s, uint32_t w, uint32_t h, const uint8_t *rgba8,
                                const oxpt_sampler *sampler);
/* A texture named by uri, its pixels supplied at prepare time by the
 * tracer's texture resolver (the reference's texture sources, decoded by its
 * TextureResolver; src/pathtrace.rs:474-478, 1261-1263). Returns its index,
 * as oxpt_host_scene_add_texture does. */
int oxpt_host_scene_add_texture_ref(oxpt_host_scene *s, const char *uri, const oxpt_sampler *sampler);
/* Append n triangles: positions 9 floats each; optional per-corner arrays
 * (normals 9, tangents 12, uv0 6, uv1 6, colors 12 floats per triangle;
 * NULL = absent); one material index per triangle. */
int oxpt_host_scene_add_triangles(oxpt_host_scene *s, size_t n, const float *pos, const float *normals,
                                  const float *tangents, const float *uv0, const float *uv1,
                                  const float *colors, const uint32_t *material);
int oxpt_host_scene_add_light(oxpt_host_scene *s, const oxpt_light *l);
int oxpt_host_scene_set_camera(oxpt_host_scene *s, const oxpt_camera *c);
size_t oxpt_host_scene_triangle_count(const oxpt_host_scene *s);

/* Equirectangular environment (EnvironmentMap): w*h RGB f32 texels. */
oxpt_env_map *oxpt_env_new(uint32_t w, uint32_t h, const float *rgb, float intensity);
void oxpt_env_free(oxpt_env_map *e);

/* Seeded scene generators. */
/* Closed room of 5 quads + 2 rotated boxes + emissive ceiling quad (Le 15),
 * Lambert colours U(0.2, 0.8), 40 degree pinhole camera. */
oxpt_host_scene *oxpt_testscene_room(uint32_t seed);
/* Double-sided emissive Lambert cube [-1,1]^3 seen from inside
 * (tests/pathtrace.rs:155-170): L = Le / (1 - rho). */
oxpt_host_scene *oxpt_testscene_furnace(float rho, float le);
/* fBm heightfield (gx x gz quads, 2 tris each) + n_spheres icospheres
 * (subdivision `subdiv`), glTF MR materials (metallic ~ Bernoulli(0.3),
 * roughness ~ U(0.05, 1)), n_emissive emissive triangles. */
oxpt_host_scene *oxpt_testscene_terrain(uint32_t seed, uint32_t gx, uint32_t gz, uint32_t n_spheres,
                                        uint32_t subdiv, uint32_t n_emissive);
/* Procedural sky: gradient + sun, w x h. */
oxpt_env_map *oxpt_env_sky(uint32_t w, uint32_t h, uint32_t seed);
/* Layered-material scene: instanced meshes baked to ~n_tris triangles,
 * transmission+volume / clearcoat / sheen / BLEND mixes, n_textures
 * value-noise RGBA textures of tex_size^2 with mips, n_lights punctual
 * lights. */
oxpt_host_scene *oxpt_testscene_layered(uint32_t seed, uint32_t n_tris, uint32_t n_textures,
                                        uint32_t tex_size, uint32_t n_lights);

/* Progressive tracer (GpuPathTracer). */
/* GpuPathTracer::new — opens device 0. */
int oxpt_tracer_new(oxpt_tracer **out);
/* Tracer on device `ordinal` (with_backend / from_device). */
int oxpt_tracer_new_on(int ordinal, oxpt_tracer **out);
void oxpt_tracer_free(oxpt_tracer *t);
const char *oxpt_tracer_last_error(const oxpt_tracer *t);
const char *oxpt_tracer_adapter_summary(const oxpt_tracer *t);
/* sync: *reset = 1 iff the accumulation was reset (src/pathtrace.rs:526-565). */
int oxpt_tracer_sync(oxpt_tracer *t, const oxpt_host_scene *scene, const oxpt_render_options *o,
                     int *reset);
int oxpt_tracer_refine(oxpt_tracer *t, uint32_t samples);
int oxpt_tracer_finish(oxpt_tracer *t);
int oxpt_tracer_image(oxpt_tracer *t, uint8_t *out_rgba8, size_t bytes);
int oxpt_tracer_hdr(oxpt_tracer *t, float *out_rgba32f, size_t bytes);
/* draw_texture (src/pathtrace.rs:687-691): resolve the current accumulation
 * with the current options into device memory (oxpt_resolve_device): RGBA8
 * (hdr = 0, w*h*4 bytes) or RGBA32F (hdr = 1); asynchronous. */
int oxpt_tracer_draw(oxpt_tracer *t, void *device_out, size_t bytes, int hdr);
/* draw_texture with the tracer's own output (src/pathtrace.rs:687-691, which
 * returns the tracer's texture): resolves into an image the tracer owns
 * (made or remade when the size or format changes) and returns it; valid
 * until the next draw_texture of another size or format, or oxpt_tracer_free.
 * Asynchronous: oxpt_tracer_finish before reading it elsewhere. */
int oxpt_tracer_draw_texture(oxpt_tracer *t, int hdr, oxpt_image **out);
/* The accumulator of the current accumulation (NULL before the first sync). */
const oxpt_accum *oxpt_tracer_accum(oxpt_tracer *t);
/* set_texture_resolver (src/pathtrace.rs:474-478, the Renderer trait's at
 * :1261-1263): the decoder of textures named by uri. It returns 0 and the
 * level-0 RGBA8 pixels (w x h x 4 bytes, read before it is called again) or
 * nonzero when it cannot decode the uri; such a texture, or any named texture
 * while no resolver is set, is absent (its material slots sample nothing, as
 * unbound slots). Installing or removing a resolver drops the decoded images
 * and re-prepares the scene on the next sync, which then resets. */
typedef int (*oxpt_texture_resolver)(void *user, const char *uri, uint32_t *w, uint32_t *h, const uint8_t **rgba8);
int oxpt_tracer_set_texture_resolver(oxpt_tracer *t, oxpt_texture_resolver fn, void *user);
void oxpt_tracer_reset(oxpt_tracer *t);
uint32_t oxpt_tracer_samples(const oxpt_tracer *t);
int oxpt_tracer_is_converged(const oxpt_tracer *t);
void oxpt_tracer_size(const oxpt_tracer *t, uint32_t *w, uint32_t *h);
void oxpt_tracer_invalidate_scene(oxpt_tracer *t);
/* set_environment: NULL removes it; resets the accumulation. The map is copied. */
int oxpt_tracer_set_environment(oxpt_tracer *t, const oxpt_env_map *env);
/* Renderer::render / render_hdr (src/pathtrace.rs:1241-1259). */
int oxpt_tracer_render(oxpt_tracer *t, const oxpt_host_scene *scene, const oxpt_render_options *o,
                       uint8_t *out_rgba8, size_t bytes);
int oxpt_tracer_render_hdr(oxpt_tracer *t, const oxpt_host_scene *scene,
                           const oxpt_render_options *o, float *out_rgba32f, size_t bytes);
/* The device context (for timing, tuning and the debug probes) and the
 * current device scene / params (valid after sync + one refine). */
oxpt_ctx *oxpt_tracer_ctx(oxpt_tracer *t);
const oxpt_scene *oxpt_tracer_device_scene(oxpt_tracer *t);
int oxpt_tracer_params(oxpt_tracer *t, oxpt_params *out);
/* Host copies of the packed words the kernels read (for the CPU oracle):
 * which = 0 geom (float4 count), 1 attrs (float4 count), 2 tables (words,
 * environment included), 3 atlas (u16 count). */
int oxpt_tracer_words(oxpt_tracer *t, int which, const void **ptr, size_t *count);
int oxpt_tracer_atlas_shape(oxpt_tracer *t, uint32_t *side, uint32_t *layers);
/* Prepare scenes on the device (oxpt_scene_build: BVH, triangle and
 * attribute words, 4-wide nodes) instead of on the host; the words and
 * images are the same. Takes effect at the next sync (the scene is
 * re-prepared). No counterpart in the reference, whose upload packs on
 * the CPU (src/pathtrace.rs:970-1239). */
int oxpt_tracer_set_device_build(oxpt_tracer *t, int enable);
/* Node counts and timings of the last device-prepared scene (zeros after
 * a host-packed one). */
int oxpt_tracer_build_info(const oxpt_tracer *t, oxpt_build_info *out);

/* Host-only packing (no device): the words and Params a sync + refine would
 * hand the kernels (upload + ensure_binding, src/pathtrace.rs:847-1239).
 * which: as oxpt_tracer_words. */
typedef struct oxpt_packed oxpt_packed;
int oxpt_pack(const oxpt_host_scene *scene, const oxpt_env_map *env, const oxpt_render_options *o,
              oxpt_packed **out);
void oxpt_packed_free(oxpt_packed *p);
int oxpt_packed_words(const oxpt_packed *p, int which, const void **ptr, size_t *count);
int oxpt_packed_params(const oxpt_packed *p, oxpt_params *out);
int oxpt_packed_info(const oxpt_packed *p, uint32_t *atlas_side, uint32_t *atlas_layers, uint32_t *bvh_depth);

/* 4-wide BVH nodes for geometry words a caller packed itself in the
 * reference's layout (binary nodes lo = (bmin, first), hi = (bmax, count),
 * node 0 the root, children the pair (first, first + 1) — the words of
 * src/pathtrace.rs:972-985): 32 words (8 float4) per wide node, to be
 * appended to geom at a float4 index that is a multiple of 8, with
 * Params.offs2[3] set to that index (0 keeps the binary walk). A node's
 * empty child slots hold count 0xFFFFFFFF and the box lo = +inf,
 * hi = -inf, which the walk's box test misses by itself. No
 * counterpart in the reference; for `-sys` callers that build their own
 * upload. Two-call: out = NULL returns the node count in *out_vec4.
 * Returns OXPT_E_INVALID when the tree exceeds the wide walk's limits
 * (leaves over 30 slots, 2^27 slots or nodes, depth over 32); the binary
 * walk then stays in use. */
int oxpt_build_wide_bvh(const float *node_words, uint32_t node_count, float *out, size_t out_vec4_cap,
                        size_t *out_vec4, uint32_t *wide_depth);

/* Serving: a continuous-batching loop over oxpt_refine_batch. Render requests
 * name a pool scene and options, refine requests add samples to a finished
 * render's resident accumulator. Requests are admitted against an HBM budget
 * that counts scenes, accumulators, readback staging and the launch scratch
 * (LRU eviction of idle scenes; a request that can never fit is rejected).
 * Scene uploads, accumulator creation and result readbacks run on a second
 * stream from an I/O thread, overlapping the running launch; a finished
 * request's resolve is queued right after the launch that finished it. Each
 * tick is one batched launch over every active request: with tick_ms > 0 it
 * is sized to about that duration from the measured rate, and the requests
 * closest to done get their remaining samples first (every request advances
 * at least one sample per tick); with tick_ms = 0 every request advances
 * spp_per_tick samples. Requests arriving meanwhile join the next tick.
 * Images equal the images rendered alone (per-pixel sample order is kept).
 * No counterpart in the reference. */
typedef struct oxpt_server oxpt_server;
typedef struct oxpt_server_config {
    uint64_t hbm_budget;    /* device bytes the server may hold (default 200 GiB) */
    uint32_t spp_per_tick;  /* tick_ms = 0: samples per request per tick (default 16);
                               tick_ms > 0: a cap per request per tick (0 = none) */
    uint32_t max_active;    /* request slots, 0 = unbounded */
    uint32_t keep_hdr;      /* also keep the RGBA32F result */
    uint32_t use_graph;     /* replay the batched step from a captured HIP graph */
    uint32_t device_build;  /* keep pool scenes as host scenes and prepare them on the device at first
                               use (oxpt_scene_build + device environment tables) instead of packing
                               them on the host in oxpt_server_add_scene; same images */
    uint32_t tick_ms;       /* target tick duration in ms (0: fixed spp_per_tick; default 0) */
    uint64_t tick_paths;    /* most paths (8x8-tile padded pixels x samples) one tick traces; bounds its
                               per-path sample buffer at 16 B each (default 2^27, at most 2^31) */
    uint32_t fail_tick;     /* testing: tick number n (1-based) fails without launching (0 = never) */
    uint32_t pad;
} oxpt_server_config;
typedef struct oxpt_request_info {
    uint32_t ok, width, height, samples;
    double latency_s;       /* arrival -> result ready */
    double queue_s;         /* arrival -> admission */
    uint64_t paths;         /* paths this request traced */
} oxpt_request_info;
typedef struct oxpt_server_stats {
    uint64_t submitted, completed, rejected, ticks, uploads, evictions, evicted_accums, paths;
    double busy_s;          /* wall time inside batched launches */
    uint64_t resident_bytes;  /* scenes + accumulators + staging + reserved launch scratch */
    uint32_t active, queued;
    uint64_t failed_ticks;
    double io_s;            /* I/O thread time in uploads and readbacks (overlaps launches) */
} oxpt_server_stats;
void oxpt_server_config_default(oxpt_server_config *c);
int oxpt_server_create(int device, const oxpt_server_config *cfg, oxpt_server **out);
void oxpt_server_destroy(oxpt_server *s);
oxpt_ctx *oxpt_server_ctx(oxpt_server *s);
const char *oxpt_server_last_error(const oxpt_server *s);
/* Pack a scene into the pool (host work, done before serving). */
int oxpt_server_add_scene(oxpt_server *s, const oxpt_host_scene *scene, const oxpt_env_map *env, uint32_t *id);
int oxpt_server_submit_render(oxpt_server *s, uint32_t scene_id, const oxpt_render_options *o, uint64_t *req);
int oxpt_server_submit_refine(oxpt_server *s, uint64_t render_req, uint32_t spp, uint64_t *req);
/* Block until the request is done (timeout_s < 0: forever); OXPT_E_BACKEND on
 * timeout or rejection. */
int oxpt_server_wait(oxpt_server *s, uint64_t req, double timeout_s, oxpt_request_info *info);
/* Copy a finished request's image: RGBA8 (hdr = 0) or RGBA32F (keep_hdr). */
int oxpt_server_result(oxpt_server *s, uint64_t req, int hdr, void *out, size_t bytes);
/* Drop a finished request's host image copies. */
int oxpt_server_release(oxpt_server *s, uint64_t req);
int oxpt_server_stats_get(oxpt_server *s, oxpt_server_stats *out);

/* Registry (register_into, src/lib.rs:117,329-338). */
#define OXPT_PATHTRACE_BACKEND_NAME "gpu-pathtrace"
/* Factory: a tracer for a registered name, OXPT_E_INVALID for unknown names. */
int oxpt_registry_make(const char *name, oxpt_tracer **out);

/* sizeof of the ABI structs, for binding checks: params, frame,
 * resolve_params, scene_desc, tex_binding, material, sampler, light, camera,
 * render_options, tuning (11 entries). */
void oxpt_abi_sizes(uint32_t out[11]);

/* Pack helper exposed for the f16 KAT (src/scene.rs:244-272). */
uint16_t oxpt_f32_to_f16(float v);
/* The packed Sobol' table and sheen LUT (product copies). */
void oxpt_sobol_directions(uint32_t out[OXPT_SOBOL_WORDS]);
void oxpt_sheen_table(float out[OXPT_SHEEN_WORDS]);

#ifdef __cplusplus
}
#endif

#endif /* OXPT_OXPT_H */
