/*
 * oxpt/layout.h — the word layouts every side of the path tracer agrees on.
 *
 * These are the device-side data formats of the reference path tracer
 * (oxideav-render-vulkan `src/pathtrace.rs` + `src/shaders/pathtrace.wgsl`),
 * kept word for word so that a scene packed once can be traced by the HIP
 * kernels and by the CPU oracle alike:
 *
 *   geom   float4[]  BVH nodes (2 float4 each) then triangles in leaf-slot
 *                    order (3 float4 each: p0+global id, p1+emissive light
 *                    index, p2+material index)      pathtrace.wgsl:54-57,638-644
 *   attrs  float4[]  12 float4 per leaf slot: 3 normals (w of the first =
 *                    flags), 3 tangents, 3 (uv0.xy, uv1.xy), 3 colours
 *                                                    pathtrace.rs:1275-1315
 *   tables uint32[]  Sobol' matrices, sheen LUT, texture descriptors,
 *                    materials, punctual lights, emissive CDF, environment
 *                                                    pathtrace.rs:1022-1210,1428-1461
 *   atlas  uint16[]  binary16 RGBA texels, square layers of side `atlas_side`,
 *                    layer-major then row-major     pathtrace.rs:1482-1627
 *   accum  float4[]  per pixel: sum of rgb over covered samples, covered count
 *                                                    pathtrace.wgsl:2064-2072
 *
 * Plain C so that C, C++ and HIP sources share it.
 */
#ifndef OXPT_LAYOUT_H
#define OXPT_LAYOUT_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Per-render constants: 11 x 16 B (pathtrace.rs:100-114, pathtrace.wgsl:17-41). */
typedef struct oxpt_params {
    float eye[4];      /* xyz eye, w = 1 for orthographic */
    float forward[4];  /* xyz forward, w = half_w */
    float side[4];     /* xyz side, w = half_h */
    float up[4];       /* xyz up, w = pixel cone spread (2 half_h / h) */
    uint32_t dims[4];  /* width, height, seed, max_bounces */
    uint32_t cfg[4];   /* rr_start, strategy (0 MIS, 1 light, 2 BSDF), n_punctual, n_emissive */
    uint32_t env[4];   /* env width, env height, has env, node count */
    uint32_t offs0[4]; /* table word offsets: sobol, sheen, materials, punctual lights */
    uint32_t offs1[4]; /* emissive lights, unused, env marginal, env conditional */
    uint32_t offs2[4]; /* env mass, env radiance, first triangle float4 in geom, first 4-wide node float4 (0: none) */
    float fparams[4];  /* ambient, clamp, P_env, unused */
} oxpt_params;

/* One dispatch's sample run (pathtrace.rs:116-123). */
typedef struct oxpt_frame {
    uint32_t first; /* first sample index */
    uint32_t count; /* consecutive samples per pixel */
    uint32_t y0;    /* first row */
    uint32_t rows;  /* rows in this dispatch */
} oxpt_frame;

/* Resolve constants (pathtrace.rs:125-131, resolve.wgsl:18-26). */
typedef struct oxpt_resolve_params {
    uint32_t cfg[4];      /* aa (1), class (3), tone map (0 clamp, 1 Reinhard, 2 ACES), hdr */
    float exposure[4];    /* exposure, samples taken, width, 0 */
    float background[4];  /* background sRGB bytes / 255 */
} oxpt_resolve_params;

#define OXPT_NONE 0xffffffffu

/* geom / attrs strides in float4 units. */
#define OXPT_NODE_VEC4 2u
#define OXPT_TRI_VEC4 3u
#define OXPT_ATTR_VEC4 12u

/* Attribute flags (w of the first normal of a slot; pathtrace.wgsl:75-80). */
#define OXPT_F_NORMALS 1u
#define OXPT_F_TANGENTS 2u
#define OXPT_F_COLORS 4u
#define OXPT_F_UV0 8u
#define OXPT_F_UV1 16u

/* Material record: 136 words (pathtrace.rs:1338-1423, pathtrace.wgsl:82-98). */
#define OXPT_MAT_STRIDE 136u
#define OXPT_M_BASE 0u        /* rgba */
#define OXPT_M_EMISSIVE 4u    /* rgb */
#define OXPT_M_FLAGS 7u       /* alpha mode bits 0-1, double-sided 4, unlit 8, emissive-lit 16 */
#define OXPT_M_METALLIC 8u
#define OXPT_M_ROUGHNESS 9u
#define OXPT_M_NORMAL_SCALE 10u
#define OXPT_M_CUTOFF 11u
#define OXPT_M_IOR 12u
#define OXPT_M_SPECULAR 13u
#define OXPT_M_SPEC_COLOR 14u /* rgb */
#define OXPT_M_TRANSMISSION 17u
#define OXPT_M_THICKNESS 18u
#define OXPT_M_ATTEN_COLOR 19u /* rgb */
#define OXPT_M_ATTEN_DIST 22u  /* < 0 = infinite */
#define OXPT_M_CLEARCOAT 23u
#define OXPT_M_CC_ROUGH 24u
#define OXPT_M_CC_NORMAL_SCALE 25u
#define OXPT_M_SHEEN_COLOR 26u /* rgb */
#define OXPT_M_SHEEN_ROUGH 29u
#define OXPT_M_SLOTS 32u       /* 13 texture slots x 8 words: desc, uv set, 2x3 transform rows */
#define OXPT_N_TEX_SLOTS 13u

#define OXPT_MF_DOUBLE_SIDED 4u
#define OXPT_MF_UNLIT 8u
#define OXPT_MF_EMISSIVE_LIT 16u

/* Texture slot order (pathtrace.wgsl:85-97). */
enum {
    OXPT_S_BASE = 0, OXPT_S_MR, OXPT_S_NORMAL, OXPT_S_EMISSIVE, OXPT_S_SPEC,
    OXPT_S_SPEC_COLOR, OXPT_S_TRANSMISSION, OXPT_S_THICKNESS, OXPT_S_CC,
    OXPT_S_CC_ROUGH, OXPT_S_CC_NORMAL, OXPT_S_SHEEN_COLOR, OXPT_S_SHEEN_ROUGH
};

/* Punctual light: 16 words (pathtrace.rs:1164-1200). */
#define OXPT_PLIGHT_STRIDE 16u
/* Emissive light record: cdf, pmf, global triangle id, leaf slot (pathtrace.rs:1201-1210). */
#define OXPT_ELIGHT_STRIDE 4u

/* Texture descriptor header (pathtrace.rs:1120-1132): levels, flags, image w,
 * image h, skipped levels, kept levels, 0, 0; then per kept level
 * {layer, x, y, w | h << 16}. */
#define OXPT_DESC_HEADER 8u

/* Table sizes. */
#define OXPT_SOBOL_WORDS 128u /* 4 dimensions x 32 direction words */
#define OXPT_SHEEN_COS 32u
#define OXPT_SHEEN_ROUGH 16u
#define OXPT_SHEEN_WORDS (OXPT_SHEEN_COS * OXPT_SHEEN_ROUGH)

/* Largest atlas layer side (pathtrace.rs:85-86). */
#define OXPT_MAX_LAYER 4096u

#ifdef __cplusplus
} /* extern "C" */
static_assert(sizeof(oxpt_params) == 176, "Params is 11 x vec4");
static_assert(sizeof(oxpt_frame) == 16, "Frame is one vec4");
static_assert(sizeof(oxpt_resolve_params) == 48, "ResolveParams is 3 x vec4");
static_assert(OXPT_M_SLOTS + 8 * OXPT_N_TEX_SLOTS <= OXPT_MAT_STRIDE, "material fits");
#endif

#endif /* OXPT_LAYOUT_H */
