/*
 * oxpt_oracle.h — CPU oracle of the path tracer. TEST INFRASTRUCTURE ONLY.
 *
 * A plain C++ restatement of the reference estimator
 * (oxideav-render-vulkan src/shaders/pathtrace.wgsl, whole file, and the
 * path-trace resolve src/shaders/resolve.wgsl:38-75,151-188), which in
 * turn mirrors the CPU `oxideav_render::PathTracer`. Only the tests, the
 * smoke check and the benchmark's CPU-baseline leg may load this library,
 * and only as the reference GPU results are compared with or as the timed
 * CPU baseline: the product path
 * (liboxpt) never calls it.
 *
 * The oracle traces the same packed words the HIP kernels read
 * (include/oxpt/layout.h); the packer is product code and is pinned by
 * the invariant tests instead.
 */
#ifndef OXPT_ORACLE_H
#define OXPT_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#include "oxpt/layout.h"

#ifdef __cplusplus
extern "C" {
#endif

/* A packed scene, as handed to the kernels. */
typedef struct oxo_scene {
    const float *geom;       /* float4 words */
    const float *attrs;      /* float4 words */
    const uint32_t *tables;
    const uint16_t *atlas;   /* may be NULL when no texture is bound */
    uint32_t atlas_side;
    uint32_t atlas_layers;
} oxo_scene;

/* Oracle options. */
enum {
    /* Evaluate the solid angle and Arvo sampling of emissive triangles in
     * f64, as the reference's CPU tracer does (README.md:117-119). Off =
     * f32, the reference kernel's arithmetic. */
    OXO_F64_LIGHTS = 1,
};

/* Algorithmic byte and event counters (no-cache model). */
typedef struct oxo_counts {
    uint64_t paths;       /* traced camera samples */
    uint64_t rays;        /* traversals */
    uint64_t nodes;       /* interior nodes visited (child-pair fetches) */
    uint64_t tris;        /* leaf triangles tested */
    uint64_t hits;        /* path vertices shaded */
    uint64_t tex_taps;
    uint64_t bytes;       /* total algorithmic bytes */
} oxo_counts;

/* Closed-form Sobol' direction numbers: 4 dimensions x 32 words. */
void oxo_sobol_directions(uint32_t out[OXPT_SOBOL_WORDS]);

/* Sheen albedo LUT by independent quadrature: 16 roughness rows x 32 cos. */
void oxo_sheen_table(float out[OXPT_SHEEN_WORDS], int samples_per_axis);

/* RNG pieces (pathtrace.wgsl:158-228). */
uint32_t oxo_pcg_hash(uint32_t x);
uint32_t oxo_hash_mix(uint32_t a, uint32_t b);
uint32_t oxo_nested_uniform_scramble(uint32_t x, uint32_t seed);
/* Scrambled integer words of pattern `p` of sample `index` of pixel (x, y):
 * the four values `to_unit` turns into the pattern's floats. */
void oxo_pattern_words(const uint32_t sobol[OXPT_SOBOL_WORDS], uint32_t seed, uint32_t x,
                       uint32_t y, uint32_t index, uint32_t p, uint32_t out[4]);
/* Batch form: n tuples. */
void oxo_pattern_words_batch(const uint32_t sobol[OXPT_SOBOL_WORDS], uint32_t seed, size_t n,
                             const uint32_t *x, const uint32_t *y, const uint32_t *index,
                             const uint32_t *p, uint32_t *out4);
/* Light-CDF indices: emissive pick and environment row/column picks for n
 * uniforms. out3 = {elight index, env row, env column}; the env pick uses u
 * for both axes. */
void oxo_light_picks(const oxo_scene *s, const oxpt_params *P, size_t n, const float *u,
                     uint32_t *out3);

/* Trace `count` samples from `first` for every pixel of rows [y0, y1) and
 * columns [x0, x1), adding into accum (float4 per pixel, width P->dims[0]).
 * threads <= 0 uses up to 16 hardware threads. counts may be NULL. */
int oxo_render(const oxo_scene *s, const oxpt_params *P, uint32_t first, uint32_t count,
               uint32_t x0, uint32_t y0, uint32_t x1, uint32_t y1, float *accum, int threads,
               int flags, oxo_counts *counts);

/* One sample of one pixel: returns covered, writes radiance[3]. */
int oxo_trace_sample(const oxo_scene *s, const oxpt_params *P, uint32_t x, uint32_t y,
                     uint32_t index, int flags, float radiance[3]);

/* Closest hit of a ray: BVH walk and brute force over every leaf slot.
 * Writes {slot, t} (slot = OXPT_NONE on a miss). */
void oxo_closest_hit(const oxo_scene *s, const oxpt_params *P, const float o[3],
                     const float d[3], int brute, uint32_t *slot, float *t);

/* Resolve (resolve.wgsl:151-188): out is RGBA8 (hdr = 0) or RGBA32F. */
void oxo_resolve(const float *accum, uint32_t w, uint32_t h, const oxpt_resolve_params *rp,
                 void *out);

/* IEEE binary32 -> binary16 (scene.rs:244-272). */
uint16_t oxo_f32_to_f16(float v);

/* BSDF probes for the invariant tests: build the layered BSDF of material
 * words `mat` (136 words, textures ignored) at normal n for view v, then
 * sample it with u (3 floats) and evaluate it at l. out: {ok, lx, ly, lz,
 * fx, fy, fz, pdf, eval_pdf_of_sampled_l, eval_f_x, eval_f_y, eval_f_z, eval_pdf}. */
void oxo_bsdf_probe(const uint32_t *mat, const float sheen_lut[OXPT_SHEEN_WORDS],
                    const float n[3], const float v[3], int front, const float u[3],
                    const float l[3], float out[13]);

/* Texel indices of one coordinate u on a level of `size` texels, wrap mode
 * 0 clamp / 1 mirror / 2 repeat: {nearest, bilinear x0, x0 + 1}, wrapped. */
void oxo_texel_index(float u, int32_t size, uint32_t mode, int32_t out[3]);
/* Microfacet distributions: kind 0 GGX D(n.h; a) (wgsl:922-930), kind 1
 * Charlie D(n.h; a) (wgsl:984-991). */
float oxo_ndf(int kind, float n_dot_h, float a);

/* Spherical-triangle probes: solid angle and an Arvo sample (f32 or f64). */
float oxo_solid_angle(const float p[3], const float a[3], const float b[3], const float c[3],
                      int f64);
void oxo_sample_spherical_triangle(const float p[3], const float a[3], const float b[3],
                                   const float c[3], float u0, float u1, int f64, float out[3]);

#ifdef __cplusplus
}
#endif

#endif
