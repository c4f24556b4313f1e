// GpuPathTracer: the progressive host driver over the device boundary,
// with the reference's API and reset rules (oxideav-render-vulkan
// src/pathtrace.rs:309-725, 847-966, 1241-1264; src/lib.rs:117,329-338).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <optional>
#include <string>

#include "scene.hpp"

namespace {

// Largest accepted output side (the reference's max_texture_dimension_2d
// check, src/pathtrace.rs:534-541).
constexpr uint32_t kMaxDim = 32768;

// Options whose change resets the accumulation (radiance_key): size,
// seed, bounces, roulette start, strategy, clamp, ambient. Background,
// tone map, exposure and the sample target are display-only
// (tests/pathtrace.rs:451-496).
bool same_radiance(const oxpt_render_options &a, const oxpt_render_options &b) {
    auto fb = [](float f) {
        uint32_t u;
        std::memcpy(&u, &f, 4);
        return u;
    };
    return a.width == b.width && a.height == b.height && a.seed == b.seed && a.max_bounces == b.max_bounces &&
           a.rr_start == b.rr_start && a.strategy == b.strategy && fb(a.clamp) == fb(b.clamp) &&
           fb(a.ambient) == fb(b.ambient);
}

}  // namespace

struct oxpt_tracer {
    oxpt_ctx *ctx = nullptr;
    std::string err;
    oxpt_render_options opts{};
    bool dirty = true;
    std::unique_ptr<oxpt::Packed> packed;
    oxpt_scene *dscene = nullptr;
    std::optional<oxpt::EnvMap> env;
    oxpt::CameraFrame cam{};
    oxpt::CameraSource cam_src;
    uint32_t w = 0, h = 0;
    oxpt_accum *accum = nullptr;
    uint32_t samples = 0;
    bool binding = false;
    oxpt_params params{};
    std::vector<uint32_t> tables;  // scene tables + environment section

    int fail(int code, const std::string &m) {
        err = m;
        return code;
    }
    int dev_fail(int code) {
        err = oxpt_last_error(ctx);
        return code;
    }
    void reset() {
        samples = 0;
        if (accum) oxpt_accum_clear(ctx, accum);
    }
    // Device scene preparation (oxpt_scene_build) instead of host BVH and
    // geometry packing; the words are the same either way.
    bool device_build = false;
    bool words_on_device = false;  // geom / attrs of `packed` not downloaded yet
    oxpt_build_info build_info{};

    int upload_device(const oxpt::HostScene &s) {
        // Host: materials, lights, texture placement, Sobol' / sheen tables.
        // Device: lights, BVH, triangle / attribute words, 4-wide nodes, atlas.
        if (dscene) {
            oxpt_scene_destroy(dscene);
            dscene = nullptr;
        }
        auto pk = std::make_unique<oxpt::Packed>();
        std::vector<oxpt::Texture> resolved;
        const std::vector<oxpt::Texture> *tl = resolve_textures(s, resolved);
        if (int rc = oxpt::device_prepare(ctx, s, *pk, &dscene, &build_info, tl)) return dev_fail(rc);
        packed = std::move(pk);
        words_on_device = true;
        cam_src = oxpt::camera_source(s);
        return OXPT_OK;
    }
    // Host copies of the device-built geometry words, on request.
    int fetch_words() {
        if (!words_on_device) return OXPT_OK;
        size_t gv = 0, av = 0;
        int rc = oxpt_scene_read(ctx, dscene, 0, nullptr, 0, &gv);
        if (!rc) rc = oxpt_scene_read(ctx, dscene, 1, nullptr, 0, &av);
        if (rc) return dev_fail(rc);
        packed->geom.resize(gv * 4);
        packed->attrs.resize(av * 4);
        rc = oxpt_scene_read(ctx, dscene, 0, packed->geom.data(), gv * 16, nullptr);
        if (!rc) rc = oxpt_scene_read(ctx, dscene, 1, packed->attrs.data(), av * 16, nullptr);
        if (rc) return dev_fail(rc);
        size_t xv = 0;
        rc = oxpt_scene_read(ctx, dscene, 3, nullptr, 0, &xv);
        if (!rc) {
            packed->atlas.resize(xv);
            rc = oxpt_scene_read(ctx, dscene, 3, packed->atlas.data(), xv * 2, nullptr);
        }
        if (rc) return dev_fail(rc);
        words_on_device = false;
        return OXPT_OK;
    }
    int upload(const oxpt::HostScene &s) {
        build_info = oxpt_build_info{};
        if (device_build && s.soup.count() > 0) return upload_device(s);
        words_on_device = false;
        std::vector<oxpt::Texture> resolved;
        const std::vector<oxpt::Texture> *tl = resolve_textures(s, resolved);
        auto pk = std::make_unique<oxpt::Packed>(oxpt::pack_scene(s, true, nullptr, true, true, tl));
        if (dscene) {
            oxpt_scene_destroy(dscene);
            dscene = nullptr;
        }
        oxpt_scene_desc d{};
        d.geom = pk->geom.data();
        d.geom_vec4 = pk->geom.size() / 4;
        d.attrs = pk->attrs.data();
        d.attrs_vec4 = pk->attrs.size() / 4;
        d.tables = pk->tables.data();
        d.table_words = pk->tables.size();
        d.atlas = pk->atlas.empty() ? nullptr : pk->atlas.data();
        d.atlas_side = pk->atlas_side;
        d.atlas_layers = pk->atlas_layers;
        int rc = oxpt_scene_upload(ctx, &d, &dscene);
        if (rc) return dev_fail(rc);
        packed = std::move(pk);
        cam_src = oxpt::camera_source(s);
        return OXPT_OK;
    }
    // Tables (scene + environment) and Params (ensure_binding).
    int ensure_binding() {
        if (binding) return OXPT_OK;
        const bool env_dev = device_build && env && words_on_device_tables_ok();
        oxpt::bind(*packed, env ? &*env : nullptr, cam, opts, w, h, tables, params, !env_dev);
        int rc;
        if (env_dev) {
            // Environment sampling tables built on the device after the scene tables.
            rc = oxpt::upload_tables(ctx, dscene, tables, &*env);
            tables_on_device = true;
        } else {
            rc = oxpt_scene_set_tables(ctx, dscene, tables.data(), tables.size());
            tables_on_device = false;
        }
        if (rc) return dev_fail(rc);
        binding = true;
        return OXPT_OK;
    }
    // The device environment builder needs a full w x h x 3 texel array.
    bool words_on_device_tables_ok() const {
        return env && env->w > 0 && env->h > 0 && env->rgb.size() == (size_t)env->w * env->h * 3;
    }
    bool tables_on_device = false;  // host `tables` lacks the environment words
    int fetch_tables() {
        if (!tables_on_device) return OXPT_OK;
        size_t tw = 0;
        int rc = oxpt_scene_read(ctx, dscene, 2, nullptr, 0, &tw);
        if (!rc) {
            tables.resize(tw);
            rc = oxpt_scene_read(ctx, dscene, 2, tables.data(), tw * 4, nullptr);
        }
        if (rc) return dev_fail(rc);
        tables_on_device = false;
        return OXPT_OK;
    }
    int sync(const oxpt_host_scene *scene, const oxpt_render_options &o, int *reset_out) {
        if (reset_out) *reset_out = 0;
        bool need_prepare = dirty || !dscene;
        bool changed = need_prepare || !same_radiance(o, opts);
        opts = o;
        if (!changed) return OXPT_OK;
        uint32_t width = std::max(o.width, 1u), height = std::max(o.height, 1u);
        if (width > kMaxDim || height > kMaxDim)
            return fail(OXPT_E_INVALID, std::to_string(width) + "x" + std::to_string(height) + " exceeds the " +
                                            std::to_string(kMaxDim) + " max dimension");
        if (need_prepare) {
            if (!scene) return fail(OXPT_E_ARG, "sync: no scene");
            if (int rc = upload(scene->s)) return rc;
            dirty = false;
        }
        cam = oxpt::resolve_camera(cam_src, width, height);
        if (w != width || h != height || !accum) {
            if (accum) oxpt_accum_destroy(accum);
            accum = nullptr;
            w = width;
            h = height;
            if (int rc = oxpt_accum_create(ctx, width, height, &accum)) return dev_fail(rc);
        }
        binding = false;
        reset();
        if (reset_out) *reset_out = 1;
        return OXPT_OK;
    }
    int refine(uint32_t n) {
        if (!dscene || n == 0) return OXPT_OK;
        n = std::min(n, 0xffffffffu - samples);
        if (n == 0) return OXPT_OK;
        if (int rc = ensure_binding()) return rc;
        if (int rc = oxpt_refine(ctx, dscene, accum, &params, samples, n)) return dev_fail(rc);
        samples += n;
        return OXPT_OK;
    }
    int resolve(bool hdr, void *out, size_t bytes, bool device = false) {
        if (!accum) {
            // Never synced: an all-background frame of the current size.
            uint32_t ww = std::max(w, 1u), hh = std::max(h, 1u);
            if (int rc = oxpt_accum_create(ctx, ww, hh, &accum)) return dev_fail(rc);
            w = ww;
            h = hh;
        }
        oxpt_resolve_params rp{};
        rp.cfg[0] = 1;
        rp.cfg[1] = 3;
        rp.cfg[2] = opts.tone_map <= 2 ? opts.tone_map : 0u;
        rp.cfg[3] = hdr ? 1u : 0u;
        rp.exposure[0] = std::max(opts.exposure, 0.0f);
        rp.exposure[1] = (float)samples;
        rp.exposure[2] = (float)w;
        for (int k = 0; k < 4; k++) rp.background[k] = (float)opts.background[k] / 255.0f;
        int rc = device ? oxpt_resolve_device(ctx, accum, &rp, out, bytes) : oxpt_resolve(ctx, accum, &rp, out, bytes);
        return rc ? dev_fail(rc) : OXPT_OK;
    }
    // set_texture_resolver: the resolver of textures named by uri, and the
    // textures it decoded (kept until the resolver changes, as the
    // reference's TextureCache keeps its decoded images).
    oxpt_texture_resolver resolver = nullptr;
    void *resolver_user = nullptr;
    std::map<std::string, oxpt::Texture> tex_cache;
    // The scene's texture list with every named texture resolved (absent,
    // 0 x 0, when the resolver gives none); nullptr when no texture is named.
    const std::vector<oxpt::Texture> *resolve_textures(const oxpt::HostScene &s, std::vector<oxpt::Texture> &out) {
        bool named = false;
        for (const auto &tx : s.textures) named = named || !tx.uri.empty();
        if (!named) return nullptr;
        out.clear();
        out.reserve(s.textures.size());
        for (const auto &tx : s.textures) {
            if (tx.uri.empty()) {
                out.push_back(tx);
                continue;
            }
            auto it = tex_cache.find(tx.uri);
            if (it == tex_cache.end()) {
                oxpt::Texture dec;
                uint32_t ww = 0, hh = 0;
                const uint8_t *px = nullptr;
                if (resolver && resolver(resolver_user, tx.uri.c_str(), &ww, &hh, &px) == 0 && px && ww > 0 && hh > 0 &&
                    ww <= 65535 && hh <= 65535) {
                    dec.w = ww;
                    dec.h = hh;
                    dec.rgba8.assign(px, px + (size_t)ww * hh * 4);
                }
                it = tex_cache.emplace(tx.uri, std::move(dec)).first;
            }
            oxpt::Texture r = it->second;
            r.sampler = tx.sampler;
            r.uri = tx.uri;
            out.push_back(std::move(r));
        }
        return &out;
    }
    oxpt_image *texture = nullptr;  // draw_texture output, tex_w x tex_h, tex_hdr
    uint32_t tex_w = 0, tex_h = 0;
    int tex_hdr = 0;
    ~oxpt_tracer() {
        if (texture) oxpt_image_destroy(texture);
        if (accum) oxpt_accum_destroy(accum);
        if (dscene) oxpt_scene_destroy(dscene);
        if (ctx) oxpt_ctx_destroy(ctx);
    }
};

extern "C" {

void oxpt_render_options_default(oxpt_render_options *o) {
    std::memset(o, 0, sizeof *o);
    o->width = 256;
    o->height = 256;
    o->samples_per_pixel = 64;
    o->max_bounces = 8;
    o->rr_start = 3;
    o->strategy = 0;
    o->seed = 0;
    o->clamp = 0.0f;
    o->ambient = 0.0f;
    o->background[0] = 16;
    o->background[1] = 16;
    o->background[2] = 20;
    o->background[3] = 255;
    o->tone_map = 0;
    o->exposure = 1.0f;
}

void oxpt_material_default(oxpt_material *m) {
    std::memset(m, 0, sizeof *m);
    for (int k = 0; k < 4; k++) m->base_color[k] = 1.0f;
    m->alpha_cutoff = 0.5f;
    m->metallic = 1.0f;
    m->roughness = 1.0f;
    m->normal_scale = 1.0f;
    m->ior = 1.5f;
    m->specular = 1.0f;
    for (int k = 0; k < 3; k++) {
        m->specular_color[k] = 1.0f;
        m->attenuation_color[k] = 1.0f;
    }
    m->clearcoat_normal_scale = 1.0f;
    for (uint32_t s = 0; s < OXPT_N_TEX_SLOTS; s++) {
        m->tex[s].texture = -1;
        m->tex[s].uv_set = 0;
        m->tex[s].xform[0] = 1.0f;
        m->tex[s].xform[4] = 1.0f;
    }
}

oxpt_host_scene *oxpt_host_scene_new(void) { return new oxpt_host_scene(); }
void oxpt_host_scene_free(oxpt_host_scene *s) { delete s; }

int oxpt_host_scene_add_material(oxpt_host_scene *s, const oxpt_material *m) {
    if (!s || !m) return OXPT_E_ARG;
    s->s.materials.push_back(*m);
    s->s.generation++;
    return (int)s->s.materials.size() - 1;
}

int oxpt_host_scene_add_texture(oxpt_host_scene *s, uint32_t w, uint32_t h, const uint8_t *rgba8,
                                const oxpt_sampler *sampler) {
    if (!s || !rgba8 || w == 0 || h == 0 || w > 65535 || h > 65535) return OXPT_E_ARG;
    oxpt::Texture t;
    t.w = w;
    t.h = h;
    t.rgba8.assign(rgba8, rgba8 + (size_t)w * h * 4);
    if (sampler) t.sampler = *sampler;
    s->s.textures.push_back(std::move(t));
    s->s.generation++;
    return (int)s->s.textures.size() - 1;
}

int oxpt_host_scene_add_texture_ref(oxpt_host_scene *s, const char *uri, const oxpt_sampler *sampler) {
    if (!s || !uri || !*uri) return OXPT_E_ARG;
    oxpt::Texture t;
    t.uri = uri;
    if (sampler) t.sampler = *sampler;
    s->s.textures.push_back(std::move(t));
    s->s.generation++;
    return (int)s->s.textures.size() - 1;
}

int oxpt_host_scene_add_triangles(oxpt_host_scene *hs, size_t n, const float *pos, const float *normals,
                                  const float *tangents, const float *uv0, const float *uv1, const float *colors,
                                  const uint32_t *material) {
    if (!hs || !pos || !material) return OXPT_E_ARG;
    oxpt::Soup &s = hs->s.soup;
    for (size_t i = 0; i < n; i++)
        if (material[i] >= hs->s.materials.size()) return OXPT_E_ARG;
    s.pos.insert(s.pos.end(), pos, pos + 9 * n);
    auto opt = [&](std::vector<float> &dst, const float *src, size_t per, float fill) {
        if (src) dst.insert(dst.end(), src, src + per * n);
        else dst.insert(dst.end(), per * n, fill);
    };
    opt(s.normals, normals, 9, 0.0f);
    opt(s.tangents, tangents, 12, 0.0f);
    opt(s.uv0, uv0, 6, 0.0f);
    opt(s.uv1, uv1, 6, 0.0f);
    opt(s.colors, colors, 12, 1.0f);
    uint8_t fl = (normals ? OXPT_F_NORMALS : 0) | (tangents ? OXPT_F_TANGENTS : 0) | (colors ? OXPT_F_COLORS : 0) |
                 (uv0 ? OXPT_F_UV0 : 0) | (uv1 ? OXPT_F_UV1 : 0);
    s.material.insert(s.material.end(), material, material + n);
    s.flags.insert(s.flags.end(), n, fl);
    hs->s.generation++;
    return OXPT_OK;
}

int oxpt_host_scene_add_light(oxpt_host_scene *s, const oxpt_light *l) {
    if (!s || !l) return OXPT_E_ARG;
    s->s.lights.push_back(*l);
    s->s.generation++;
    return OXPT_OK;
}

int oxpt_host_scene_set_camera(oxpt_host_scene *s, const oxpt_camera *c) {
    if (!s || !c) return OXPT_E_ARG;
    s->s.camera = *c;
    s->s.has_camera = true;
    s->s.generation++;
    return OXPT_OK;
}

size_t oxpt_host_scene_triangle_count(const oxpt_host_scene *s) { return s ? s->s.soup.count() : 0; }

oxpt_env_map *oxpt_env_new(uint32_t w, uint32_t h, const float *rgb, float intensity) {
    if (!rgb || w == 0 || h == 0) return nullptr;
    auto *e = new oxpt_env_map();
    e->e.w = w;
    e->e.h = h;
    e->e.rgb.assign(rgb, rgb + (size_t)w * h * 3);
    e->e.intensity = intensity;
    return e;
}
void oxpt_env_free(oxpt_env_map *e) { delete e; }

int oxpt_tracer_new_on(int ordinal, oxpt_tracer **out) {
    if (!out) return OXPT_E_ARG;
    *out = nullptr;
    oxpt_ctx *c = nullptr;
    int rc = oxpt_ctx_create(ordinal, &c);
    if (rc) return rc;
    auto *t = new oxpt_tracer();
    t->ctx = c;
    oxpt_render_options_default(&t->opts);
    *out = t;
    return OXPT_OK;
}
int oxpt_tracer_new(oxpt_tracer **out) { return oxpt_tracer_new_on(0, out); }
void oxpt_tracer_free(oxpt_tracer *t) { delete t; }
const char *oxpt_tracer_last_error(const oxpt_tracer *t) { return t ? t->err.c_str() : "null tracer"; }
const char *oxpt_tracer_adapter_summary(const oxpt_tracer *t) { return t ? oxpt_ctx_summary(t->ctx) : ""; }

int oxpt_tracer_sync(oxpt_tracer *t, const oxpt_host_scene *scene, const oxpt_render_options *o, int *reset) {
    if (!t || !o) return OXPT_E_ARG;
    return t->sync(scene, *o, reset);
}
int oxpt_tracer_refine(oxpt_tracer *t, uint32_t n) { return t ? t->refine(n) : OXPT_E_ARG; }
int oxpt_tracer_finish(oxpt_tracer *t) {
    if (!t) return OXPT_E_ARG;
    int rc = oxpt_finish(t->ctx);
    return rc ? t->dev_fail(rc) : OXPT_OK;
}
int oxpt_tracer_image(oxpt_tracer *t, uint8_t *out, size_t bytes) { return t ? t->resolve(false, out, bytes) : OXPT_E_ARG; }
int oxpt_tracer_hdr(oxpt_tracer *t, float *out, size_t bytes) { return t ? t->resolve(true, out, bytes) : OXPT_E_ARG; }
int oxpt_tracer_draw(oxpt_tracer *t, void *device_out, size_t bytes, int hdr) {
    return t ? t->resolve(hdr != 0, device_out, bytes, true) : OXPT_E_ARG;
}
int oxpt_tracer_draw_texture(oxpt_tracer *t, int hdr, oxpt_image **out) {
    if (!t || !out) return OXPT_E_ARG;
    *out = nullptr;
    // An accumulator exists after the first resolve (resolve makes a
    // background frame when never synced); the image follows its size.
    uint32_t ww = std::max(t->w, 1u), hh = std::max(t->h, 1u);
    const size_t want = (size_t)ww * hh * (hdr ? 16u : 4u);
    if (t->texture && (t->tex_w != ww || t->tex_h != hh || t->tex_hdr != (hdr ? 1 : 0))) {
        oxpt_finish(t->ctx);
        oxpt_image_destroy(t->texture);
        t->texture = nullptr;
    }
    if (!t->texture) {
        if (int rc = oxpt_image_create(t->ctx, ww, hh, hdr, &t->texture)) return t->dev_fail(rc);
        t->tex_w = ww;
        t->tex_h = hh;
        t->tex_hdr = hdr ? 1 : 0;
    }
    if (int rc = t->resolve(hdr != 0, oxpt_image_device_ptr(t->texture), want, true)) return rc;
    *out = t->texture;
    return OXPT_OK;
}
const oxpt_accum *oxpt_tracer_accum(oxpt_tracer *t) { return t ? t->accum : nullptr; }
int oxpt_tracer_set_texture_resolver(oxpt_tracer *t, oxpt_texture_resolver fn, void *user) {
    if (!t) return OXPT_E_ARG;
    t->resolver = fn;
    t->resolver_user = user;
    t->tex_cache.clear();
    t->dirty = true;  // re-prepare on the next sync (src/pathtrace.rs:474-478)
    return OXPT_OK;
}
void oxpt_tracer_reset(oxpt_tracer *t) {
    if (t) t->reset();
}
uint32_t oxpt_tracer_samples(const oxpt_tracer *t) { return t ? t->samples : 0; }
int oxpt_tracer_is_converged(const oxpt_tracer *t) { return t && t->samples >= t->opts.samples_per_pixel ? 1 : 0; }
void oxpt_tracer_size(const oxpt_tracer *t, uint32_t *w, uint32_t *h) {
    if (w) *w = t ? t->w : 0;
    if (h) *h = t ? t->h : 0;
}
void oxpt_tracer_invalidate_scene(oxpt_tracer *t) {
    if (t) t->dirty = true;
}
int oxpt_tracer_set_environment(oxpt_tracer *t, const oxpt_env_map *e) {
    if (!t) return OXPT_E_ARG;
    if (e) t->env = e->e;
    else t->env.reset();
    t->binding = false;
    t->reset();
    return OXPT_OK;
}
int oxpt_tracer_render(oxpt_tracer *t, const oxpt_host_scene *scene, const oxpt_render_options *o, uint8_t *out,
                       size_t bytes) {
    if (!t || !o) return OXPT_E_ARG;
    t->dirty = true;
    if (int rc = t->sync(scene, *o, nullptr)) return rc;
    if (int rc = t->refine(std::max(o->samples_per_pixel, 1u))) return rc;
    return t->resolve(false, out, bytes);
}
int oxpt_tracer_render_hdr(oxpt_tracer *t, const oxpt_host_scene *scene, const oxpt_render_options *o, float *out,
                           size_t bytes) {
    if (!t || !o) return OXPT_E_ARG;
    t->dirty = true;
    if (int rc = t->sync(scene, *o, nullptr)) return rc;
    if (int rc = t->refine(std::max(o->samples_per_pixel, 1u))) return rc;
    return t->resolve(true, out, bytes);
}
oxpt_ctx *oxpt_tracer_ctx(oxpt_tracer *t) { return t ? t->ctx : nullptr; }
const oxpt_scene *oxpt_tracer_device_scene(oxpt_tracer *t) {
    if (!t || !t->dscene) return nullptr;
    if (t->ensure_binding()) return nullptr;
    return t->dscene;
}
int oxpt_tracer_params(oxpt_tracer *t, oxpt_params *out) {
    if (!t || !out) return OXPT_E_ARG;
    if (!t->dscene) return t->fail(OXPT_E_ARG, "params: not synced");
    if (int rc = t->ensure_binding()) return rc;
    *out = t->params;
    return OXPT_OK;
}
int oxpt_tracer_words(oxpt_tracer *t, int which, const void **ptr, size_t *count) {
    if (!t || !ptr || !count) return OXPT_E_ARG;
    if (!t->packed) return t->fail(OXPT_E_ARG, "words: not synced");
    if (int rc = t->ensure_binding()) return rc;
    if (which <= 1 || which == 3)
        if (int rc = t->fetch_words()) return rc;
    if (which == 2)
        if (int rc = t->fetch_tables()) return rc;
    switch (which) {
        case 0: *ptr = t->packed->geom.data(); *count = t->packed->geom.size() / 4; return OXPT_OK;
        case 1: *ptr = t->packed->attrs.data(); *count = t->packed->attrs.size() / 4; return OXPT_OK;
        case 2: *ptr = t->tables.data(); *count = t->tables.size(); return OXPT_OK;
        case 3: *ptr = t->packed->atlas.data(); *count = t->packed->atlas.size(); return OXPT_OK;
        default: return OXPT_E_ARG;
    }
}
int oxpt_tracer_set_device_build(oxpt_tracer *t, int enable) {
    if (!t) return OXPT_E_ARG;
    if (t->device_build != (enable != 0)) t->dirty = true;  // re-prepare on the next sync
    t->device_build = enable != 0;
    return OXPT_OK;
}
int oxpt_tracer_build_info(const oxpt_tracer *t, oxpt_build_info *out) {
    if (!t || !out) return OXPT_E_ARG;
    *out = t->build_info;
    return OXPT_OK;
}
int oxpt_tracer_atlas_shape(oxpt_tracer *t, uint32_t *side, uint32_t *layers) {
    if (!t || !t->packed) return OXPT_E_ARG;
    if (side) *side = t->packed->atlas_side;
    if (layers) *layers = t->packed->atlas_layers;
    return OXPT_OK;
}

int oxpt_registry_make(const char *name, oxpt_tracer **out) {
    if (!name || !out) return OXPT_E_ARG;
    *out = nullptr;
    if (std::strcmp(name, OXPT_PATHTRACE_BACKEND_NAME) != 0) return OXPT_E_INVALID;
    return oxpt_tracer_new(out);
}

int oxpt_pack(const oxpt_host_scene *scene, const oxpt_env_map *env, const oxpt_render_options *o,
              oxpt_packed **out) {
    if (!scene || !o || !out) return OXPT_E_ARG;
    *out = nullptr;
    uint32_t w = std::max(o->width, 1u), h = std::max(o->height, 1u);
    if (w > kMaxDim || h > kMaxDim) return OXPT_E_INVALID;
    auto *p = new oxpt_packed();
    p->pk = oxpt::pack_scene(scene->s);
    oxpt::CameraFrame cam = oxpt::resolve_camera(oxpt::camera_source(scene->s), w, h);
    oxpt::bind(p->pk, env ? &env->e : nullptr, cam, *o, w, h, p->tables, p->P);
    *out = p;
    return OXPT_OK;
}
void oxpt_packed_free(oxpt_packed *p) { delete p; }
int oxpt_packed_words(const oxpt_packed *p, int which, const void **ptr, size_t *count) {
    if (!p || !ptr || !count) return OXPT_E_ARG;
    switch (which) {
        case 0: *ptr = p->pk.geom.data(); *count = p->pk.geom.size() / 4; return OXPT_OK;
        case 1: *ptr = p->pk.attrs.data(); *count = p->pk.attrs.size() / 4; return OXPT_OK;
        case 2: *ptr = p->tables.data(); *count = p->tables.size(); return OXPT_OK;
        case 3: *ptr = p->pk.atlas.data(); *count = p->pk.atlas.size(); return OXPT_OK;
        default: return OXPT_E_ARG;
    }
}
int oxpt_packed_params(const oxpt_packed *p, oxpt_params *out) {
    if (!p || !out) return OXPT_E_ARG;
    *out = p->P;
    return OXPT_OK;
}
int oxpt_packed_info(const oxpt_packed *p, uint32_t *atlas_side, uint32_t *atlas_layers, uint32_t *bvh_depth) {
    if (!p) return OXPT_E_ARG;
    if (atlas_side) *atlas_side = p->pk.atlas_side;
    if (atlas_layers) *atlas_layers = p->pk.atlas_layers;
    if (bvh_depth) *bvh_depth = p->pk.bvh_depth;
    return OXPT_OK;
}

int oxpt_build_wide_bvh(const float *node_words, uint32_t node_count, float *out, size_t out_vec4_cap,
                        size_t *out_vec4, uint32_t *wide_depth) {
    if (!node_words || node_count == 0 || !out_vec4) return OXPT_E_ARG;
    std::vector<uint32_t> nw((size_t)node_count * 8);
    std::memcpy(nw.data(), node_words, nw.size() * 4);
    std::vector<float> wide;
    uint32_t depth = oxpt::build_wide(nw, wide);
    if (wide.empty()) return OXPT_E_INVALID;
    *out_vec4 = wide.size() / 4;
    if (wide_depth) *wide_depth = depth;
    if (!out) return OXPT_OK;
    if (out_vec4_cap < wide.size() / 4) return OXPT_E_ARG;
    std::memcpy(out, wide.data(), wide.size() * 4);
    return OXPT_OK;
}

void oxpt_abi_sizes(uint32_t out[11]) {
    const uint32_t v[11] = {sizeof(oxpt_params), sizeof(oxpt_frame), sizeof(oxpt_resolve_params),
                            sizeof(oxpt_scene_desc), sizeof(oxpt_tex_binding), sizeof(oxpt_material),
                            sizeof(oxpt_sampler), sizeof(oxpt_light), sizeof(oxpt_camera),
                            sizeof(oxpt_render_options), sizeof(oxpt_tuning)};
    std::memcpy(out, v, sizeof v);
}

uint16_t oxpt_f32_to_f16(float v) { return oxpt::f32_to_f16(v); }
void oxpt_sobol_directions(uint32_t out[OXPT_SOBOL_WORDS]) { oxpt::sobol_directions(out); }
void oxpt_sheen_table(float out[OXPT_SHEEN_WORDS]) { oxpt::sheen_table(out); }

}  // extern "C"
