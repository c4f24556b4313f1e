// Device scene preparation from a host scene: the host packs materials,
// lights and texture placement (pack_scene without geometry, texels or
// emissive records); oxpt_scene_build builds the BVH, triangle / attribute
// words, 4-wide nodes, emissive lights and atlas texels on the device. The
// resulting words are the host packer's (GpuPathTracer::upload,
// oxideav-render-vulkan src/pathtrace.rs:970-1239).
#include "scene.hpp"

namespace oxpt {

int device_prepare(oxpt_ctx *ctx, const HostScene &s, Packed &pk, oxpt_scene **out, oxpt_build_info *info,
                   const std::vector<Texture> *textures) {
    pk = pack_scene(s, false, nullptr, false, false, textures);
    const std::vector<Texture> &tex_list = textures ? *textures : s.textures;
    const Soup &so = s.soup;
    const std::vector<float> mat_lum = material_emission(s);
    oxpt_soup_desc d{};
    d.n_tris = so.count();
    d.pos = so.pos.data();
    d.normals = so.normals.empty() ? nullptr : so.normals.data();
    d.tangents = so.tangents.empty() ? nullptr : so.tangents.data();
    d.uv0 = so.uv0.empty() ? nullptr : so.uv0.data();
    d.uv1 = so.uv1.empty() ? nullptr : so.uv1.data();
    d.colors = so.colors.empty() ? nullptr : so.colors.data();
    d.material = so.material.data();
    d.flags = so.flags.data();
    d.tables = pk.tables.data();
    d.table_words = pk.tables.size();
    d.atlas_side = pk.atlas_side;
    d.atlas_layers = pk.atlas_layers;
    d.material_emission = mat_lum.data();
    d.n_materials = (uint32_t)mat_lum.size();
    std::vector<oxpt_texture_desc> tex(tex_list.size());
    for (size_t i = 0; i < tex_list.size(); i++)
        tex[i] = {tex_list[i].w, tex_list[i].h, tex_list[i].rgba8.empty() ? nullptr : tex_list[i].rgba8.data()};
    float lut[512];
    texel_decode_lut(lut);
    oxpt_atlas_desc ad{tex.size(), tex.data(), pk.chains.size(), pk.chains.data(), lut};
    d.atlas_build = pk.chains.empty() ? nullptr : &ad;
    oxpt_build_info bi{};
    int rc = oxpt_scene_build(ctx, &d, nullptr, out, &bi);
    if (rc) return rc;
    // The device tables end with the emissive records it found.
    size_t tw = 0;
    rc = oxpt_scene_read(ctx, *out, 2, nullptr, 0, &tw);
    if (!rc) {
        pk.tables.resize(tw);
        rc = oxpt_scene_read(ctx, *out, 2, pk.tables.data(), tw * 4, nullptr);
    }
    if (rc) {
        oxpt_scene_destroy(*out);
        *out = nullptr;
        return rc;
    }
    pk.n_emissive = bi.n_emissive;
    pk.off_elights = bi.elight_offset;
    pk.node_count = bi.node_count;
    pk.tri_base = bi.tri_base;
    pk.wide_base = bi.wide_base;
    pk.bvh_depth = bi.bvh_depth;
    pk.wide_depth = bi.wide_depth;
    if (info) *info = bi;
    return OXPT_OK;
}

int upload_tables(oxpt_ctx *ctx, oxpt_scene *dev, const std::vector<uint32_t> &scene_tables, const EnvMap *env) {
    if (env && env->w > 0 && env->h > 0 && env->rgb.size() == (size_t)env->w * env->h * 3) {
        const std::vector<double> row_w = env_row_weights(env->h);
        oxpt_env_desc ed{env->w, env->h, env->rgb.data(), env->intensity, 0u, row_w.data()};
        return oxpt_scene_set_tables_env(ctx, dev, scene_tables.data(), scene_tables.size(), &ed);
    }
    return oxpt_scene_set_tables(ctx, dev, scene_tables.data(), scene_tables.size());
}

}  // namespace oxpt
