// Seeded scene generators for the benchmark workloads and the tests.
// Each generator draws from its own PCG32 stream (O'Neill 2014,
// XSH-RR output) seeded from the caller's seed, so a (generator, seed)
// pair always yields the same triangle soup.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "scene.hpp"

namespace oxpt {
namespace {

struct Pcg32 {
    uint64_t state = 0, inc = 0;
    Pcg32(uint64_t seed, uint64_t stream) {
        inc = (stream << 1u) | 1u;
        next();
        state += seed;
        next();
    }
    uint32_t next() {
        uint64_t old = state;
        state = old * 6364136223846793005ULL + inc;
        uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
        uint32_t rot = (uint32_t)(old >> 59u);
        return (xs >> rot) | (xs << ((32 - rot) & 31));
    }
    float uniform() { return (float)(next() >> 8) * (1.0f / 16777216.0f); }
    float range(float a, float b) { return a + (b - a) * uniform(); }
};

struct V {
    float x, y, z;
};
inline V operator+(V a, V b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V operator-(V a, V b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V operator*(V a, float s) { return {a.x * s, a.y * s, a.z * s}; }
inline V cross(V a, V b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline float dot(V a, V b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V norm(V a) {
    float l = std::sqrt(dot(a, a));
    return l > 0 ? a * (1.0f / l) : V{0, 0, 1};
}

// Appends one triangle; n/uv may be null.
void tri(Soup &s, V a, V b, V c, uint32_t mat, const V *n = nullptr, const float *uv = nullptr) {
    for (V p : {a, b, c}) s.pos.insert(s.pos.end(), {p.x, p.y, p.z});
    uint8_t fl = 0;
    if (n) {
        for (int i = 0; i < 3; i++) s.normals.insert(s.normals.end(), {n[i].x, n[i].y, n[i].z});
        fl |= OXPT_F_NORMALS;
    } else {
        s.normals.insert(s.normals.end(), 9, 0.0f);
    }
    if (uv) {
        s.uv0.insert(s.uv0.end(), uv, uv + 6);
        fl |= OXPT_F_UV0;
    } else {
        s.uv0.insert(s.uv0.end(), 6, 0.0f);
    }
    s.tangents.insert(s.tangents.end(), 12, 0.0f);
    s.uv1.insert(s.uv1.end(), 6, 0.0f);
    s.colors.insert(s.colors.end(), 12, 1.0f);
    s.material.push_back(mat);
    s.flags.push_back(fl);
}

// Quad a-b-c-d (counter-clockwise seen from the front) as two triangles.
void quad(Soup &s, V a, V b, V c, V d, uint32_t mat, bool uvs = false) {
    float uv1[6] = {0, 0, 1, 0, 1, 1}, uv2[6] = {0, 0, 1, 1, 0, 1};
    tri(s, a, b, c, mat, nullptr, uvs ? uv1 : nullptr);
    tri(s, a, c, d, mat, nullptr, uvs ? uv2 : nullptr);
}

// Axis-aligned box of half extents `h` rotated by `ang` about +y, base at
// `base` (y = base.y), outward faces.
void box(Soup &s, V base, V h, float ang, uint32_t mat) {
    float c = std::cos(ang), sn = std::sin(ang);
    auto P = [&](float x, float y, float z) {
        return V{base.x + c * x + sn * z, base.y + y, base.z - sn * x + c * z};
    };
    V p[8];
    for (int i = 0; i < 8; i++)
        p[i] = P((i & 1) ? h.x : -h.x, (i & 2) ? 2 * h.y : 0.0f, (i & 4) ? h.z : -h.z);
    // faces: -x, +x, -y, +y, -z, +z (CCW seen from outside)
    quad(s, p[0], p[4], p[6], p[2], mat);
    quad(s, p[1], p[3], p[7], p[5], mat);
    quad(s, p[0], p[1], p[5], p[4], mat);
    quad(s, p[2], p[6], p[7], p[3], mat);
    quad(s, p[0], p[2], p[3], p[1], mat);
    quad(s, p[4], p[5], p[7], p[6], mat);
}

oxpt_material lambert(float r, float g, float b) {
    oxpt_material m;
    oxpt_material_default(&m);
    m.base_color[0] = r;
    m.base_color[1] = g;
    m.base_color[2] = b;
    m.metallic = 0.0f;
    m.roughness = 1.0f;
    m.has_specular = 1;
    m.specular = 0.0f;
    return m;
}

void set_camera(HostScene &sc, V eye, V target, float yfov) {
    sc.has_camera = true;
    sc.camera.eye[0] = eye.x;
    sc.camera.eye[1] = eye.y;
    sc.camera.eye[2] = eye.z;
    sc.camera.target[0] = target.x;
    sc.camera.target[1] = target.y;
    sc.camera.target[2] = target.z;
    sc.camera.up[0] = 0;
    sc.camera.up[1] = 1;
    sc.camera.up[2] = 0;
    sc.camera.yfov = yfov;
    sc.camera.orthographic = 0;
    sc.camera.ymag = 1.0f;
}

// Icosphere: icosahedron subdivided `sub` times, projected to the sphere.
void icosphere(std::vector<V> &out_tris, uint32_t sub) {
    const float t = (1.0f + std::sqrt(5.0f)) / 2.0f;
    V v[12] = {{-1, t, 0}, {1, t, 0}, {-1, -t, 0}, {1, -t, 0}, {0, -1, t}, {0, 1, t},
               {0, -1, -t}, {0, 1, -t}, {t, 0, -1}, {t, 0, 1}, {-t, 0, -1}, {-t, 0, 1}};
    int f[20][3] = {{0, 11, 5}, {0, 5, 1}, {0, 1, 7}, {0, 7, 10}, {0, 10, 11}, {1, 5, 9}, {5, 11, 4},
                    {11, 10, 2}, {10, 7, 6}, {7, 1, 8}, {3, 9, 4}, {3, 4, 2}, {3, 2, 6}, {3, 6, 8},
                    {3, 8, 9}, {4, 9, 5}, {2, 4, 11}, {6, 2, 10}, {8, 6, 7}, {9, 8, 1}};
    std::vector<V> tris;
    for (auto &fc : f) tris.insert(tris.end(), {norm(v[fc[0]]), norm(v[fc[1]]), norm(v[fc[2]])});
    for (uint32_t s = 0; s < sub; s++) {
        std::vector<V> nt;
        nt.reserve(tris.size() * 4);
        for (size_t i = 0; i < tris.size(); i += 3) {
            V a = tris[i], b = tris[i + 1], c = tris[i + 2];
            V ab = norm((a + b) * 0.5f), bc = norm((b + c) * 0.5f), ca = norm((c + a) * 0.5f);
            nt.insert(nt.end(), {a, ab, ca, ab, b, bc, ca, bc, c, ab, bc, ca});
        }
        tris.swap(nt);
    }
    out_tris = tris;
}

// Value noise on an integer lattice hashed from the seed, smoothstep blend.
struct ValueNoise {
    uint32_t seed;
    float lattice(int x, int y) const {
        uint32_t h = (uint32_t)x * 0x8da6b343u ^ (uint32_t)y * 0xd8163841u ^ seed * 0xcb1ab31fu;
        h ^= h >> 15;
        h *= 0x2c1b3c6du;
        h ^= h >> 12;
        h *= 0x297a2d39u;
        h ^= h >> 15;
        return (float)(h >> 8) * (1.0f / 16777216.0f);
    }
    float at(float x, float y) const {
        int x0 = (int)std::floor(x), y0 = (int)std::floor(y);
        float fx = x - x0, fy = y - y0;
        fx = fx * fx * (3 - 2 * fx);
        fy = fy * fy * (3 - 2 * fy);
        float a = lattice(x0, y0), b = lattice(x0 + 1, y0), c = lattice(x0, y0 + 1), d = lattice(x0 + 1, y0 + 1);
        return (a * (1 - fx) + b * fx) * (1 - fy) + (c * (1 - fx) + d * fx) * fy;
    }
    float fbm(float x, float y, int oct) const {
        float s = 0, amp = 0.5f, f = 1;
        for (int o = 0; o < oct; o++) {
            s += amp * (at(x * f, y * f) * 2.0f - 1.0f);
            f *= 2.0f;
            amp *= 0.5f;
        }
        return s;
    }
};

}  // namespace
}  // namespace oxpt

using namespace oxpt;

extern "C" {

oxpt_host_scene *oxpt_testscene_room(uint32_t seed) {
    auto *hs = new oxpt_host_scene();
    HostScene &sc = hs->s;
    Pcg32 rng(seed, 0x524f4f4dULL);  // "ROOM"
    // 7 Lambert materials (floor, ceiling, back, left, right, box A, box B)
    // with base colours U(0.2, 0.8)^3, then the light.
    for (int i = 0; i < 7; i++) {
        float r = rng.range(0.2f, 0.8f), g = rng.range(0.2f, 0.8f), b = rng.range(0.2f, 0.8f);
        sc.materials.push_back(lambert(r, g, b));
    }
    oxpt_material lm = lambert(0, 0, 0);
    lm.emissive[0] = lm.emissive[1] = lm.emissive[2] = 15.0f;
    sc.materials.push_back(lm);
    Soup &s = sc.soup;
    // Room x in [-1, 1], y in [0, 2], z in [-1, 1], open towards +z.
    quad(s, {-1, 0, 1}, {1, 0, 1}, {1, 0, -1}, {-1, 0, -1}, 0);    // floor, +y
    quad(s, {-1, 2, -1}, {1, 2, -1}, {1, 2, 1}, {-1, 2, 1}, 1);    // ceiling, -y
    quad(s, {-1, 0, -1}, {1, 0, -1}, {1, 2, -1}, {-1, 2, -1}, 2);  // back, +z
    quad(s, {-1, 0, 1}, {-1, 0, -1}, {-1, 2, -1}, {-1, 2, 1}, 3);  // left, +x
    quad(s, {1, 0, -1}, {1, 0, 1}, {1, 2, 1}, {1, 2, -1}, 4);      // right, -x
    float a1 = rng.range(0.1f, 0.5f), a2 = -rng.range(0.1f, 0.5f);
    box(s, {-0.35f, 0.0f, -0.3f}, {0.28f, 0.6f, 0.28f}, a1, 5);    // tall box
    box(s, {0.4f, 0.0f, 0.3f}, {0.28f, 0.28f, 0.28f}, a2, 6);      // short box
    quad(s, {-0.25f, 1.98f, -0.25f}, {0.25f, 1.98f, -0.25f}, {0.25f, 1.98f, 0.25f}, {-0.25f, 1.98f, 0.25f}, 7);
    set_camera(sc, {0.0f, 1.0f, 3.7f}, {0.0f, 1.0f, 0.0f}, 40.0f * 3.14159265f / 180.0f);
    sc.generation++;
    return hs;
}

oxpt_host_scene *oxpt_testscene_furnace(float rho, float le) {
    auto *hs = new oxpt_host_scene();
    HostScene &sc = hs->s;
    oxpt_material m = lambert(rho, rho, rho);
    m.emissive[0] = m.emissive[1] = m.emissive[2] = le;
    m.double_sided = 1;
    sc.materials.push_back(m);
    // Cube [-1, 1]^3 (tests/pathtrace.rs:155-170), faces outward.
    box(sc.soup, {0.0f, -1.0f, 0.0f}, {1.0f, 1.0f, 1.0f}, 0.0f, 0);
    V eye = {0.2f, 0.1f, 0.4f}, dir = {0.0f, -0.3f, -1.0f};
    set_camera(sc, eye, eye + dir, 1.2f);
    sc.generation++;
    return hs;
}

oxpt_host_scene *oxpt_testscene_terrain(uint32_t seed, uint32_t gx, uint32_t gz, uint32_t n_spheres,
                                        uint32_t subdiv, uint32_t n_emissive) {
    auto *hs = new oxpt_host_scene();
    HostScene &sc = hs->s;
    Pcg32 rng(seed, 0x54455252ULL);  // "TERR"
    // 32 glTF metallic-roughness materials: base U(0.05, 0.95)^3,
    // metallic ~ Bernoulli(0.3), roughness ~ U(0.05, 1).
    const uint32_t NM = 32;
    for (uint32_t i = 0; i < NM; i++) {
        oxpt_material m;
        oxpt_material_default(&m);
        m.base_color[0] = rng.range(0.05f, 0.95f);
        m.base_color[1] = rng.range(0.05f, 0.95f);
        m.base_color[2] = rng.range(0.05f, 0.95f);
        m.metallic = rng.uniform() < 0.3f ? 1.0f : 0.0f;
        m.roughness = rng.range(0.05f, 1.0f);
        sc.materials.push_back(m);
    }
    uint32_t emat = (uint32_t)sc.materials.size();
    for (int i = 0; i < 4; i++) {
        oxpt_material m;
        oxpt_material_default(&m);
        m.base_color[0] = m.base_color[1] = m.base_color[2] = 0.0f;
        m.metallic = 0.0f;
        m.emissive[0] = rng.range(20.0f, 60.0f);
        m.emissive[1] = rng.range(15.0f, 45.0f);
        m.emissive[2] = rng.range(10.0f, 30.0f);
        m.double_sided = 1;
        sc.materials.push_back(m);
    }
    Soup &s = sc.soup;
    size_t total = (size_t)gx * gz * 2 + (size_t)n_spheres * 20 * (1u << (2 * subdiv)) + n_emissive;
    s.pos.reserve(total * 9);
    s.normals.reserve(total * 9);
    s.tangents.reserve(total * 12);
    s.uv0.reserve(total * 6);
    s.uv1.reserve(total * 6);
    s.colors.reserve(total * 12);
    s.material.reserve(total);
    s.flags.reserve(total);
    // fBm heightfield over (gx x gz) quads of side `cell`, vertex normals
    // from central differences; one material per 64 x 64 patch.
    const float cell = 0.1f, amp = 4.0f;
    const float X0 = -0.5f * gx * cell, Z0 = -0.5f * gz * cell;
    ValueNoise noise{seed * 2654435761u + 17u};
    std::vector<float> hgt((size_t)(gx + 1) * (gz + 1));
    for (uint32_t j = 0; j <= gz; j++)
        for (uint32_t i = 0; i <= gx; i++)
            hgt[(size_t)j * (gx + 1) + i] = amp * noise.fbm(i * cell * 0.08f, j * cell * 0.08f, 7);
    auto H = [&](int i, int j) {
        i = std::min(std::max(i, 0), (int)gx);
        j = std::min(std::max(j, 0), (int)gz);
        return hgt[(size_t)j * (gx + 1) + i];
    };
    auto Pv = [&](uint32_t i, uint32_t j) { return V{X0 + i * cell, H(i, j), Z0 + j * cell}; };
    auto Nv = [&](uint32_t i, uint32_t j) {
        float dx = (H(i + 1, j) - H((int)i - 1, j)) / (2 * cell), dz = (H(i, j + 1) - H(i, (int)j - 1)) / (2 * cell);
        return norm(V{-dx, 1.0f, -dz});
    };
    std::vector<uint32_t> patch_mat(((gx + 63) / 64) * ((gz + 63) / 64));
    for (auto &pm : patch_mat) pm = rng.next() % NM;
    uint32_t px = (gx + 63) / 64;
    for (uint32_t j = 0; j < gz; j++)
        for (uint32_t i = 0; i < gx; i++) {
            uint32_t mat = patch_mat[(j / 64) * px + i / 64];
            V a = Pv(i, j), b = Pv(i, j + 1), c = Pv(i + 1, j + 1), d = Pv(i + 1, j);
            V na[3] = {Nv(i, j), Nv(i, j + 1), Nv(i + 1, j + 1)};
            V nb[3] = {Nv(i, j), Nv(i + 1, j + 1), Nv(i + 1, j)};
            tri(s, a, b, c, mat, na);
            tri(s, a, c, d, mat, nb);
        }
    // Icospheres resting on the terrain.
    std::vector<V> unit;
    icosphere(unit, subdiv);
    for (uint32_t k = 0; k < n_spheres; k++) {
        float r = rng.range(0.3f, 2.0f);
        float x = X0 + rng.uniform() * gx * cell, z = Z0 + rng.uniform() * gz * cell;
        float y = H((int)((x - X0) / cell), (int)((z - Z0) / cell)) + r * 0.8f;
        uint32_t mat = rng.next() % NM;
        V c = {x, y, z};
        for (size_t i = 0; i < unit.size(); i += 3) {
            V n[3] = {unit[i], unit[i + 1], unit[i + 2]};
            tri(s, c + n[0] * r, c + n[1] * r, c + n[2] * r, mat, n);
        }
    }
    // Emissive triangles hovering above the terrain (double-sided lanterns).
    for (uint32_t k = 0; k < n_emissive; k++) {
        float x = X0 + rng.uniform() * gx * cell, z = Z0 + rng.uniform() * gz * cell;
        float y = H((int)((x - X0) / cell), (int)((z - Z0) / cell)) + rng.range(1.5f, 4.0f);
        float sz = rng.range(0.3f, 0.8f), ang = rng.range(0.0f, 6.2831853f);
        V c = {x, y, z};
        V u = {std::cos(ang) * sz, 0.0f, std::sin(ang) * sz};
        V w = {0.0f, sz, 0.0f};
        tri(s, c - u, c + u, c + w, emat + (rng.next() & 3));
    }
    // Camera at the low-z end looking across the terrain.
    float ey = H(gx / 2, gz / 8) + 6.0f;
    set_camera(sc, V{0.0f, ey, Z0 + gz * cell * 0.12f}, V{0.0f, ey - 5.0f, Z0 + gz * cell * 0.6f},
               50.0f * 3.14159265f / 180.0f);
    sc.generation++;
    return hs;
}

oxpt_env_map *oxpt_env_sky(uint32_t w, uint32_t h, uint32_t seed) {
    auto *em = new oxpt_env_map();
    EnvMap &e = em->e;
    e.w = w;
    e.h = h;
    e.rgb.resize((size_t)w * h * 3);
    Pcg32 rng(seed, 0x534b59ULL);  // "SKY"
    float sun_phi = rng.range(0.0f, 6.2831853f), sun_th = rng.range(0.35f, 0.9f);  // from zenith
    float sd[3] = {std::sin(sun_th) * std::sin(sun_phi), std::cos(sun_th), -std::sin(sun_th) * std::cos(sun_phi)};
    const float PI = 3.14159265f;
    for (uint32_t y = 0; y < h; y++)
        for (uint32_t x = 0; x < w; x++) {
            // Texel centre direction (the kernel's equirectangular mapping).
            float u = (x + 0.5f) / w, v = (y + 0.5f) / h;
            float phi = 2 * PI * (u - 0.5f), th = PI * v;
            float d[3] = {std::sin(th) * std::sin(phi), std::cos(th), -std::sin(th) * std::cos(phi)};
            float c[3];
            if (d[1] >= 0.0f) {
                float t = std::pow(1.0f - d[1], 3.0f);
                c[0] = 0.25f + 0.65f * t;
                c[1] = 0.45f + 0.45f * t;
                c[2] = 0.95f + 0.05f * t;
            } else {
                float t = std::min(1.0f, -d[1] * 4.0f);
                c[0] = 0.35f * (1 - t) + 0.12f * t;
                c[1] = 0.30f * (1 - t) + 0.10f * t;
                c[2] = 0.28f * (1 - t) + 0.08f * t;
            }
            float cs = d[0] * sd[0] + d[1] * sd[1] + d[2] * sd[2];
            if (cs > 0.9996f) {  // ~1.6 degree disc
                c[0] = 3000.0f;
                c[1] = 2700.0f;
                c[2] = 2300.0f;
            } else if (cs > 0.99f) {
                float g = (cs - 0.99f) / 0.0096f;
                for (int k = 0; k < 3; k++) c[k] += 20.0f * g * g;
            }
            for (int k = 0; k < 3; k++) e.rgb[3 * ((size_t)y * w + x) + k] = c[k];
        }
    e.intensity = 1.0f;
    return em;
}

oxpt_host_scene *oxpt_testscene_layered(uint32_t seed, uint32_t n_tris, uint32_t n_textures, uint32_t tex_size,
                                        uint32_t n_lights) {
    auto *hs = new oxpt_host_scene();
    HostScene &sc = hs->s;
    Pcg32 rng(seed, 0x4c415952ULL);  // "LAYR"
    // Value-noise RGBA textures, repeat-wrapped, trilinear.
    for (uint32_t t = 0; t < n_textures; t++) {
        Texture tx;
        tx.w = tx.h = tex_size;
        tx.rgba8.resize((size_t)tex_size * tex_size * 4);
        ValueNoise vn{seed * 977u + t * 131u + 7u};
        float f = 8.0f / tex_size;
        for (uint32_t y = 0; y < tex_size; y++)
            for (uint32_t x = 0; x < tex_size; x++)
                for (int c = 0; c < 4; c++) {
                    float v = 0.5f + 0.5f * vn.fbm(x * f + 31.7f * c, y * f - 11.3f * c, 5);
                    if (c == 3) v = 0.35f + 0.65f * v;
                    tx.rgba8[4 * ((size_t)y * tex_size + x) + c] = (uint8_t)std::min(255.0f, std::max(0.0f, v * 255.0f + 0.5f));
                }
        tx.sampler = {2, 2, 1, 5};
        sc.textures.push_back(std::move(tx));
    }
    // Material families: 30% transmission + volume, 20% clearcoat, 20%
    // sheen, 10% alpha BLEND, 20% plain metallic-roughness.
    const uint32_t NM = 40;
    for (uint32_t i = 0; i < NM; i++) {
        oxpt_material m;
        oxpt_material_default(&m);
        m.base_color[0] = rng.range(0.2f, 0.95f);
        m.base_color[1] = rng.range(0.2f, 0.95f);
        m.base_color[2] = rng.range(0.2f, 0.95f);
        m.metallic = 0.0f;
        m.roughness = rng.range(0.05f, 0.9f);
        float fam = (float)i / NM;
        if (fam < 0.3f) {
            m.has_transmission = 1;
            m.transmission = rng.range(0.7f, 1.0f);
            m.has_volume = 1;
            m.thickness = rng.range(0.2f, 1.0f);
            m.attenuation_color[0] = rng.range(0.5f, 1.0f);
            m.attenuation_color[1] = rng.range(0.5f, 1.0f);
            m.attenuation_color[2] = rng.range(0.5f, 1.0f);
            m.attenuation_distance = rng.range(0.5f, 3.0f);
            m.roughness = rng.range(0.02f, 0.3f);
        } else if (fam < 0.5f) {
            m.has_clearcoat = 1;
            m.clearcoat = rng.range(0.5f, 1.0f);
            m.clearcoat_roughness = rng.range(0.02f, 0.4f);
            m.clearcoat_normal_scale = 1.0f;
        } else if (fam < 0.7f) {
            m.has_sheen = 1;
            m.sheen_color[0] = rng.range(0.4f, 1.0f);
            m.sheen_color[1] = rng.range(0.4f, 1.0f);
            m.sheen_color[2] = rng.range(0.4f, 1.0f);
            m.sheen_roughness = rng.range(0.2f, 0.9f);
        } else if (fam < 0.8f) {
            m.alpha_mode = 2;
            m.base_color[3] = rng.range(0.3f, 0.8f);
            m.double_sided = 1;
        } else {
            m.metallic = rng.uniform() < 0.5f ? 1.0f : 0.0f;
        }
        if (n_textures > 0) {
            m.tex[OXPT_S_BASE].texture = (int32_t)(rng.next() % n_textures);
            m.tex[OXPT_S_BASE].uv_set = 0;
            if (rng.uniform() < 0.5f) {
                m.tex[OXPT_S_MR].texture = (int32_t)(rng.next() % n_textures);
                m.tex[OXPT_S_MR].uv_set = 0;
            }
        }
        sc.materials.push_back(m);
    }
    // Ground plane (textured) + instanced icospheres / boxes until n_tris.
    Soup &s = sc.soup;
    oxpt_material gm;
    oxpt_material_default(&gm);
    gm.metallic = 0.0f;
    gm.roughness = 0.7f;
    if (n_textures > 0) {
        gm.tex[OXPT_S_BASE].texture = 0;
        gm.tex[OXPT_S_BASE].xform[0] = 16.0f;
        gm.tex[OXPT_S_BASE].xform[4] = 16.0f;
    }
    uint32_t gmat = (uint32_t)sc.materials.size();
    sc.materials.push_back(gm);
    const float E = 60.0f;
    quad(s, {-E, 0, E}, {E, 0, E}, {E, 0, -E}, {-E, 0, -E}, gmat, true);
    std::vector<V> unit;
    icosphere(unit, 4);  // 5120 tris per instance
    while (s.count() + unit.size() / 3 <= n_tris || s.count() < 3) {
        float r = rng.range(0.2f, 1.2f);
        V c = {rng.range(-E * 0.9f, E * 0.9f), r * rng.range(0.9f, 3.0f), rng.range(-E * 0.9f, E * 0.9f)};
        uint32_t mat = rng.next() % NM;
        for (size_t i = 0; i < unit.size(); i += 3) {
            V n[3] = {unit[i], unit[i + 1], unit[i + 2]};
            float uv[6];
            for (int k = 0; k < 3; k++) {
                uv[2 * k] = 0.5f + std::atan2(n[k].x, -n[k].z) / 6.2831853f;
                uv[2 * k + 1] = std::acos(std::max(-1.0f, std::min(1.0f, n[k].y))) / 3.14159265f;
            }
            tri(s, c + n[0] * r, c + n[1] * r, c + n[2] * r, mat, n, uv);
        }
        if (unit.size() / 3 > n_tris) break;
    }
    // Punctual lights: points and spots above the field.
    for (uint32_t i = 0; i < n_lights; i++) {
        oxpt_light l;
        std::memset(&l, 0, sizeof(l));
        l.kind = (i % 3 == 2) ? 2u : 1u;
        l.position[0] = rng.range(-E * 0.8f, E * 0.8f);
        l.position[1] = rng.range(4.0f, 10.0f);
        l.position[2] = rng.range(-E * 0.8f, E * 0.8f);
        l.direction[0] = 0;
        l.direction[1] = -1;
        l.direction[2] = 0;
        l.color[0] = rng.range(0.7f, 1.0f);
        l.color[1] = rng.range(0.7f, 1.0f);
        l.color[2] = rng.range(0.7f, 1.0f);
        l.intensity = rng.range(100.0f, 400.0f);
        l.range = 0.0f;
        l.inner_cone = 0.3f;
        l.outer_cone = 0.7f;
        sc.lights.push_back(l);
    }
    set_camera(sc, V{0.0f, 6.0f, E * 0.95f}, V{0.0f, 1.0f, 0.0f}, 50.0f * 3.14159265f / 180.0f);
    sc.generation++;
    return hs;
}

}  // extern "C"
