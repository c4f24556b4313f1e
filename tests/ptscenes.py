"""Test scenes built through the scene API, after the reference's test scenes
(tests/pathtrace.rs:145-289 and oxideav-render's testscenes): extension
materials, alpha planes, a textured checker floor, a sphere grid."""
import numpy as np

import oxpt


def uv_sphere(r, seg=24, rings=12):
    """Triangles (n, 9), normals (n, 9), uvs (n, 6) of a UV sphere."""
    tris, nrm, uvs = [], [], []
    for i in range(rings):
        t0, t1 = np.pi * i / rings, np.pi * (i + 1) / rings
        for j in range(seg):
            p0, p1 = 2 * np.pi * j / seg, 2 * np.pi * (j + 1) / seg

            def P(t, p):
                return np.array([np.sin(t) * np.cos(p), np.cos(t), np.sin(t) * np.sin(p)])

            a, b, c, d = P(t0, p0), P(t0, p1), P(t1, p1), P(t1, p0)
            ua, ub, uc, ud = (j / seg, i / rings), ((j + 1) / seg, i / rings), ((j + 1) / seg, (i + 1) / rings), (
                j / seg, (i + 1) / rings)
            for q in ((a, c, b, ua, uc, ub), (a, d, c, ua, ud, uc)):
                if np.linalg.norm(np.cross(q[1] - q[0], q[2] - q[0])) < 1e-9:
                    continue
                tris.append(np.concatenate([q[0] * r, q[1] * r, q[2] * r]))
                nrm.append(np.concatenate(q[:3]))
                uvs.append(np.concatenate(q[3:]))
    return np.array(tris, np.float32), np.array(nrm, np.float32), np.array(uvs, np.float32)


def quad(a, b, c, d):
    """Two triangles a-b-c, a-c-d with uvs."""
    pos = np.array([np.concatenate([a, b, c]), np.concatenate([a, c, d])], np.float32)
    uv = np.array([[0, 0, 1, 0, 1, 1], [0, 0, 1, 1, 0, 1]], np.float32)
    return pos, uv


def lambert(rgb, **kw):
    return oxpt.material(base_color=(*rgb, 1.0), metallic=0.0, roughness=1.0, has_specular=1, specular=0.0, **kw)


def diffuse(rgb, roughness, **kw):
    return oxpt.material(base_color=(*rgb, 1.0), metallic=0.0, roughness=roughness, **kw)


def materials_scene():
    """Floor + five spheres (volume glass, thin transmission, clearcoat, sheen,
    metal) under an emissive quad and a point light (tests/pathtrace.rs:185-269)."""
    s = oxpt.Scene()
    floor = s.add_material(diffuse((0.7, 0.7, 0.7), 0.8))
    pos, uv = quad([-3, 0, 3], [3, 0, 3], [3, 0, -3], [-3, 0, -3])
    s.add_triangles(pos, floor, uv0=uv)
    light = s.add_material(lambert((0, 0, 0), emissive=(6.0, 5.5, 5.0)))
    pos, uv = quad([-0.8, 2.5, -0.8], [0.8, 2.5, -0.8], [0.8, 2.5, 0.8], [-0.8, 2.5, 0.8])
    s.add_triangles(pos[:, [0, 1, 2, 6, 7, 8, 3, 4, 5]], light)  # facing down
    mats = [
        diffuse((0.95, 0.95, 1.0), 0.1, has_transmission=1, transmission=1.0, has_volume=1, thickness=1.0,
                attenuation_distance=1.5, attenuation_color=(0.6, 0.8, 0.9)),
        diffuse((0.9, 0.7, 0.5), 0.3, has_transmission=1, transmission=0.8),
        diffuse((0.6, 0.1, 0.1), 0.6, has_clearcoat=1, clearcoat=1.0, clearcoat_roughness=0.15),
        diffuse((0.1, 0.1, 0.3), 0.9, has_sheen=1, sheen_color=(0.9, 0.8, 0.6), sheen_roughness=0.5),
        oxpt.material(base_color=(0.95, 0.8, 0.5, 1), metallic=1.0, roughness=0.35),
    ]
    tris, nrm, uvs = uv_sphere(0.45)
    for i, m in enumerate(mats):
        mid = s.add_material(m)
        off = np.tile(np.array([-1.6 + 0.8 * i, 0.45, 0.0], np.float32), 3)
        s.add_triangles(tris + off, mid, normals=nrm, uv0=uvs)
    s.add_light("point", position=(2.0, 2.0, 2.0), color=(1.0, 0.9, 0.8), intensity=6.0)
    s.add_light("spot", position=(-2.0, 3.0, 1.0), direction=(0.5, -1.0, -0.3), color=(0.8, 0.9, 1.0),
                intensity=12.0, range=8.0, inner_cone=0.2, outer_cone=0.6)
    s.add_light("directional", direction=(-0.3, -1.0, -0.2), color=(1, 1, 1), intensity=0.3)
    s.set_camera((0.0, 1.6, 4.2), (0.0, 0.4, 0.0), yfov=0.8)
    return s


def checker(n=256, cells=16):
    y, x = np.mgrid[0:n, 0:n]
    c = (((x * cells // n) + (y * cells // n)) % 2).astype(np.uint8)
    img = np.zeros((n, n, 4), np.uint8)
    img[..., 0] = 40 + 180 * c
    img[..., 1] = 60 + 150 * c
    img[..., 2] = 200 - 150 * c
    img[..., 3] = 255
    return img


def checker_floor(sampler=(2, 2, 1, 5)):
    """A large textured floor seen at a grazing angle: ray-cone LOD + trilinear."""
    s = oxpt.Scene()
    t = s.add_texture(checker(), sampler)
    m = oxpt.material(base_color=(1, 1, 1, 1), metallic=0.0, roughness=0.8)
    m.tex[0].texture = t
    m.tex[0].xform[0] = 8.0
    m.tex[0].xform[4] = 8.0
    mid = s.add_material(m)
    pos, uv = quad([-20, 0, 20], [20, 0, 20], [20, 0, -20], [-20, 0, -20])
    s.add_triangles(pos, mid, uv0=uv)
    s.add_light("directional", direction=(-0.4, -1.0, -0.3), color=(1, 1, 1), intensity=2.0)
    s.set_camera((0.0, 1.0, 6.0), (0.0, 0.0, -10.0), yfov=0.9)
    return s


def alpha_planes():
    """Opaque backdrop, a MASK plane with a textured alpha and a BLEND plane."""
    s = oxpt.Scene()
    back = s.add_material(lambert((0.7, 0.6, 0.5)))
    pos, uv = quad([-3, -2, -1], [3, -2, -1], [3, 2, -1], [-3, 2, -1])
    s.add_triangles(pos, back, uv0=uv)
    tex = np.zeros((64, 64, 4), np.uint8)
    y, x = np.mgrid[0:64, 0:64]
    tex[..., :3] = 200
    tex[..., 3] = np.where(((x // 8) + (y // 8)) % 2 == 0, 255, 30).astype(np.uint8)
    t = s.add_texture(tex, (2, 2, 0, 0))
    mm = oxpt.material(base_color=(0.2, 0.8, 0.3, 1.0), metallic=0.0, roughness=0.6, alpha_mode=1,
                       alpha_cutoff=0.5, double_sided=1)
    mm.tex[0].texture = t
    mask = s.add_material(mm)
    pos, uv = quad([-2.5, -1.5, 0], [-0.2, -1.5, 0], [-0.2, 1.5, 0], [-2.5, 1.5, 0])
    s.add_triangles(pos, mask, uv0=uv)
    blend = s.add_material(oxpt.material(base_color=(0.9, 0.2, 0.2, 0.45), metallic=0.0, roughness=0.5,
                                         alpha_mode=2, double_sided=1))
    pos, uv = quad([0.2, -1.5, 0.5], [2.5, -1.5, 0.5], [2.5, 1.5, 0.5], [0.2, 1.5, 0.5])
    s.add_triangles(pos, blend, uv0=uv)
    lm = s.add_material(lambert((0, 0, 0), emissive=(5, 5, 5)))
    pos, _ = quad([-1, 3, 2], [1, 3, 2], [1, 3, 0], [-1, 3, 0])
    s.add_triangles(pos[:, [0, 1, 2, 6, 7, 8, 3, 4, 5]], lm)
    s.add_light("point", position=(0, 1, 3), color=(1, 1, 1), intensity=4.0)
    s.set_camera((0.0, 0.0, 4.0), (0.0, 0.0, 0.0), yfov=0.9)
    return s


def sphere_grid(nx=4, ny=3):
    """Metallic x roughness grid of spheres over a floor, lit by an emissive quad."""
    s = oxpt.Scene()
    floor = s.add_material(diffuse((0.6, 0.6, 0.6), 0.9))
    pos, uv = quad([-4, -0.5, 3], [4, -0.5, 3], [4, -0.5, -3], [-4, -0.5, -3])
    s.add_triangles(pos, floor, uv0=uv)
    tris, nrm, uvs = uv_sphere(0.35, 20, 10)
    for i in range(nx):
        for j in range(ny):
            m = s.add_material(oxpt.material(base_color=(0.8, 0.3 + 0.2 * j, 0.2, 1), metallic=i / max(nx - 1, 1),
                                             roughness=0.1 + 0.8 * j / max(ny - 1, 1)))
            off = np.tile(np.array([(i - (nx - 1) / 2) * 0.9, (j - (ny - 1) / 2) * 0.9 + 0.4, 0], np.float32), 3)
            s.add_triangles(tris + off, m, normals=nrm, uv0=uvs)
    lm = s.add_material(lambert((0, 0, 0), emissive=(8, 8, 8)))
    pos, _ = quad([-1, 3, 1], [1, 3, 1], [1, 3, -1], [-1, 3, -1])
    s.add_triangles(pos[:, [0, 1, 2, 6, 7, 8, 3, 4, 5]], lm)
    s.set_camera((0.0, 0.5, 4.5), (0.0, 0.4, 0.0), yfov=0.8)
    return s
