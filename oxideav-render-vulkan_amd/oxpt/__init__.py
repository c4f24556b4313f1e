"""Python mirror of the path tracer's API over the C ABI of ``liboxpt.so``.

The compute path is the HIP library; this module only binds it with
``ctypes`` so tests and the benchmark read like the reference's own
(oxideav-render-vulkan ``tests/pathtrace.rs``, ``examples/pt_bench.rs``):

* :class:`GpuPathTracer` — ``sync`` / ``refine`` / ``finish`` / ``image`` /
  ``hdr`` / ``reset`` / ``samples`` / ``is_converged`` / ``size`` /
  ``invalidate_scene`` / ``set_environment`` / ``render`` / ``render_hdr``
  (src/pathtrace.rs:309-725, 1241-1264);
* :data:`PATHTRACE_BACKEND_NAME` and :func:`register_into` (src/lib.rs:117,329-338);
* :class:`BackendError` / :class:`InvalidOptions` for the reference's
  ``Error::Backend`` / ``Error::InvalidOptions``.

Loading fails loudly when the library has not been built: there is no
fallback path.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field
from typing import Callable, Dict, Optional, Tuple

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "liboxpt.so")

PATHTRACE_BACKEND_NAME = "gpu-pathtrace"

OK, E_BACKEND, E_INVALID, E_ARG = 0, -1, -2, -3


class Error(Exception):
    """Base error (the reference's ``oxideav_render::Error``)."""


class BackendError(Error):
    """``Error::Backend``: no device, limits, allocation or launch failure."""


class InvalidOptions(Error):
    """``Error::InvalidOptions``: e.g. dimensions beyond the device limit."""


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it first (python -c 'import __graft_entry__ as g; g.build()')"
        )
    return C.CDLL(LIB_PATH)


lib = _load()

u32, i32, f32, u64, sz = C.c_uint32, C.c_int32, C.c_float, C.c_uint64, C.c_size_t
P = C.POINTER


class Params(C.Structure):
    _fields_ = [
        ("eye", f32 * 4), ("forward", f32 * 4), ("side", f32 * 4), ("up", f32 * 4),
        ("dims", u32 * 4), ("cfg", u32 * 4), ("env", u32 * 4),
        ("offs0", u32 * 4), ("offs1", u32 * 4), ("offs2", u32 * 4), ("fparams", f32 * 4),
    ]


class Frame(C.Structure):
    _fields_ = [("first", u32), ("count", u32), ("y0", u32), ("rows", u32)]


class ResolveParams(C.Structure):
    _fields_ = [("cfg", u32 * 4), ("exposure", f32 * 4), ("background", f32 * 4)]


class SceneDesc(C.Structure):
    _fields_ = [
        ("geom", P(f32)), ("geom_vec4", sz), ("attrs", P(f32)), ("attrs_vec4", sz),
        ("tables", P(u32)), ("table_words", sz), ("atlas", P(C.c_uint16)),
        ("atlas_side", u32), ("atlas_layers", u32),
    ]


class SoupDesc(C.Structure):
    _fields_ = [
        ("n_tris", sz), ("pos", P(f32)), ("normals", P(f32)), ("tangents", P(f32)), ("uv0", P(f32)),
        ("uv1", P(f32)), ("colors", P(f32)), ("material", P(u32)), ("flags", P(C.c_uint8)),
        ("light_of", P(u32)), ("n_emissive", u32), ("elight_offset", u32), ("tables", P(u32)),
        ("table_words", sz), ("atlas", P(C.c_uint16)), ("atlas_side", u32), ("atlas_layers", u32),
        ("atlas_build", C.c_void_p), ("material_emission", P(f32)), ("n_materials", u32),
    ]


class BuildInfo(C.Structure):
    _fields_ = [("node_count", u32), ("tri_base", u32), ("wide_base", u32), ("bvh_depth", u32),
                ("wide_depth", u32), ("pad", u32), ("upload_ms", C.c_double), ("build_ms", C.c_double),
                ("atlas_ms", C.c_double), ("n_emissive", u32), ("elight_offset", u32)]


class TexBinding(C.Structure):
    _fields_ = [("texture", i32), ("uv_set", u32), ("xform", f32 * 6)]


class Material(C.Structure):
    _fields_ = [
        ("base_color", f32 * 4), ("emissive", f32 * 3), ("alpha_mode", u32), ("alpha_cutoff", f32),
        ("double_sided", u32), ("unlit", u32),
        ("metallic", f32), ("roughness", f32), ("normal_scale", f32), ("ior", f32),
        ("has_specular", u32), ("specular", f32), ("specular_color", f32 * 3),
        ("has_transmission", u32), ("transmission", f32),
        ("has_volume", u32), ("thickness", f32), ("attenuation_color", f32 * 3), ("attenuation_distance", f32),
        ("has_clearcoat", u32), ("clearcoat", f32), ("clearcoat_roughness", f32), ("clearcoat_normal_scale", f32),
        ("has_sheen", u32), ("sheen_color", f32 * 3), ("sheen_roughness", f32),
        ("tex", TexBinding * 13),
    ]


class Sampler(C.Structure):
    _fields_ = [("wrap_s", u32), ("wrap_t", u32), ("mag", u32), ("min", u32)]


class Light(C.Structure):
    _fields_ = [
        ("kind", u32), ("position", f32 * 3), ("direction", f32 * 3), ("color", f32 * 3),
        ("intensity", f32), ("range", f32), ("inner_cone", f32), ("outer_cone", f32),
    ]


class Camera(C.Structure):
    _fields_ = [
        ("eye", f32 * 3), ("target", f32 * 3), ("up", f32 * 3), ("yfov", f32),
        ("orthographic", u32), ("ymag", f32),
    ]


class COptions(C.Structure):
    _fields_ = [
        ("width", u32), ("height", u32), ("samples_per_pixel", u32), ("max_bounces", u32),
        ("rr_start", u32), ("strategy", u32), ("seed", u32), ("clamp", f32), ("ambient", f32),
        ("background", C.c_uint8 * 4), ("tone_map", u32), ("exposure", f32),
    ]


class Tuning(C.Structure):
    _fields_ = [("block", u32), ("waves_per_cu", u32), ("max_paths", u32), ("kernel", u32), ("wf_variant", u32),
                ("stack_lds", u32), ("bvh", u32), ("tail", u32), ("early", u32), ("opaque", u32)]


STACK_PRIVATE = 0xFFFF  # oxpt_tuning.stack_lds: whole traversal stack in private memory
TAIL_NONE = 0xFFFF  # oxpt_tuning.tail: persistent kernel traces every pixel as a whole run
EARLY_OFF = 0xFFFF  # oxpt_tuning.early: the general megakernel forms even without punctual lights
OPAQUE_OFF = 0xFFFF  # oxpt_tuning.opaque: the general traversal even for all-opaque scenes


class BatchItem(C.Structure):
    _fields_ = [("scene", C.c_void_p), ("accum", C.c_void_p), ("params", Params), ("first", u32), ("count", u32)]


class ServerConfig(C.Structure):
    _fields_ = [("hbm_budget", u64), ("spp_per_tick", u32), ("max_active", u32), ("keep_hdr", u32),
                ("use_graph", u32), ("device_build", u32), ("tick_ms", u32), ("tick_paths", u64),
                ("fail_tick", u32), ("pad", u32)]


class RequestInfo(C.Structure):
    _fields_ = [("ok", u32), ("width", u32), ("height", u32), ("samples", u32), ("latency_s", C.c_double),
                ("queue_s", C.c_double), ("paths", u64)]


class ServerStats(C.Structure):
    _fields_ = [("submitted", u64), ("completed", u64), ("rejected", u64), ("ticks", u64), ("uploads", u64),
                ("evictions", u64), ("evicted_accums", u64), ("paths", u64), ("busy_s", C.c_double),
                ("resident_bytes", u64), ("active", u32), ("queued", u32), ("failed_ticks", u64),
                ("io_s", C.c_double)]


ABI_STRUCTS = [Params, Frame, ResolveParams, SceneDesc, TexBinding, Material, Sampler, Light, Camera,
               COptions, Tuning]

_vp = C.c_void_p


def _sig(name, res, *args):
    fn = getattr(lib, name)
    fn.restype = res
    fn.argtypes = list(args)
    return fn


# Layer 1.
_sig("oxpt_ctx_create", C.c_int, C.c_int, P(_vp))
_sig("oxpt_ctx_create_peer", C.c_int, _vp, P(_vp))
_sig("oxpt_image_create", C.c_int, _vp, u32, u32, C.c_int, P(_vp))
_sig("oxpt_image_destroy", None, _vp)
_sig("oxpt_image_device_ptr", _vp, _vp)
_sig("oxpt_image_bytes", sz, _vp)
_sig("oxpt_image_resolve", C.c_int, _vp, _vp, P(ResolveParams), _vp)
_sig("oxpt_image_export", C.c_int, _vp, _vp, _vp, sz)
_sig("oxpt_tracer_draw_texture", C.c_int, _vp, C.c_int, P(_vp))
_sig("oxpt_tracer_accum", _vp, _vp)
TEXTURE_RESOLVER = C.CFUNCTYPE(C.c_int, _vp, C.c_char_p, P(u32), P(u32), P(P(C.c_uint8)))
_sig("oxpt_tracer_set_texture_resolver", C.c_int, _vp, TEXTURE_RESOLVER, _vp)
_sig("oxpt_host_scene_add_texture_ref", C.c_int, _vp, C.c_char_p, P(Sampler))
_sig("oxpt_refine_batch_noalloc", C.c_int, _vp, _vp, sz)
_sig("oxpt_autotune_batch", C.c_int, _vp, _vp, sz, u32, P(Tuning), P(C.c_double))
_sig("oxpt_ctx_destroy", None, _vp)
_sig("oxpt_last_error", C.c_char_p, _vp)
_sig("oxpt_ctx_summary", C.c_char_p, _vp)
_sig("oxpt_ctx_last_kernel", C.c_char_p, _vp)
_sig("oxpt_ctx_set_budget", C.c_int, _vp, u64)
_sig("oxpt_ctx_bytes_in_use", u64, _vp)
_sig("oxpt_scene_upload", C.c_int, _vp, P(SceneDesc), P(_vp))
_sig("oxpt_scene_set_tables", C.c_int, _vp, _vp, P(u32), sz)
_sig("oxpt_scene_set_tables_env", C.c_int, _vp, _vp, P(u32), sz, _vp)
_sig("oxpt_scene_destroy", None, _vp)
_sig("oxpt_scene_build", C.c_int, _vp, P(SoupDesc), P(u32), P(_vp), P(BuildInfo))
_sig("oxpt_scene_read", C.c_int, _vp, _vp, C.c_int, _vp, sz, P(sz))
_sig("oxpt_scene_bytes", u64, _vp)
_sig("oxpt_accum_create", C.c_int, _vp, u32, u32, P(_vp))
_sig("oxpt_accum_destroy", None, _vp)
_sig("oxpt_accum_clear", C.c_int, _vp, _vp)
_sig("oxpt_accum_read", C.c_int, _vp, _vp, P(f32), sz)
_sig("oxpt_accum_write", C.c_int, _vp, _vp, P(f32), sz)
_sig("oxpt_refine", C.c_int, _vp, _vp, _vp, P(Params), u32, u32)
_sig("oxpt_finish", C.c_int, _vp)
_sig("oxpt_resolve", C.c_int, _vp, _vp, P(ResolveParams), _vp, sz)
_sig("oxpt_resolve_device", C.c_int, _vp, _vp, P(ResolveParams), _vp, sz)
_sig("oxpt_ctx_timing", C.c_int, _vp, C.c_int, P(u64), P(C.c_double))
_sig("oxpt_refine_batch", C.c_int, _vp, P(BatchItem), sz)
_sig("oxpt_ctx_set_tuning", C.c_int, _vp, P(Tuning))
_sig("oxpt_ctx_get_tuning", C.c_int, _vp, P(Tuning))
_sig("oxpt_autotune", C.c_int, _vp, _vp, P(Params), u32, P(Tuning), P(C.c_double))
_sig("oxpt_debug_pattern_words", C.c_int, _vp, _vp, u32, sz, P(u32), P(u32), P(u32), P(u32), P(u32))
_sig("oxpt_debug_light_picks", C.c_int, _vp, _vp, P(Params), sz, P(f32), P(u32))
_sig("oxpt_debug_closest_hit", C.c_int, _vp, _vp, P(Params), sz, P(f32), P(f32), P(u32))
_sig("oxpt_udiv_magic", None, u32, P(u32), P(u32))
_sig("oxpt_debug_udiv", C.c_int, _vp, sz, P(u32), P(u32), P(u32))
_sig("oxpt_debug_texel_index", C.c_int, _vp, sz, P(f32), P(C.c_int32), P(u32), P(C.c_int32))
_sig("oxpt_build_wide_bvh", C.c_int, P(f32), u32, P(f32), sz, P(sz), P(u32))
# Layer 2.
_sig("oxpt_render_options_default", None, P(COptions))
_sig("oxpt_material_default", None, P(Material))
_sig("oxpt_host_scene_new", _vp)
_sig("oxpt_host_scene_free", None, _vp)
_sig("oxpt_host_scene_add_material", C.c_int, _vp, P(Material))
_sig("oxpt_host_scene_add_texture", C.c_int, _vp, u32, u32, P(C.c_uint8), P(Sampler))
_sig("oxpt_host_scene_add_triangles", C.c_int, _vp, sz, P(f32), P(f32), P(f32), P(f32), P(f32), P(f32), P(u32))
_sig("oxpt_host_scene_add_light", C.c_int, _vp, P(Light))
_sig("oxpt_host_scene_set_camera", C.c_int, _vp, P(Camera))
_sig("oxpt_host_scene_triangle_count", sz, _vp)
_sig("oxpt_env_new", _vp, u32, u32, P(f32), f32)
_sig("oxpt_env_free", None, _vp)
_sig("oxpt_testscene_room", _vp, u32)
_sig("oxpt_testscene_furnace", _vp, f32, f32)
_sig("oxpt_testscene_terrain", _vp, u32, u32, u32, u32, u32, u32)
_sig("oxpt_env_sky", _vp, u32, u32, u32)
_sig("oxpt_testscene_layered", _vp, u32, u32, u32, u32, u32)
_sig("oxpt_tracer_new", C.c_int, P(_vp))
_sig("oxpt_tracer_new_on", C.c_int, C.c_int, P(_vp))
_sig("oxpt_tracer_free", None, _vp)
_sig("oxpt_tracer_last_error", C.c_char_p, _vp)
_sig("oxpt_tracer_adapter_summary", C.c_char_p, _vp)
_sig("oxpt_tracer_sync", C.c_int, _vp, _vp, P(COptions), P(C.c_int))
_sig("oxpt_tracer_refine", C.c_int, _vp, u32)
_sig("oxpt_tracer_finish", C.c_int, _vp)
_sig("oxpt_tracer_image", C.c_int, _vp, P(C.c_uint8), sz)
_sig("oxpt_tracer_hdr", C.c_int, _vp, P(f32), sz)
_sig("oxpt_tracer_draw", C.c_int, _vp, _vp, sz, C.c_int)
_sig("oxpt_tracer_reset", None, _vp)
_sig("oxpt_tracer_samples", u32, _vp)
_sig("oxpt_tracer_is_converged", C.c_int, _vp)
_sig("oxpt_tracer_size", None, _vp, P(u32), P(u32))
_sig("oxpt_tracer_invalidate_scene", None, _vp)
_sig("oxpt_tracer_set_environment", C.c_int, _vp, _vp)
_sig("oxpt_tracer_render", C.c_int, _vp, _vp, P(COptions), P(C.c_uint8), sz)
_sig("oxpt_tracer_render_hdr", C.c_int, _vp, _vp, P(COptions), P(f32), sz)
_sig("oxpt_tracer_ctx", _vp, _vp)
_sig("oxpt_tracer_device_scene", _vp, _vp)
_sig("oxpt_tracer_params", C.c_int, _vp, P(Params))
_sig("oxpt_tracer_words", C.c_int, _vp, C.c_int, P(_vp), P(sz))
_sig("oxpt_tracer_atlas_shape", C.c_int, _vp, P(u32), P(u32))
_sig("oxpt_tracer_set_device_build", C.c_int, _vp, C.c_int)
_sig("oxpt_tracer_build_info", C.c_int, _vp, P(BuildInfo))
_sig("oxpt_pack", C.c_int, _vp, _vp, P(COptions), P(_vp))
_sig("oxpt_packed_free", None, _vp)
_sig("oxpt_packed_words", C.c_int, _vp, C.c_int, P(_vp), P(sz))
_sig("oxpt_packed_params", C.c_int, _vp, P(Params))
_sig("oxpt_packed_info", C.c_int, _vp, P(u32), P(u32), P(u32))
_sig("oxpt_registry_make", C.c_int, C.c_char_p, P(_vp))
_sig("oxpt_ctx_set_graph", C.c_int, _vp, C.c_int)
_sig("oxpt_ctx_graph_launches", u64, _vp)
_sig("oxpt_server_config_default", None, P(ServerConfig))
_sig("oxpt_server_create", C.c_int, C.c_int, P(ServerConfig), P(_vp))
_sig("oxpt_server_destroy", None, _vp)
_sig("oxpt_server_ctx", _vp, _vp)
_sig("oxpt_server_last_error", C.c_char_p, _vp)
_sig("oxpt_server_add_scene", C.c_int, _vp, _vp, _vp, P(u32))
_sig("oxpt_server_submit_render", C.c_int, _vp, u32, P(COptions), P(u64))
_sig("oxpt_server_submit_refine", C.c_int, _vp, u64, u32, P(u64))
_sig("oxpt_server_wait", C.c_int, _vp, u64, C.c_double, P(RequestInfo))
_sig("oxpt_server_result", C.c_int, _vp, u64, C.c_int, _vp, sz)
_sig("oxpt_server_release", C.c_int, _vp, u64)
_sig("oxpt_server_stats_get", C.c_int, _vp, P(ServerStats))
_sig("oxpt_abi_sizes", None, P(u32))
_sig("oxpt_f32_to_f16", C.c_uint16, f32)
_sig("oxpt_sobol_directions", None, P(u32))
_sig("oxpt_sheen_table", None, P(f32))


def abi_sizes() -> list:
    out = (u32 * 11)()
    lib.oxpt_abi_sizes(out)
    return list(out)


def _ptr(a: np.ndarray, ct):
    return a.ctypes.data_as(P(ct))


def _raise(code: int, msg: str):
    if code == E_INVALID:
        raise InvalidOptions(msg)
    if code == E_BACKEND:
        raise BackendError(msg)
    raise Error(f"{msg} (status {code})")


STRATEGIES = {"mis": 0, "light": 1, "bsdf": 2}
TONE_MAPS = {"clamp": 0, "reinhard": 1, "aces": 2}


@dataclass
class RenderOptions:
    """RenderOptions + PathTraceOptions fields the tracer reads."""

    width: int = 256
    height: int = 256
    samples_per_pixel: int = 64
    max_bounces: int = 8
    rr_start: int = 3
    strategy: str = "mis"
    seed: int = 0
    clamp: float = 0.0
    ambient: float = 0.0
    background: Tuple[int, int, int, int] = (16, 16, 20, 255)
    tone_map: str = "clamp"
    exposure: float = 1.0

    def to_c(self) -> COptions:
        o = COptions()
        lib.oxpt_render_options_default(C.byref(o))
        o.width, o.height = self.width, self.height
        o.samples_per_pixel, o.max_bounces, o.rr_start = self.samples_per_pixel, self.max_bounces, self.rr_start
        o.strategy = STRATEGIES[self.strategy] if isinstance(self.strategy, str) else int(self.strategy)
        o.seed, o.clamp, o.ambient = self.seed, self.clamp, self.ambient
        for k in range(4):
            o.background[k] = self.background[k]
        o.tone_map = TONE_MAPS[self.tone_map] if isinstance(self.tone_map, str) else int(self.tone_map)
        o.exposure = self.exposure
        return o


def material(**kw) -> Material:
    """A glTF material with defaults; keyword fields override."""
    m = Material()
    lib.oxpt_material_default(C.byref(m))
    for k, v in kw.items():
        cur = getattr(m, k)
        if isinstance(cur, C.Array):
            for i, x in enumerate(v):
                cur[i] = x
        else:
            setattr(m, k, v)
    return m


class Scene:
    """Host scene (triangle soup, materials, textures, lights, camera)."""

    def __init__(self, handle: Optional[int] = None):
        self.h = handle if handle is not None else lib.oxpt_host_scene_new()
        if not self.h:
            raise Error("scene allocation failed")

    def __del__(self):
        if getattr(self, "h", None):
            lib.oxpt_host_scene_free(self.h)
            self.h = None

    @classmethod
    def room(cls, seed: int = 0) -> "Scene":
        return cls(lib.oxpt_testscene_room(seed))

    @classmethod
    def furnace(cls, rho: float, le: float) -> "Scene":
        return cls(lib.oxpt_testscene_furnace(rho, le))

    @classmethod
    def terrain(cls, seed=1, gx=1024, gz=512, n_spheres=200, subdiv=3, n_emissive=64) -> "Scene":
        return cls(lib.oxpt_testscene_terrain(seed, gx, gz, n_spheres, subdiv, n_emissive))

    @classmethod
    def layered(cls, seed=2, n_tris=4_000_000, n_textures=16, tex_size=2048, n_lights=16) -> "Scene":
        return cls(lib.oxpt_testscene_layered(seed, n_tris, n_textures, tex_size, n_lights))

    def add_material(self, m: Material) -> int:
        return lib.oxpt_host_scene_add_material(self.h, C.byref(m))

    def add_texture(self, rgba8: np.ndarray, sampler: Tuple[int, int, int, int] = (2, 2, 1, 5)) -> int:
        a = np.ascontiguousarray(rgba8, dtype=np.uint8)
        s = Sampler(*sampler)
        return lib.oxpt_host_scene_add_texture(self.h, a.shape[1], a.shape[0], _ptr(a, C.c_uint8), C.byref(s))

    def add_texture_ref(self, uri: str, sampler: Tuple[int, int, int, int] = (2, 2, 1, 5)) -> int:
        """A texture named by uri; a tracer's texture resolver supplies its pixels."""
        sm = Sampler(*sampler)
        i = lib.oxpt_host_scene_add_texture_ref(self.h, uri.encode(), C.byref(sm))
        if i < 0:
            _raise(i, "add_texture_ref")
        return i

    def add_triangles(self, pos, material, normals=None, tangents=None, uv0=None, uv1=None, colors=None):
        pos = np.ascontiguousarray(pos, dtype=np.float32).reshape(-1, 9)
        n = pos.shape[0]
        mat = np.ascontiguousarray(np.broadcast_to(np.asarray(material, dtype=np.uint32), (n,)))

        def opt(a, per):
            if a is None:
                return None, None
            arr = np.ascontiguousarray(a, dtype=np.float32).reshape(n, per)
            return arr, _ptr(arr, f32)

        keep = [opt(normals, 9), opt(tangents, 12), opt(uv0, 6), opt(uv1, 6), opt(colors, 12)]
        rc = lib.oxpt_host_scene_add_triangles(self.h, n, _ptr(pos, f32), *[k[1] for k in keep], _ptr(mat, u32))
        if rc:
            _raise(rc, "add_triangles")

    def add_light(self, kind="point", position=(0, 0, 0), direction=(0, -1, 0), color=(1, 1, 1), intensity=1.0,
                  range=0.0, inner_cone=0.0, outer_cone=0.7853982):
        k = {"directional": 0, "point": 1, "spot": 2}[kind] if isinstance(kind, str) else int(kind)
        l = Light(k, (f32 * 3)(*position), (f32 * 3)(*direction), (f32 * 3)(*color), intensity, range, inner_cone,
                  outer_cone)
        lib.oxpt_host_scene_add_light(self.h, C.byref(l))

    def set_camera(self, eye, target, yfov=0.8, up=(0, 1, 0), orthographic=False, ymag=1.0):
        c = Camera((f32 * 3)(*eye), (f32 * 3)(*target), (f32 * 3)(*up), yfov, 1 if orthographic else 0, ymag)
        lib.oxpt_host_scene_set_camera(self.h, C.byref(c))

    def triangle_count(self) -> int:
        return lib.oxpt_host_scene_triangle_count(self.h)


class EnvironmentMap:
    """Equirectangular HDR environment (EnvironmentMap)."""

    def __init__(self, rgb: Optional[np.ndarray] = None, intensity: float = 1.0, handle=None):
        if handle is not None:
            self.h = handle
        else:
            a = np.ascontiguousarray(rgb, dtype=np.float32)
            self.h = lib.oxpt_env_new(a.shape[1], a.shape[0], _ptr(a, f32), intensity)
        if not self.h:
            raise Error("environment allocation failed")

    def __del__(self):
        if getattr(self, "h", None):
            lib.oxpt_env_free(self.h)
            self.h = None

    @classmethod
    def sky(cls, w=1024, h=512, seed=0) -> "EnvironmentMap":
        return cls(handle=lib.oxpt_env_sky(w, h, seed))


class Packed:
    """Host-only packed words + Params (no device), for the CPU oracle."""

    def __init__(self, scene: Scene, opts: RenderOptions, env: Optional[EnvironmentMap] = None):
        h = _vp()
        o = opts.to_c()
        rc = lib.oxpt_pack(scene.h, env.h if env else None, C.byref(o), C.byref(h))
        if rc:
            _raise(rc, "pack")
        self.h = h.value
        self.geom = self._words(0, np.float32, 4)
        self.attrs = self._words(1, np.float32, 4)
        self.tables = self._words(2, np.uint32, 1)
        self.atlas = self._words(3, np.uint16, 1)
        p = Params()
        lib.oxpt_packed_params(self.h, C.byref(p))
        self.params = p
        side, layers, depth = u32(), u32(), u32()
        lib.oxpt_packed_info(self.h, C.byref(side), C.byref(layers), C.byref(depth))
        self.atlas_side, self.atlas_layers, self.bvh_depth = side.value, layers.value, depth.value

    def _words(self, which, dt, per):
        ptr, n = _vp(), sz()
        lib.oxpt_packed_words(self.h, which, C.byref(ptr), C.byref(n))
        if n.value == 0 or not ptr.value:
            return np.zeros(0, dtype=dt)
        ct = {np.float32: f32, np.uint32: u32, np.uint16: C.c_uint16}[dt]
        arr = np.ctypeslib.as_array(C.cast(ptr.value, P(ct)), shape=(n.value * per,))
        return arr.copy()

    def __del__(self):
        if getattr(self, "h", None):
            lib.oxpt_packed_free(self.h)
            self.h = None


class GpuPathTracer:
    """Progressive GPU path tracer (GpuPathTracer, src/pathtrace.rs:186-725)."""

    def __init__(self, device: int = 0, _handle: Optional[int] = None):
        if _handle is None:
            h = _vp()
            rc = lib.oxpt_tracer_new_on(device, C.byref(h))
            if rc:
                _raise(rc, "no HIP device: the GPU path tracer needs a gfx950 device")
            _handle = h.value
        self.h = _handle
        self._opts: Optional[RenderOptions] = None

    def __del__(self):
        if getattr(self, "h", None):
            lib.oxpt_tracer_free(self.h)
            self.h = None

    def _check(self, rc, what):
        if rc:
            _raise(rc, f"{what}: {lib.oxpt_tracer_last_error(self.h).decode()}")

    def adapter_summary(self) -> str:
        return lib.oxpt_tracer_adapter_summary(self.h).decode()

    def sync(self, scene: Scene, opts: RenderOptions) -> bool:
        o = opts.to_c()
        r = C.c_int(0)
        self._check(lib.oxpt_tracer_sync(self.h, scene.h if scene else None, C.byref(o), C.byref(r)), "sync")
        self._opts = opts
        return bool(r.value)

    def refine(self, samples: int) -> None:
        self._check(lib.oxpt_tracer_refine(self.h, samples), "refine")

    def finish(self) -> None:
        self._check(lib.oxpt_tracer_finish(self.h), "finish")

    def size(self) -> Tuple[int, int]:
        w, h = u32(), u32()
        lib.oxpt_tracer_size(self.h, C.byref(w), C.byref(h))
        return w.value, h.value

    def image(self) -> np.ndarray:
        w, h = self.size()
        w, h = max(w, 1), max(h, 1)
        out = np.zeros((h, w, 4), dtype=np.uint8)
        self._check(lib.oxpt_tracer_image(self.h, _ptr(out, C.c_uint8), out.nbytes), "image")
        return out

    def hdr(self) -> np.ndarray:
        w, h = self.size()
        w, h = max(w, 1), max(h, 1)
        out = np.zeros((h, w, 4), dtype=np.float32)
        self._check(lib.oxpt_tracer_hdr(self.h, _ptr(out, f32), out.nbytes), "hdr")
        return out

    def reset(self) -> None:
        lib.oxpt_tracer_reset(self.h)

    def samples(self) -> int:
        return lib.oxpt_tracer_samples(self.h)

    def is_converged(self) -> bool:
        return bool(lib.oxpt_tracer_is_converged(self.h))

    def invalidate_scene(self) -> None:
        lib.oxpt_tracer_invalidate_scene(self.h)

    def set_environment(self, env: Optional[EnvironmentMap]) -> None:
        self._check(lib.oxpt_tracer_set_environment(self.h, env.h if env else None), "set_environment")

    def render(self, scene: Scene, opts: RenderOptions) -> np.ndarray:
        o = opts.to_c()
        w, h = max(opts.width, 1), max(opts.height, 1)
        out = np.zeros((h, w, 4), dtype=np.uint8)
        self._check(lib.oxpt_tracer_render(self.h, scene.h, C.byref(o), _ptr(out, C.c_uint8), out.nbytes), "render")
        self._opts = opts
        return out

    def render_hdr(self, scene: Scene, opts: RenderOptions) -> np.ndarray:
        o = opts.to_c()
        w, h = max(opts.width, 1), max(opts.height, 1)
        out = np.zeros((h, w, 4), dtype=np.float32)
        self._check(lib.oxpt_tracer_render_hdr(self.h, scene.h, C.byref(o), _ptr(out, f32), out.nbytes),
                    "render_hdr")
        self._opts = opts
        return out

    # ---- device access (timing, tuning, parity probes) -------------------
    @property
    def ctx(self) -> int:
        return lib.oxpt_tracer_ctx(self.h)

    def device_scene(self) -> int:
        s = lib.oxpt_tracer_device_scene(self.h)
        if not s:
            raise Error("device scene: not synced")
        return s

    def params(self) -> Params:
        p = Params()
        self._check(lib.oxpt_tracer_params(self.h, C.byref(p)), "params")
        return p

    def words(self, which: int) -> np.ndarray:
        ptr, n = _vp(), sz()
        self._check(lib.oxpt_tracer_words(self.h, which, C.byref(ptr), C.byref(n)), "words")
        dt, ct, per = [(np.float32, f32, 4), (np.float32, f32, 4), (np.uint32, u32, 1), (np.uint16, C.c_uint16, 1)][which]
        if n.value == 0:
            return np.zeros(0, dtype=dt)
        return np.ctypeslib.as_array(C.cast(ptr.value, P(ct)), shape=(n.value * per,)).copy()

    def atlas_shape(self) -> Tuple[int, int]:
        s, l = u32(), u32()
        self._check(lib.oxpt_tracer_atlas_shape(self.h, C.byref(s), C.byref(l)), "atlas_shape")
        return s.value, l.value

    def draw(self, device_ptr: int, nbytes: int, hdr: bool = False) -> None:
        """draw_texture: resolve into device memory at device_ptr (RGBA8 or
        RGBA32F rows), asynchronously; finish() orders host reads after it."""
        self._check(lib.oxpt_tracer_draw(self.h, _vp(device_ptr), nbytes, 1 if hdr else 0), "draw")

    def set_texture_resolver(self, resolver: Optional[Callable[[str], Optional[np.ndarray]]]) -> None:
        """set_texture_resolver: resolver(uri) -> (h, w, 4) uint8 RGBA8 level 0, or
        None when it cannot decode uri. None removes it. The scene is re-prepared
        (and the accumulation reset) on the next sync."""
        if resolver is None:
            self._resolver = None
            self._check(lib.oxpt_tracer_set_texture_resolver(self.h, TEXTURE_RESOLVER(), None), "set_texture_resolver")
            return
        keep = {}

        def trampoline(_user, uri, w, h, px):
            try:
                img = resolver(uri.decode())
            except Exception:  # a failing decoder leaves the texture absent
                return 1
            if img is None:
                return 1
            a = np.ascontiguousarray(img, dtype=np.uint8)
            if a.ndim != 3 or a.shape[2] != 4:
                return 1
            keep["last"] = a  # read by the tracer before the next call
            w[0], h[0] = a.shape[1], a.shape[0]
            px[0] = a.ctypes.data_as(P(C.c_uint8))
            return 0

        cb = TEXTURE_RESOLVER(trampoline)
        self._resolver = (cb, keep)
        self._check(lib.oxpt_tracer_set_texture_resolver(self.h, cb, None), "set_texture_resolver")

    def draw_texture(self, hdr: bool = False) -> "Image":
        """draw_texture into the tracer's own image (resolved asynchronously;
        finish() before reading it elsewhere). The image stays the tracer's."""
        h = _vp()
        self._check(lib.oxpt_tracer_draw_texture(self.h, 1 if hdr else 0, C.byref(h)), "draw_texture")
        w, hh = self.size()
        return Image(h.value, self.ctx, max(w, 1), max(hh, 1), hdr)

    def accum(self) -> np.ndarray:
        """The current accumulator (float4 per pixel: rgb sums, covered count)."""
        a = lib.oxpt_tracer_accum(self.h)
        if not a:
            raise Error("accum: not synced")
        w, h = self.size()
        out = np.zeros((h, w, 4), dtype=np.float32)
        rc = lib.oxpt_accum_read(self.ctx, a, _ptr(out, f32), out.nbytes)
        if rc:
            _raise(rc, lib.oxpt_last_error(self.ctx).decode())
        return out

    def set_device_build(self, enable: bool = True) -> None:
        """Prepare scenes on the device (BVH, triangle / attribute words,
        4-wide nodes) from the next sync on; the words are the host packer's."""
        self._check(lib.oxpt_tracer_set_device_build(self.h, 1 if enable else 0), "set_device_build")

    def build_info(self) -> BuildInfo:
        b = BuildInfo()
        self._check(lib.oxpt_tracer_build_info(self.h, C.byref(b)), "build_info")
        return b

    def scene_words(self, which: int) -> np.ndarray:
        """The resident scene's words as the device holds them (oxpt_scene_read)."""
        scene = self.device_scene()
        n = sz()
        rc = lib.oxpt_scene_read(self.ctx, scene, which, None, 0, C.byref(n))
        if rc:
            _raise(rc, lib.oxpt_last_error(self.ctx).decode())
        dt, per, elem = [(np.float32, 4, 16), (np.float32, 4, 16), (np.uint32, 1, 4), (np.uint16, 1, 2)][which]
        out = np.zeros(n.value * per, dtype=dt)
        rc = lib.oxpt_scene_read(self.ctx, scene, which, out.ctypes.data_as(_vp), n.value * elem, None)
        if rc:
            _raise(rc, lib.oxpt_last_error(self.ctx).decode())
        return out

    def timing(self, reset: bool = False) -> Tuple[int, float]:
        n, ms = u64(), C.c_double()
        lib.oxpt_ctx_timing(self.ctx, 1 if reset else 0, C.byref(n), C.byref(ms))
        return n.value, ms.value

    def last_kernel(self) -> str:
        """Name of the trace kernel the last refine launched."""
        return lib.oxpt_ctx_last_kernel(self.ctx).decode()

    def set_tuning(self, block=0, waves_per_cu=0, max_paths=0, kernel=0, wf_variant=0, stack_lds=0, bvh=0,
                   tail=0, early=0, opaque=0) -> None:
        t = Tuning(block, waves_per_cu, max_paths, kernel, wf_variant, stack_lds, bvh, tail, early, opaque)
        rc = lib.oxpt_ctx_set_tuning(self.ctx, C.byref(t))
        if rc:
            _raise(rc, lib.oxpt_last_error(self.ctx).decode())

    def get_tuning(self) -> Tuning:
        t = Tuning()
        lib.oxpt_ctx_get_tuning(self.ctx, C.byref(t))
        return t

    def autotune(self, spp_per_trial: int = 4) -> Tuple[Tuning, float]:
        """Pick the fastest kernel form for the synced scene (oxpt_autotune);
        valid after sync + one refine. Images do not change."""
        t, ms = Tuning(), C.c_double()
        P = self.params()
        rc = lib.oxpt_autotune(self.ctx, self.device_scene(), C.byref(P), spp_per_trial, C.byref(t), C.byref(ms))
        if rc:
            _raise(rc, lib.oxpt_last_error(self.ctx).decode())
        return t, ms.value

    def debug_pattern_words(self, seed, x, y, index, p) -> np.ndarray:
        arrs = [np.ascontiguousarray(a, dtype=np.uint32) for a in (x, y, index, p)]
        n = arrs[0].size
        out = np.zeros((n, 4), dtype=np.uint32)
        rc = lib.oxpt_debug_pattern_words(self.ctx, self.device_scene(), seed, n, *[_ptr(a, u32) for a in arrs],
                                          _ptr(out, u32))
        if rc:
            _raise(rc, lib.oxpt_last_error(self.ctx).decode())
        return out

    def debug_light_picks(self, u) -> np.ndarray:
        u = np.ascontiguousarray(u, dtype=np.float32)
        out = np.zeros((u.size, 3), dtype=np.uint32)
        p = self.params()
        rc = lib.oxpt_debug_light_picks(self.ctx, self.device_scene(), C.byref(p), u.size, _ptr(u, f32), _ptr(out, u32))
        if rc:
            _raise(rc, lib.oxpt_last_error(self.ctx).decode())
        return out

    def debug_udiv(self, num, den) -> np.ndarray:
        """num // den on the device with the claim step's multiply-shift division."""
        num = np.ascontiguousarray(num, dtype=np.uint32)
        den = np.ascontiguousarray(den, dtype=np.uint32)
        out = np.zeros(num.size, dtype=np.uint32)
        rc = lib.oxpt_debug_udiv(self.ctx, num.size, _ptr(num, u32), _ptr(den, u32), _ptr(out, u32))
        if rc:
            _raise(rc, lib.oxpt_last_error(self.ctx).decode())
        return out

    def debug_texel_index(self, u, size, mode) -> np.ndarray:
        """{nearest, bilinear x0, x0 + 1} wrapped texel indices as the kernels form them."""
        u = np.ascontiguousarray(u, dtype=np.float32)
        size = np.ascontiguousarray(size, dtype=np.int32)
        mode = np.ascontiguousarray(mode, dtype=np.uint32)
        out = np.zeros((u.size, 3), dtype=np.int32)
        rc = lib.oxpt_debug_texel_index(self.ctx, u.size, _ptr(u, f32), _ptr(size, C.c_int32), _ptr(mode, u32),
                                        _ptr(out, C.c_int32))
        if rc:
            _raise(rc, lib.oxpt_last_error(self.ctx).decode())
        return out

    def debug_closest_hit(self, o, d) -> np.ndarray:
        o = np.ascontiguousarray(o, dtype=np.float32).reshape(-1, 3)
        d = np.ascontiguousarray(d, dtype=np.float32).reshape(-1, 3)
        out = np.zeros((o.shape[0], 2), dtype=np.uint32)
        p = self.params()
        rc = lib.oxpt_debug_closest_hit(self.ctx, self.device_scene(), C.byref(p), o.shape[0], _ptr(o, f32),
                                        _ptr(d, f32), _ptr(out, u32))
        if rc:
            _raise(rc, lib.oxpt_last_error(self.ctx).decode())
        return out


_hip = None


def _hip_runtime():
    """The HIP runtime liboxpt.so itself links (already loaded with it)."""
    global _hip
    if _hip is None:
        _hip = C.CDLL("libamdhip64.so.7")
        _hip.hipMalloc.argtypes = [P(_vp), sz]
        _hip.hipFree.argtypes = [_vp]
        _hip.hipMemcpy.argtypes = [_vp, _vp, sz, C.c_int]
        _hip.hipSetDevice.argtypes = [C.c_int]
    return _hip


class Image:
    """A display image the library owns (oxpt_image): device rows of RGBA8 or
    RGBA32F. Not freed here: a tracer's draw_texture image belongs to the tracer."""

    def __init__(self, handle: int, ctx: int, w: int, h: int, hdr: bool):
        self.h, self.ctx, self.w, self.hgt, self.hdr = handle, ctx, w, h, hdr

    @property
    def ptr(self) -> int:
        return lib.oxpt_image_device_ptr(self.h)

    @property
    def nbytes(self) -> int:
        return lib.oxpt_image_bytes(self.h)

    def read(self) -> np.ndarray:
        """Device -> host copy (after the producer's finish)."""
        out = np.empty((self.hgt, self.w, 4), dtype=np.float32 if self.hdr else np.uint8)
        if _hip_runtime().hipMemcpy(out.ctypes.data_as(_vp), _vp(self.ptr), out.nbytes, 2) != 0:
            raise BackendError("hipMemcpy failed")
        return out

    def export(self) -> bytes:
        """The 64-byte HIP IPC handle another process opens with hipIpcOpenMemHandle."""
        buf = (C.c_uint8 * 64)()
        rc = lib.oxpt_image_export(self.ctx, self.h, buf, 64)
        if rc:
            _raise(rc, lib.oxpt_last_error(self.ctx).decode())
        return bytes(buf)


class DeviceBuffer:
    """Raw device memory on one GPU (a draw target for GpuPathTracer.draw),
    allocated through liboxpt's own HIP runtime."""

    def __init__(self, nbytes: int, device: int = 0):
        hip = _hip_runtime()
        hip.hipSetDevice(device)
        p = _vp()
        if hip.hipMalloc(C.byref(p), nbytes) != 0 or not p.value:
            raise BackendError(f"hipMalloc of {nbytes} bytes failed")
        self.ptr, self.nbytes = p.value, nbytes

    def read(self, dtype=np.uint8) -> np.ndarray:
        out = np.empty(self.nbytes // np.dtype(dtype).itemsize, dtype=dtype)
        if _hip_runtime().hipMemcpy(out.ctypes.data_as(_vp), _vp(self.ptr), out.nbytes, 2) != 0:  # device->host
            raise BackendError("hipMemcpy failed")
        return out

    def __del__(self):
        if getattr(self, "ptr", None):
            _hip_runtime().hipFree(_vp(self.ptr))
            self.ptr = None


class Device:
    """The device boundary (layer 1): resident scenes and accumulators on one context."""

    def __init__(self, ordinal: int = 0):
        h = _vp()
        rc = lib.oxpt_ctx_create(ordinal, C.byref(h))
        if rc:
            _raise(rc, "no HIP device")
        self.h = h.value
        self._keep = []

    def __del__(self):
        if getattr(self, "h", None):
            lib.oxpt_ctx_destroy(self.h)
            self.h = None

    def _check(self, rc, what):
        if rc:
            _raise(rc, f"{what}: {lib.oxpt_last_error(self.h).decode()}")

    def summary(self) -> str:
        return lib.oxpt_ctx_summary(self.h).decode()

    def upload(self, pk: "Packed") -> "DeviceScene":
        return DeviceScene(self, pk)

    def accum(self, w: int, h: int) -> "Accum":
        return Accum(self, w, h)

    def refine(self, scene: "DeviceScene", acc: "Accum", params: Params, first: int, count: int) -> None:
        self._check(lib.oxpt_refine(self.h, scene.h, acc.h, C.byref(params), first, count), "refine")

    def refine_batch(self, items) -> None:
        """items: iterable of (DeviceScene, Accum, Params, first, count)."""
        arr = (BatchItem * len(items))()
        for i, (sc, acc, prm, first, count) in enumerate(items):
            arr[i] = BatchItem(sc.h, acc.h, prm, first, count)
        self._check(lib.oxpt_refine_batch(self.h, arr, len(items)), "refine_batch")

    def refine_batch_noalloc(self, items) -> None:
        """oxpt_refine_batch_noalloc: the same, with no device allocation."""
        arr = (BatchItem * len(items))()
        for i, (sc, acc, prm, first, count) in enumerate(items):
            arr[i] = BatchItem(sc.h, acc.h, prm, first, count)
        self._check(lib.oxpt_refine_batch_noalloc(self.h, arr, len(items)), "refine_batch_noalloc")

    def autotune_batch(self, items, spp_per_trial: int = 16) -> Tuple[Tuning, float]:
        """Pick the fastest form / samples per dispatch / stack split for these items (oxpt_autotune_batch)."""
        arr = (BatchItem * len(items))()
        for i, (sc, acc, prm, first, count) in enumerate(items):
            arr[i] = BatchItem(sc.h, acc.h, prm, first, count)
        t, ms = Tuning(), C.c_double()
        self._check(lib.oxpt_autotune_batch(self.h, arr, len(items), spp_per_trial, C.byref(t), C.byref(ms)),
                    "autotune_batch")
        return t, ms.value

    def last_kernel(self) -> str:
        return lib.oxpt_ctx_last_kernel(self.h).decode()

    def finish(self) -> None:
        self._check(lib.oxpt_finish(self.h), "finish")

    def timing(self, reset: bool = False):
        n, ms = u64(), C.c_double()
        lib.oxpt_ctx_timing(self.h, 1 if reset else 0, C.byref(n), C.byref(ms))
        return n.value, ms.value

    def set_tuning(self, block=0, waves_per_cu=0, max_paths=0, kernel=0, wf_variant=0, stack_lds=0, bvh=0,
                   tail=0, early=0, opaque=0) -> None:
        t = Tuning(block, waves_per_cu, max_paths, kernel, wf_variant, stack_lds, bvh, tail, early, opaque)
        self._check(lib.oxpt_ctx_set_tuning(self.h, C.byref(t)), "set_tuning")

    def get_tuning(self) -> Tuning:
        t = Tuning()
        lib.oxpt_ctx_get_tuning(self.h, C.byref(t))
        return t

    def set_budget(self, nbytes: int) -> None:
        self._check(lib.oxpt_ctx_set_budget(self.h, nbytes), "set_budget")

    def set_graph(self, enable: bool) -> None:
        """Replay the batched step from a captured HIP graph."""
        self._check(lib.oxpt_ctx_set_graph(self.h, 1 if enable else 0), "set_graph")

    def graph_launches(self) -> int:
        return lib.oxpt_ctx_graph_launches(self.h)

    def bytes_in_use(self) -> int:
        return lib.oxpt_ctx_bytes_in_use(self.h)


class DeviceScene:
    def __init__(self, dev: Device, pk: "Packed"):
        self.dev = dev
        d = SceneDesc()
        d.geom, d.geom_vec4 = _ptr(pk.geom, f32), pk.geom.size // 4
        d.attrs, d.attrs_vec4 = _ptr(pk.attrs, f32), pk.attrs.size // 4
        d.tables, d.table_words = _ptr(pk.tables, u32), pk.tables.size
        d.atlas = _ptr(pk.atlas, C.c_uint16) if pk.atlas.size else None
        d.atlas_side, d.atlas_layers = pk.atlas_side, pk.atlas_layers
        h = _vp()
        dev._check(lib.oxpt_scene_upload(dev.h, C.byref(d), C.byref(h)), "scene_upload")
        self.h = h.value
        self.params = pk.params

    def __del__(self):
        if getattr(self, "h", None) and getattr(self.dev, "h", None):
            lib.oxpt_scene_destroy(self.h)
            self.h = None


class Accum:
    def __init__(self, dev: Device, w: int, h: int):
        self.dev, self.w, self.h_ = dev, w, h
        hh = _vp()
        dev._check(lib.oxpt_accum_create(dev.h, w, h, C.byref(hh)), "accum_create")
        self.h = hh.value

    def __del__(self):
        if getattr(self, "h", None) and getattr(self.dev, "h", None):
            lib.oxpt_accum_destroy(self.h)
            self.h = None

    def clear(self) -> None:
        self.dev._check(lib.oxpt_accum_clear(self.dev.h, self.h), "accum_clear")

    def read(self) -> np.ndarray:
        out = np.zeros((self.h_, self.w, 4), dtype=np.float32)
        self.dev._check(lib.oxpt_accum_read(self.dev.h, self.h, _ptr(out, f32), out.nbytes), "accum_read")
        return out

    def resolve(self, rp: ResolveParams) -> np.ndarray:
        hdr = rp.cfg[3] == 1
        out = np.zeros((self.h_, self.w, 4), dtype=np.float32 if hdr else np.uint8)
        self.dev._check(lib.oxpt_resolve(self.dev.h, self.h, C.byref(rp), out.ctypes.data_as(C.c_void_p), out.nbytes),
                        "resolve")
        return out


class Server:
    """Continuous-batching render server over one device (include/oxpt/oxpt.h, serving)."""

    def __init__(self, device=0, hbm_budget=200 << 30, spp_per_tick=16, max_active=0, keep_hdr=False,
                 use_graph=False, device_build=False, tick_ms=0, tick_paths=0, fail_tick=0):
        cfg = ServerConfig()
        lib.oxpt_server_config_default(C.byref(cfg))
        cfg.hbm_budget, cfg.spp_per_tick, cfg.max_active = hbm_budget, spp_per_tick, max_active
        cfg.keep_hdr, cfg.use_graph = 1 if keep_hdr else 0, 1 if use_graph else 0
        cfg.device_build = 1 if device_build else 0
        cfg.tick_ms, cfg.fail_tick = tick_ms, fail_tick
        if tick_paths:
            cfg.tick_paths = tick_paths
        h = _vp()
        rc = lib.oxpt_server_create(device, C.byref(cfg), C.byref(h))
        if rc:
            _raise(rc, "server: no HIP device")
        self.h = h.value

    def close(self):
        if getattr(self, "h", None):
            lib.oxpt_server_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def _check(self, rc, what):
        if rc:
            _raise(rc, f"{what}: {lib.oxpt_server_last_error(self.h).decode()}")

    def add_scene(self, scene: Scene, env: Optional[EnvironmentMap] = None) -> int:
        i = u32()
        self._check(lib.oxpt_server_add_scene(self.h, scene.h, env.h if env else None, C.byref(i)), "add_scene")
        return i.value

    def submit_render(self, scene_id: int, opts: RenderOptions) -> int:
        r = u64()
        o = opts.to_c()
        self._check(lib.oxpt_server_submit_render(self.h, scene_id, C.byref(o), C.byref(r)), "submit_render")
        return r.value

    def submit_refine(self, render_req: int, spp: int) -> int:
        r = u64()
        self._check(lib.oxpt_server_submit_refine(self.h, render_req, spp, C.byref(r)), "submit_refine")
        return r.value

    def wait(self, req: int, timeout: float = -1.0) -> RequestInfo:
        info = RequestInfo()
        self._check(lib.oxpt_server_wait(self.h, req, timeout, C.byref(info)), f"wait({req})")
        return info

    def result(self, req: int, hdr: bool = False) -> np.ndarray:
        info = RequestInfo()
        lib.oxpt_server_wait(self.h, req, 0.0, C.byref(info))
        out = np.zeros((info.height, info.width, 4), dtype=np.float32 if hdr else np.uint8)
        self._check(lib.oxpt_server_result(self.h, req, 1 if hdr else 0, out.ctypes.data_as(C.c_void_p), out.nbytes),
                    "result")
        return out

    def release(self, req: int) -> None:
        lib.oxpt_server_release(self.h, req)

    def stats(self) -> ServerStats:
        st = ServerStats()
        lib.oxpt_server_stats_get(self.h, C.byref(st))
        return st

    def graph_launches(self) -> int:
        return lib.oxpt_ctx_graph_launches(lib.oxpt_server_ctx(self.h))

    def kernel_time(self, reset: bool = False):
        """(launches, total ms) of the batched steps the server launched, from
        the HIP events around each launch on its context's stream."""
        n, ms = u64(), C.c_double()
        lib.oxpt_ctx_timing(lib.oxpt_server_ctx(self.h), 1 if reset else 0, C.byref(n), C.byref(ms))
        return n.value, ms.value


class RenderRegistry:
    """Name -> factory registry (oxideav_render::RenderRegistry)."""

    def __init__(self):
        self._f: Dict[str, Callable[[], GpuPathTracer]] = {}

    def register(self, name: str, factory: Callable[[], GpuPathTracer]) -> None:
        self._f[name] = factory

    def make(self, name: str) -> GpuPathTracer:
        if name not in self._f:
            raise InvalidOptions(f"no renderer named {name!r}")
        return self._f[name]()


def _make_registered(name: str) -> GpuPathTracer:
    h = _vp()
    rc = lib.oxpt_registry_make(name.encode(), C.byref(h))
    if rc:
        _raise(rc, f"registry: {name}")
    return GpuPathTracer(_handle=h.value)


def register_into(registry: RenderRegistry) -> None:
    """Register the path tracer under PATHTRACE_BACKEND_NAME (devices open lazily, per make)."""
    registry.register(PATHTRACE_BACKEND_NAME, lambda: _make_registered(PATHTRACE_BACKEND_NAME))


def f32_to_f16(v: float) -> int:
    return lib.oxpt_f32_to_f16(v)


def sobol_directions() -> np.ndarray:
    out = np.zeros(128, dtype=np.uint32)
    lib.oxpt_sobol_directions(_ptr(out, u32))
    return out


def sheen_table() -> np.ndarray:
    out = np.zeros(512, dtype=np.float32)
    lib.oxpt_sheen_table(_ptr(out, f32))
    return out


def gpu_available() -> bool:
    """True when a HIP device can be opened (without keeping it)."""
    h = _vp()
    if lib.oxpt_ctx_create(0, C.byref(h)) != 0:
        return False
    lib.oxpt_ctx_destroy(h.value)
    return True


def build_wide_bvh(node_words: np.ndarray) -> Tuple[np.ndarray, int]:
    """4-wide nodes (float32, 32 per node) for binary node words in the
    reference layout (8 words per node); see oxpt_build_wide_bvh."""
    nw = np.ascontiguousarray(node_words, dtype=np.float32)
    n = nw.size // 8
    cnt, depth = sz(), u32()
    rc = lib.oxpt_build_wide_bvh(_ptr(nw, f32), n, None, 0, C.byref(cnt), C.byref(depth))
    if rc:
        _raise(rc, "build_wide_bvh")
    out = np.zeros(cnt.value * 4, dtype=np.float32)
    rc = lib.oxpt_build_wide_bvh(_ptr(nw, f32), n, _ptr(out, f32), cnt.value, C.byref(cnt), C.byref(depth))
    if rc:
        _raise(rc, "build_wide_bvh")
    return out, depth.value


def udiv_magic(d: int) -> Tuple[int, int]:
    """The claim step's multiply-shift pair (m, l) for divisor d (host only)."""
    m, l = u32(), u32()
    lib.oxpt_udiv_magic(d, C.byref(m), C.byref(l))
    return m.value, l.value
