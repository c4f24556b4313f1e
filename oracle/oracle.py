"""ctypes binding of the CPU oracle (liboxpt_oracle.so). TEST INFRASTRUCTURE.

Only tests/, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg
import this module, and only as the reference the GPU results are compared
with or as the timed CPU baseline.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboxpt_oracle.so")

F64_LIGHTS = 1

u32, f32, u64, sz = C.c_uint32, C.c_float, C.c_uint64, C.c_size_t
P = C.POINTER


def build() -> None:
    """Compile the oracle with its Makefile (gcc)."""
    import subprocess

    subprocess.run(["make", "-s", "-C", HERE, "-B"], check=True)


class OScene(C.Structure):
    _fields_ = [
        ("geom", P(f32)), ("attrs", P(f32)), ("tables", P(u32)), ("atlas", P(C.c_uint16)),
        ("atlas_side", u32), ("atlas_layers", u32),
    ]


class Counts(C.Structure):
    _fields_ = [("paths", u64), ("rays", u64), ("nodes", u64), ("tris", u64), ("hits", u64), ("tex_taps", u64),
                ("bytes", u64)]


def _load():
    if not os.path.exists(LIB_PATH):
        build()
    lib = C.CDLL(LIB_PATH)

    def sig(name, res, *args):
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = list(args)

    vp = C.c_void_p
    sig("oxo_sobol_directions", None, P(u32))
    sig("oxo_sheen_table", None, P(f32), C.c_int)
    sig("oxo_pcg_hash", u32, u32)
    sig("oxo_hash_mix", u32, u32, u32)
    sig("oxo_nested_uniform_scramble", u32, u32, u32)
    sig("oxo_pattern_words", None, P(u32), u32, u32, u32, u32, u32, P(u32))
    sig("oxo_pattern_words_batch", None, P(u32), u32, sz, P(u32), P(u32), P(u32), P(u32), P(u32))
    sig("oxo_light_picks", None, P(OScene), vp, sz, P(f32), P(u32))
    sig("oxo_render", C.c_int, P(OScene), vp, u32, u32, u32, u32, u32, u32, P(f32), C.c_int, C.c_int, P(Counts))
    sig("oxo_trace_sample", C.c_int, P(OScene), vp, u32, u32, u32, C.c_int, P(f32))
    sig("oxo_closest_hit", None, P(OScene), vp, P(f32), P(f32), C.c_int, P(u32), P(f32))
    sig("oxo_resolve", None, P(f32), u32, u32, vp, vp)
    sig("oxo_f32_to_f16", C.c_uint16, f32)
    sig("oxo_bsdf_probe", None, P(u32), P(f32), P(f32), P(f32), C.c_int, P(f32), P(f32), P(f32))
    sig("oxo_solid_angle", f32, P(f32), P(f32), P(f32), P(f32), C.c_int)
    sig("oxo_sample_spherical_triangle", None, P(f32), P(f32), P(f32), P(f32), f32, f32, C.c_int, P(f32))
    sig("oxo_texel_index", None, f32, C.c_int32, u32, P(C.c_int32))
    sig("oxo_ndf", f32, C.c_int, f32, f32)
    return lib


lib = _load()


def _p(a, ct):
    return a.ctypes.data_as(P(ct))


def _v3(x):
    return (f32 * 3)(*[float(v) for v in x])


class Scene:
    """The oracle's view of packed words: geom, attrs, tables, atlas + Params."""

    def __init__(self, geom, attrs, tables, atlas, atlas_side, atlas_layers, params):
        self.geom = np.ascontiguousarray(geom, dtype=np.float32)
        self.attrs = np.ascontiguousarray(attrs, dtype=np.float32)
        self.tables = np.ascontiguousarray(tables, dtype=np.uint32)
        self.atlas = np.ascontiguousarray(atlas if atlas is not None and len(atlas) else np.zeros(4, np.uint16),
                                          dtype=np.uint16)
        self.params = params  # a ctypes Params (176 bytes)
        self.s = OScene(_p(self.geom, f32), _p(self.attrs, f32), _p(self.tables, u32), _p(self.atlas, C.c_uint16),
                        max(int(atlas_side), 1), max(int(atlas_layers), 1))

    @classmethod
    def from_packed(cls, pk) -> "Scene":
        return cls(pk.geom, pk.attrs, pk.tables, pk.atlas, pk.atlas_side, pk.atlas_layers, pk.params)

    @property
    def width(self):
        return self.params.dims[0]

    @property
    def height(self):
        return self.params.dims[1]

    def _pp(self):
        return C.cast(C.pointer(self.params), C.c_void_p)

    def render(self, first=0, count=1, region=None, accum=None, threads=0, flags=0, counts=False):
        """Trace samples [first, first + count) over region (x0, y0, x1, y1); returns the float4 accumulator."""
        w, h = self.width, self.height
        if accum is None:
            accum = np.zeros((h, w, 4), dtype=np.float32)
        x0, y0, x1, y1 = region if region is not None else (0, 0, w, h)
        cnt = Counts() if counts else None
        lib.oxo_render(C.byref(self.s), self._pp(), first, count, x0, y0, x1, y1, _p(accum, f32), threads, flags,
                       C.byref(cnt) if cnt is not None else None)
        return (accum, cnt) if counts else accum

    def trace_sample(self, x, y, index, flags=0):
        out = (f32 * 3)()
        cov = lib.oxo_trace_sample(C.byref(self.s), self._pp(), x, y, index, flags, out)
        return bool(cov), np.array(list(out), dtype=np.float32)

    def closest_hit(self, o, d, brute=False):
        slot, t = u32(), f32()
        lib.oxo_closest_hit(C.byref(self.s), self._pp(), _v3(o), _v3(d), 1 if brute else 0, C.byref(slot), C.byref(t))
        return slot.value, t.value

    def light_picks(self, u):
        u = np.ascontiguousarray(u, dtype=np.float32)
        out = np.zeros((u.size, 3), dtype=np.uint32)
        lib.oxo_light_picks(C.byref(self.s), self._pp(), u.size, _p(u, f32), _p(out, u32))
        return out


def sobol_directions() -> np.ndarray:
    out = np.zeros(128, dtype=np.uint32)
    lib.oxo_sobol_directions(_p(out, u32))
    return out


def sheen_table(n=256) -> np.ndarray:
    out = np.zeros(512, dtype=np.float32)
    lib.oxo_sheen_table(_p(out, f32), n)
    return out


def pattern_words(sobol, seed, x, y, index, p) -> np.ndarray:
    arrs = [np.ascontiguousarray(a, dtype=np.uint32) for a in (x, y, index, p)]
    sob = np.ascontiguousarray(sobol, dtype=np.uint32)
    n = arrs[0].size
    out = np.zeros((n, 4), dtype=np.uint32)
    lib.oxo_pattern_words_batch(_p(sob, u32), seed, n, *[_p(a, u32) for a in arrs], _p(out, u32))
    return out


def resolve(accum, rp) -> np.ndarray:
    """rp: a ctypes ResolveParams; RGBA8 (cfg[3] == 0) or RGBA32F."""
    accum = np.ascontiguousarray(accum, dtype=np.float32)
    h, w = accum.shape[:2]
    hdr = rp.cfg[3] == 1
    out = np.zeros((h, w, 4), dtype=np.float32 if hdr else np.uint8)
    lib.oxo_resolve(_p(accum, f32), w, h, C.cast(C.pointer(rp), C.c_void_p), out.ctypes.data_as(C.c_void_p))
    return out


def f32_to_f16(v: float) -> int:
    return lib.oxo_f32_to_f16(v)


def bsdf_probe(mat_words, lut, n, v, front, u, l):
    mw = np.ascontiguousarray(mat_words, dtype=np.uint32)
    lt = np.ascontiguousarray(lut, dtype=np.float32)
    out = (f32 * 13)()
    lib.oxo_bsdf_probe(_p(mw, u32), _p(lt, f32), _v3(n), _v3(v), 1 if front else 0, _v3(u), _v3(l), out)
    return np.array(list(out), dtype=np.float32)


def solid_angle(p, a, b, c, f64=False) -> float:
    return lib.oxo_solid_angle(_v3(p), _v3(a), _v3(b), _v3(c), 1 if f64 else 0)


def sample_spherical_triangle(p, a, b, c, u0, u1, f64=False) -> np.ndarray:
    out = (f32 * 3)()
    lib.oxo_sample_spherical_triangle(_v3(p), _v3(a), _v3(b), _v3(c), u0, u1, 1 if f64 else 0, out)
    return np.array(list(out), dtype=np.float32)


def texel_index(u: float, size: int, mode: int) -> tuple:
    """{nearest, bilinear x0, x0 + 1} texel indices of coordinate u, wrapped."""
    out = (C.c_int32 * 3)()
    lib.oxo_texel_index(float(u), int(size), int(mode), out)
    return tuple(out)


def ndf(kind: int, n_dot_h: float, a: float) -> float:
    """kind 0: GGX D; kind 1: Charlie D."""
    return lib.oxo_ndf(kind, float(n_dot_h), float(a))
