"""draw_texture (src/pathtrace.rs:687-691): the tracer resolves into an image
it owns; the bytes equal the CPU oracle's resolve (resolve.wgsl:151-188) of
the same accumulator, and another process reads the same device memory
through the exported IPC handle (no copy through the host)."""
import subprocess
import sys

import numpy as np
import pytest

import oracle
import oxpt
from ptutil import oracle_frame, parity, resolve_params

pytestmark = pytest.mark.gpu


def _scene():
    return oxpt.Scene.terrain(seed=3, gx=96, gz=48, n_spheres=12, subdiv=2, n_emissive=64)


def test_draw_texture_equals_the_oracle_resolve(gpu):
    env = oxpt.EnvironmentMap.sky(128, 64, 4)
    opts = oxpt.RenderOptions(width=120, height=72, samples_per_pixel=8, max_bounces=8, tone_map="reinhard",
                              exposure=1.4, background=(30, 50, 80, 180))
    gpu.set_environment(env)
    gpu.sync(gpu_scene := _scene(), opts)
    gpu.refine(3)  # a partial accumulation: coverage and compositing matter
    acc = gpu.accum()
    for hdr in (False, True):
        img = gpu.draw_texture(hdr)
        again = gpu.draw_texture(hdr)
        assert again.h == img.h and again.ptr == img.ptr  # the tracer's one image, reused
        gpu.finish()
        got = img.read()
        want = oracle.resolve(acc, resolve_params(opts, 3, hdr))
        if hdr:
            assert np.allclose(got, want, rtol=1e-6, atol=1e-7)
        else:
            d = np.abs(got.astype(np.int32) - want.astype(np.int32))
            assert d.max() <= 1 and (d > 0).mean() < 1e-3, (d.max(), (d > 0).mean())
    # The converged draw against the oracle's own render at equal spp.
    gpu.refine(5)
    gh = gpu.draw_texture(True)
    gpu.finish()
    gh = gh.read()
    gi = gpu.draw_texture(False)
    gpu.finish()
    gi = gi.read()
    _, ci, ch = oracle_frame(gpu_scene, opts, env)
    parity("draw", gh, gi, ch, ci, 60.0)


CHILD = r"""
import ctypes as C, sys
hip = C.CDLL("libamdhip64.so.7")
class H(C.Structure):
    _fields_ = [("reserved", C.c_char * 64)]
hip.hipIpcOpenMemHandle.argtypes = [C.POINTER(C.c_void_p), H, C.c_uint]
hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
h = H()
C.memmove(C.byref(h), bytes.fromhex(sys.argv[1]), 64)
n = int(sys.argv[2])
assert hip.hipSetDevice(0) == 0
p = C.c_void_p()
rc = hip.hipIpcOpenMemHandle(C.byref(p), h, 1)
assert rc == 0, f"hipIpcOpenMemHandle: {rc}"
buf = C.create_string_buffer(n)
assert hip.hipMemcpy(buf, p, n, 2) == 0
hip.hipIpcCloseMemHandle(p)
sys.stdout.buffer.write(buf.raw)
"""


def test_draw_texture_export_is_read_by_another_process(gpu):
    opts = oxpt.RenderOptions(width=96, height=64, samples_per_pixel=4, max_bounces=6)
    gpu.set_environment(None)
    gpu.sync(_scene(), opts)
    gpu.refine(4)
    img = gpu.draw_texture(False)
    gpu.finish()
    local = img.read()
    handle = img.export()
    assert len(handle) == 64
    print("child process opens the exported image", flush=True)
    out = subprocess.run([sys.executable, "-c", CHILD, handle.hex(), str(img.nbytes)], capture_output=True,
                         timeout=120)
    assert out.returncode == 0, out.stderr.decode()[-2000:]
    remote = np.frombuffer(out.stdout, dtype=np.uint8).reshape(local.shape)
    assert np.array_equal(remote, local)
