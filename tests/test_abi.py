"""The C ABI loads and exports every symbol include/oxpt/*.h declares (no GPU calls)."""
import ctypes as C
import os
import re

import numpy as np

import oxpt

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in ("oxpt.h",):
        text = open(os.path.join(ROOT, "include", "oxpt", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^[A-Za-z_][\w \*]*?\b(oxpt_\w+)\s*\(", text, flags=re.M):
            if not m.group(0).lstrip().startswith(("typedef", "#")):
                names.add(m.group(1))
    return sorted(names)


def test_every_declared_symbol_is_exported():
    names = declared_functions()
    assert len(names) > 60
    missing = [n for n in names if not hasattr(oxpt.lib, n)]
    assert not missing, missing


def test_struct_sizes_match_the_header():
    assert oxpt.abi_sizes() == [C.sizeof(s) for s in oxpt.ABI_STRUCTS]
    assert C.sizeof(oxpt.Params) == 176 and C.sizeof(oxpt.Frame) == 16


def test_registry_rejects_unknown_names_without_a_device():
    h = C.c_void_p()
    assert oxpt.lib.oxpt_registry_make(b"gpu", C.byref(h)) == oxpt.E_INVALID
    assert not h.value


def test_device_calls_reject_null_handles():
    assert oxpt.lib.oxpt_refine(None, None, None, None, 0, 1) == oxpt.E_ARG
    assert oxpt.lib.oxpt_finish(None) == oxpt.E_ARG
    assert oxpt.lib.oxpt_tracer_refine(None, 1) == oxpt.E_ARG


def test_missing_device_is_a_backend_error():
    if oxpt.gpu_available():
        import pytest

        pytest.skip("a HIP device is present")
    import pytest

    with pytest.raises(oxpt.BackendError):
        oxpt.GpuPathTracer(0)


def test_build_wide_bvh_matches_the_packed_scene():
    """oxpt_build_wide_bvh over the binary node words reproduces the 4-wide
    nodes the packer appends (for callers that pack geom themselves)."""
    scene = oxpt.Scene.terrain(seed=4, gx=40, gz=20, n_spheres=6, subdiv=2, n_emissive=8)
    pk = oxpt.Packed(scene, oxpt.RenderOptions(width=8, height=8))
    P = pk.params
    nodes = pk.geom[: 8 * P.env[3]]
    wide, depth = oxpt.build_wide_bvh(nodes)
    assert depth >= 1
    assert np.array_equal(wide.view(np.uint32), pk.geom[4 * P.offs2[3]: 4 * P.offs2[3] + wide.size].view(np.uint32))
    assert 4 * P.offs2[3] + wide.size == pk.geom.size
