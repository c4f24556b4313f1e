"""oxpt_refine_batch_noalloc (the batched refine without device allocation)
against oxpt_refine_batch, through the C ABI: bitwise equal accumulators on
two calls in a row (the second smaller), with a scene carrying punctual
lights; oversize requests are refused with an error."""
import ctypes as C

import numpy as np
import pytest

import oxpt
from oxpt import workloads

pytestmark = pytest.mark.gpu

lib = oxpt.lib
lib.oxpt_refine_batch_noalloc.restype = C.c_int
lib.oxpt_refine_batch_noalloc.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t]


def _items(entries):
    arr = (oxpt.BatchItem * len(entries))()
    for i, (sc, acc, prm, first, count) in enumerate(entries):
        arr[i] = oxpt.BatchItem(sc.h, acc.h, prm, first, count)
    return arr


def _noalloc(dev, entries):
    arr = _items(entries)
    rc = lib.oxpt_refine_batch_noalloc(dev.h, arr, len(entries))
    if rc:
        oxpt._raise(rc, lib.oxpt_last_error(dev.h).decode())


def test_noalloc_batch_equals_batch_on_two_calls():
    dev = oxpt.Device(0)
    jobs = workloads.uneven_batch(n=5, seed=41, threads=8, scale=0.02)
    lights = oxpt.Scene.layered(seed=9, n_tris=6000, n_textures=1, tex_size=32, n_lights=3)
    lopts = oxpt.RenderOptions(width=72, height=40, samples_per_pixel=6, max_bounces=8)
    packs = [j.packed for j in jobs] + [oxpt.Packed(lights, lopts)]
    spps = [j.opts.samples_per_pixel for j in jobs] + [6]
    ds = [dev.upload(pk) for pk in packs]
    sizes = [(pk.params.dims[0], pk.params.dims[1]) for pk in packs]
    for call, keep in enumerate((len(ds), 3)):  # second call: fewer, smaller items
        sel = list(range(len(ds)))[-keep:] if call else list(range(keep))
        ref = [dev.accum(*sizes[i]) for i in sel]
        got = [dev.accum(*sizes[i]) for i in sel]
        counts = [max(1, spps[i] // (1 + call)) for i in sel]
        dev.refine_batch([(ds[i], a, ds[i].params, 0, n) for i, a, n in zip(sel, ref, counts)])
        _noalloc(dev, [(ds[i], a, ds[i].params, 0, n) for i, a, n in zip(sel, got, counts)])
        dev.finish()
        for i, a, b in zip(sel, ref, got):
            assert np.array_equal(a.read().view(np.uint32), b.read().view(np.uint32)), (call, i)
    # Progressive: continuing the same accumulators from sample n equals one run.
    a = dev.accum(*sizes[0])
    b = dev.accum(*sizes[0])
    dev.refine_batch([(ds[0], a, ds[0].params, 0, 5)])
    _noalloc(dev, [(ds[0], b, ds[0].params, 0, 2)])
    _noalloc(dev, [(ds[0], b, ds[0].params, 2, 3)])
    assert np.array_equal(a.read().view(np.uint32), b.read().view(np.uint32))


def test_noalloc_batch_refuses_what_its_scratch_cannot_hold():
    dev = oxpt.Device(0)
    pk = oxpt.Packed(oxpt.Scene.room(0), oxpt.RenderOptions(width=2056, height=2048, samples_per_pixel=1))
    s = dev.upload(pk)
    big = dev.accum(2056, 2048)
    with pytest.raises(oxpt.Error, match="4194304 paths"):
        _noalloc(dev, [(s, big, s.params, 0, 1)])
    small = oxpt.Packed(oxpt.Scene.room(0), oxpt.RenderOptions(width=8, height=8, samples_per_pixel=1))
    ss = dev.upload(small)
    accs = [dev.accum(8, 8) for _ in range(257)]
    with pytest.raises(oxpt.Error, match="256 items"):
        _noalloc(dev, [(ss, a, ss.params, 0, 1) for a in accs])
    # Exactly at the limit still runs.
    fit = oxpt.Packed(oxpt.Scene.room(0), oxpt.RenderOptions(width=2048, height=2048, samples_per_pixel=1))
    fs = dev.upload(fit)
    fa, fb = dev.accum(2048, 2048), dev.accum(2048, 2048)
    _noalloc(dev, [(fs, fa, fs.params, 0, 2)])
    dev.refine_batch([(fs, fb, fs.params, 0, 2)])
    dev.finish()
    assert np.array_equal(fa.read().view(np.uint32), fb.read().view(np.uint32))
