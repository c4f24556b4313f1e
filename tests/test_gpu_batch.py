"""Batched form: one persistent launch over scenes of uneven size equals each scene alone."""
import numpy as np
import pytest

import oxpt
from oxpt import workloads

pytestmark = pytest.mark.gpu


def _setup(n=12, scale=0.01):
    dev = oxpt.Device(0)
    jobs = workloads.uneven_batch(n=n, seed=5, threads=8, scale=scale)
    ds = [dev.upload(j.packed) for j in jobs]
    acc = [dev.accum(j.opts.width, j.opts.height) for j in jobs]
    return dev, jobs, ds, acc


def test_batch_equals_each_scene_alone():
    dev, jobs, ds, acc = _setup()
    alone = []
    for j, s, a in zip(jobs, ds, acc):
        dev.refine(s, a, s.params, 0, j.opts.samples_per_pixel)
        alone.append(a.read())
    for a in acc:
        a.clear()
    dev.refine_batch([(s, a, s.params, 0, j.opts.samples_per_pixel) for j, s, a in zip(jobs, ds, acc)])
    for i, a in enumerate(acc):
        got = a.read()
        assert np.array_equal(got.view(np.uint32), alone[i].view(np.uint32)), f"job {i} differs"
    # Rounds: a small path cap splits every item's samples over several
    # launches; the accumulators stay bit-identical.
    for a in acc:
        a.clear()
    dev.set_tuning(max_paths=20000)
    dev.refine_batch([(s, a, s.params, 0, j.opts.samples_per_pixel) for j, s, a in zip(jobs, ds, acc)])
    n_launch, _ = dev.timing()
    dev.set_tuning(max_paths=1 << 27)
    for i, a in enumerate(acc):
        assert np.array_equal(a.read().view(np.uint32), alone[i].view(np.uint32)), f"job {i} differs (rounds)"
    assert n_launch > len(jobs)


def test_batch_matches_plain_kernel_and_progressive_split():
    dev, jobs, ds, acc = _setup(n=6)
    dev.set_tuning(kernel=1)
    ref = []
    for j, s, a in zip(jobs, ds, acc):
        dev.refine(s, a, s.params, 0, j.opts.samples_per_pixel)
        ref.append(a.read())
        a.clear()
    dev.set_tuning(kernel=0)
    # Progressive: first half of the samples, then the rest, in two batches.
    half = [max(1, j.opts.samples_per_pixel // 2) for j in jobs]
    dev.refine_batch([(s, a, s.params, 0, h) for s, a, h in zip(ds, acc, half)])
    dev.refine_batch([(s, a, s.params, h, j.opts.samples_per_pixel - h)
                      for j, s, a, h in zip(jobs, ds, acc, half) if j.opts.samples_per_pixel > h])
    for i, a in enumerate(acc):
        assert np.array_equal(a.read().view(np.uint32), ref[i].view(np.uint32)), f"job {i}"


def test_batch_rejects_bad_items():
    dev, jobs, ds, acc = _setup(n=2)
    with pytest.raises(oxpt.Error):
        dev.refine_batch([(ds[0], acc[0], ds[0].params, 0, 1), (ds[1], acc[0], ds[1].params, 0, 1)])
    other = oxpt.Device(0)
    oacc = other.accum(jobs[0].opts.width, jobs[0].opts.height)
    with pytest.raises(oxpt.Error):
        dev.refine_batch([(ds[0], oacc, ds[0].params, 0, 1)])


def test_batch_forms_bitwise():
    """Every batched form (occupancy builds, the wavefront engine, the binary
    tree walk, graph replay) gives each job's image alone."""
    dev, jobs, ds, acc = _setup(n=8)
    alone = []
    for j, s, a in zip(jobs, ds, acc):
        dev.refine(s, a, s.params, 0, j.opts.samples_per_pixel)
        alone.append(a.read())
    items = [(s, a, s.params, 0, j.opts.samples_per_pixel) for j, s, a in zip(jobs, ds, acc)]
    for kernel, bvh, graph in ((3, 0, False), (4, 0, False), (8, 0, False), (7, 0, False), (0, 2, False),
                               (0, 0, True), (0, 0, True)):
        for a in acc:
            a.clear()
        dev.set_tuning(kernel=kernel, bvh=bvh)
        dev.set_graph(graph)
        dev.refine_batch(items)
        for i, a in enumerate(acc):
            assert np.array_equal(a.read().view(np.uint32), alone[i].view(np.uint32)), (kernel, bvh, graph, i)
    dev.set_graph(False)
    dev.set_tuning(kernel=0)


def test_batch_pixel_runs_and_tails_equal_alone():
    """A batched job is traced as whole-pixel runs (added straight into the
    accumulator) and a path-by-path tail (added by the reduce in sample
    order). Every split gives each job's image alone, bit for bit: all
    pixel runs (TAIL_NONE), the default (small jobs path by path), quarter
    and whole-frame tails, and with rounds of a few samples."""
    dev, jobs, ds, acc = _setup(n=10, scale=0.02)
    items = [(s, a, s.params, 0, j.opts.samples_per_pixel) for j, s, a in zip(jobs, ds, acc)]
    alone = []
    for j, s, a in zip(jobs, ds, acc):
        a.clear()
        dev.refine(s, a, s.params, 0, j.opts.samples_per_pixel)
        alone.append(a.read().view(np.uint32).copy())
    for tail, max_paths in ((oxpt.TAIL_NONE, 1 << 30), (0, 1 << 30), (4, 1 << 30), (16, 1 << 30),
                            (oxpt.TAIL_NONE, 60000), (4, 60000)):
        dev.set_tuning(tail=tail, max_paths=max_paths)
        for a in acc:
            a.clear()
        dev.refine_batch(items)
        for i, a in enumerate(acc):
            assert np.array_equal(a.read().view(np.uint32), alone[i]), (tail, max_paths, i)
    dev.set_tuning(max_paths=1 << 30)
