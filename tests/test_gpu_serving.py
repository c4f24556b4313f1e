"""Serving: continuous batching keeps each image equal to the image rendered alone."""
import threading
import time

import numpy as np
import pytest

import oxpt
from oxpt import workloads

pytestmark = pytest.mark.gpu


def _pool(n=6, scale=0.01, seed=21):
    return workloads.uneven_batch(n=n, seed=seed, threads=8, scale=scale)


@pytest.mark.parametrize("use_graph,device_build", [(False, False), (True, False), (False, True)],
                         ids=["eager", "graph", "device_build"])
def test_served_images_equal_images_alone(use_graph, device_build):
    """device_build: pool scenes are kept as host scenes and prepared on the
    device at first use (BVH, words, lights, environment tables)."""
    jobs = _pool()
    srv = oxpt.Server(0, spp_per_tick=4, keep_hdr=True, use_graph=use_graph, device_build=device_build)
    ids = [srv.add_scene(j.scene, j.env) for j in jobs]
    reqs, refines = [], []

    # Staggered arrivals from another thread, so requests join running ticks.
    def submit():
        rng = np.random.default_rng(1)
        for k, (i, j) in enumerate(zip(ids, jobs)):
            reqs.append(srv.submit_render(i, j.opts))
            time.sleep(float(rng.uniform(0.0, 0.02)))
        for k in (0, 2, 4):
            refines.append((k, srv.submit_refine(reqs[k], 4)))

    t = threading.Thread(target=submit)
    t.start()
    t.join()
    infos = [srv.wait(r, 300.0) for r in reqs]
    rinfo = [srv.wait(r, 300.0) for _, r in refines]
    assert all(i.ok for i in infos) and all(i.ok for i in rinfo)
    st = srv.stats()
    assert st.completed == len(reqs) + len(refines) and st.ticks > 1
    if use_graph:
        assert srv.graph_launches() >= st.ticks
    refined = {k: srv.result(r, hdr=True) for k, r in refines}
    served = [srv.result(r, hdr=True) for r in reqs]
    # Refines waited for their render; their image is render + 4 samples.
    pt = oxpt.GpuPathTracer(0)
    for k, j in enumerate(jobs):
        pt.set_environment(j.env)
        alone = pt.render_hdr(j.scene, j.opts)
        if k in refined:
            # The render's own image was resolved before its refine ran.
            assert np.array_equal(served[k].view(np.uint32), alone.view(np.uint32)), f"render {k}"
            pt.refine(4)
            more = pt.hdr()
            assert np.array_equal(refined[k].view(np.uint32), more.view(np.uint32)), f"refine {k}"
        else:
            assert np.array_equal(served[k].view(np.uint32), alone.view(np.uint32)), f"render {k}"
    srv.close()


def test_admission_respects_the_budget_and_evicts():
    jobs = _pool(n=4, seed=22)
    # Budget for about one scene at a time (plus the launch scratch the
    # server reserves up front): requests are admitted one by one, idle
    # scenes are evicted, and every request still completes.
    one = max(j.packed.geom.nbytes + j.packed.attrs.nbytes + j.packed.tables.nbytes for j in jobs)
    px = max(j.opts.width * j.opts.height for j in jobs)
    probe = oxpt.Server(0, tick_paths=1 << 20)
    reserve = probe.stats().resident_bytes
    probe.close()
    budget = reserve + int(one * 1.5 + px * 20 * 1.2) + (8 << 20)
    srv = oxpt.Server(0, hbm_budget=budget, spp_per_tick=8, tick_paths=1 << 20)
    ids = [srv.add_scene(j.scene, j.env) for j in jobs]
    reqs = [srv.submit_render(i, j.opts) for i, j in zip(ids, jobs)]
    for r in reqs:
        assert srv.wait(r, 300.0).ok
    st = srv.stats()
    assert st.evictions >= 1
    assert st.resident_bytes <= budget
    srv.close()


def test_request_larger_than_the_budget_is_rejected():
    """A request that cannot fit even in an empty server fails at once
    instead of blocking the queue behind it."""
    jobs = _pool(n=2, seed=23)
    srv = oxpt.Server(0, hbm_budget=1 << 16, spp_per_tick=4)
    i = srv.add_scene(jobs[0].scene, jobs[0].env)
    r = srv.submit_render(i, jobs[0].opts)
    with pytest.raises(oxpt.BackendError):
        srv.wait(r, 30.0)
    assert srv.stats().rejected == 1
    srv.close()


def test_failed_tick_gives_back_its_memory():
    """A tick that fails (forced here before its launch) fails every request
    in it and frees their accumulators; the budget returns to what a server
    that never failed holds, and later requests on the same scene complete
    with the image rendered alone."""
    jobs = _pool(n=1, seed=24)
    j = jobs[0]
    opts = oxpt.RenderOptions(width=j.opts.width, height=j.opts.height, samples_per_pixel=12, max_bounces=12,
                              seed=j.opts.seed)
    resident = []
    for fail in (2, 0):
        srv = oxpt.Server(0, spp_per_tick=4, keep_hdr=True, fail_tick=fail)
        i = srv.add_scene(j.scene, j.env)
        if fail:
            r1 = srv.submit_render(i, opts)
            with pytest.raises(oxpt.BackendError, match="injected tick failure"):
                srv.wait(r1, 120.0)
            st = srv.stats()
            assert st.failed_ticks == 1 and st.rejected == 1 and st.active == 0
        r2 = srv.submit_render(i, opts)
        assert srv.wait(r2, 120.0).ok
        img = srv.result(r2, hdr=True)
        resident.append(srv.stats().resident_bytes)
        srv.close()
    assert resident[0] == resident[1], resident
    pt = oxpt.GpuPathTracer(0)
    pt.set_environment(j.env)
    alone = pt.render_hdr(j.scene, opts)
    assert np.array_equal(img.view(np.uint32), alone.view(np.uint32))


def test_adaptive_ticks_keep_images():
    """tick_ms > 0: ticks sized by the measured rate, shortest remaining work
    first, uneven samples per request per tick — images still equal alone."""
    jobs = _pool(n=5, seed=25)
    srv = oxpt.Server(0, keep_hdr=True, tick_ms=2, spp_per_tick=0)
    ids = [srv.add_scene(j.scene, j.env) for j in jobs]
    reqs = [srv.submit_render(i, j.opts) for i, j in zip(ids, jobs)]
    ref = srv.submit_refine(reqs[1], 3)
    for r in reqs + [ref]:
        assert srv.wait(r, 300.0).ok
    assert srv.stats().ticks >= 2
    pt = oxpt.GpuPathTracer(0)
    for k, (j, r) in enumerate(zip(jobs, reqs)):
        pt.set_environment(j.env)
        alone = pt.render_hdr(j.scene, j.opts)
        assert np.array_equal(srv.result(r, hdr=True).view(np.uint32), alone.view(np.uint32)), k
        if k == 1:
            pt.refine(3)
            assert np.array_equal(srv.result(ref, hdr=True).view(np.uint32), pt.hdr().view(np.uint32))
    srv.close()
