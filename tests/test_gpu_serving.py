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
    # Budget for about one scene at a time: requests are admitted one by one,
    # idle scenes are evicted, and every request still completes.
    one = max(j.packed.geom.nbytes + j.packed.attrs.nbytes + j.packed.tables.nbytes for j in jobs)
    acc = max(j.opts.width * j.opts.height * 16 for j in jobs)
    srv = oxpt.Server(0, hbm_budget=int(one * 1.5 + acc * 1.2) + (8 << 20), spp_per_tick=8)
    ids = [srv.add_scene(j.scene, j.env) for j in jobs]
    reqs = [srv.submit_render(i, j.opts) for i, j in zip(ids, jobs)]
    for r in reqs:
        assert srv.wait(r, 300.0).ok
    st = srv.stats()
    assert st.evictions >= 1
    assert st.resident_bytes <= int(one * 1.5 + acc * 1.2) + (8 << 20)
    srv.close()


def test_request_larger_than_the_budget_waits_then_fails_cleanly():
    jobs = _pool(n=1, seed=23)
    srv = oxpt.Server(0, hbm_budget=1 << 16, spp_per_tick=4)
    i = srv.add_scene(jobs[0].scene, jobs[0].env)
    r = srv.submit_render(i, jobs[0].opts)
    with pytest.raises(oxpt.BackendError):
        srv.wait(r, 2.0)  # never admitted: times out
    srv.close()
