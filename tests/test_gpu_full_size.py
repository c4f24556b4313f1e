"""The uneven-batch and serving workloads at full size, on the GPU.

* Uneven batch: the full 64-scene workload (triangle counts log-uniform over
  1 k..2 M, sides 128..1024, spp in {4, 16, 64, 256}, depth 12) in one batched
  refine equals each scene refined alone, bit for bit; the smallest, median
  and largest jobs match the CPU oracle on seeded 64 x 64 tiles at full spp
  with the reference's parity method (tests/pathtrace.rs:88-142).
* Serving: about twenty full-size requests drawn from the same distribution,
  with refine(4) on some, arriving while others run: each served image
  equals the image rendered alone; one matches the oracle on tiles.
"""
import threading
import time

import numpy as np
import pytest

import oracle
import oxpt
from oxpt import workloads
from ptutil import parity, resolve_params

pytestmark = pytest.mark.gpu


def _tiles(w, h, n, seed):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        x0 = int(rng.integers(0, max(w - 64, 0) + 1))
        y0 = int(rng.integers(0, max(h - 64, 0) + 1))
        out.append((x0, y0, min(x0 + 64, w), min(y0 + 64, h)))
    return out


def _oracle_tiles(job, gpu_acc, spp, name, seed):
    """GPU accumulator vs the oracle on two tiles at `spp` samples."""
    w, h = job.opts.width, job.opts.height
    osc = oracle.Scene.from_packed(job.packed)
    cacc = np.zeros((h, w, 4), np.float32)
    tl = _tiles(w, h, 2, seed)
    for r in tl:
        osc.render(0, spp, region=r, accum=cacc)
    rl, rh = resolve_params(job.opts, spp, False), resolve_params(job.opts, spp, True)
    g_img, g_hdr = oracle.resolve(gpu_acc, rl), oracle.resolve(gpu_acc, rh)
    c_img, c_hdr = oracle.resolve(cacc, rl), oracle.resolve(cacc, rh)

    def cat(a):
        return np.concatenate([a[r[1]:r[3], r[0]:r[2]] for r in tl], axis=1)

    return parity(name, cat(g_hdr), cat(g_img), cat(c_hdr), cat(c_img), 60.0)


@pytest.fixture(scope="module")
def uneven():
    t = time.perf_counter()
    jobs = workloads.uneven_batch(n=64, seed=3, threads=16)
    print(f"64 scenes packed in {time.perf_counter() - t:.1f} s, "
          f"{sum(j.paths for j in jobs) / 1e9:.2f} G paths", flush=True)
    return jobs


def test_full_uneven_batch_equals_alone_and_oracle(uneven):
    jobs = uneven
    dev = oxpt.Device(0)
    ds = [dev.upload(j.packed) for j in jobs]
    acc = [dev.accum(j.opts.width, j.opts.height) for j in jobs]
    dev.refine_batch([(s, a, s.params, 0, j.opts.samples_per_pixel) for j, s, a in zip(jobs, ds, acc)])
    dev.finish()
    batched = [a.read() for a in acc]
    for i, (j, s, a) in enumerate(zip(jobs, ds, acc)):
        a.clear()
        dev.refine(s, a, s.params, 0, j.opts.samples_per_pixel)
        alone = a.read()
        assert np.array_equal(alone.view(np.uint32), batched[i].view(np.uint32)), f"job {i} differs"
    order = sorted(range(len(jobs)), key=lambda i: jobs[i].paths)
    for name, i in (("smallest", order[0]), ("median", order[len(order) // 2]), ("largest", order[-1])):
        j = jobs[i]
        p = _oracle_tiles(j, batched[i], j.opts.samples_per_pixel, f"uneven-{name}", 100 + i)
        print(f"job {i} ({name}: {j.n_tris} tris, {j.opts.width}^2, {j.opts.samples_per_pixel} spp): "
              f"PSNR {p:.2f} dB on tiles", flush=True)


def test_full_size_serving_equals_alone(uneven):
    pool = [uneven[k] for k in range(0, 64, 8)]  # 8 full-size scenes of the workload
    srv = oxpt.Server(0, keep_hdr=True, tick_ms=8, spp_per_tick=0)
    ids = [srv.add_scene(j.scene, j.env) for j in pool]
    rng = np.random.default_rng(17)
    picks = [int(rng.integers(0, len(pool))) for _ in range(20)]
    renders, refines = [], []

    def submit():
        for k, p in enumerate(picks):
            renders.append(srv.submit_render(ids[p], pool[p].opts))
            if k % 4 == 1:
                refines.append((k, srv.submit_refine(renders[k], 4)))
            time.sleep(float(rng.exponential(0.02)))

    th = threading.Thread(target=submit)
    th.start()
    th.join()
    for r in renders + [r for _, r in refines]:
        assert srv.wait(r, 600.0).ok
    st = srv.stats()
    print(f"served {st.completed} requests in {st.ticks} ticks, busy {st.busy_s:.2f} s, io {st.io_s:.2f} s",
          flush=True)
    refined = dict(refines)
    dev = oxpt.Device(0)
    uploaded = {}
    for k, p in enumerate(picks):
        j = pool[p]
        if p not in uploaded:
            uploaded[p] = dev.upload(j.packed)
        s = uploaded[p]
        a = dev.accum(j.opts.width, j.opts.height)
        spp = j.opts.samples_per_pixel
        dev.refine(s, a, s.params, 0, spp)
        alone = a.resolve(resolve_params(j.opts, spp, True))
        assert np.array_equal(srv.result(renders[k], hdr=True).view(np.uint32), alone.view(np.uint32)), k
        if k in refined:
            dev.refine(s, a, s.params, spp, 4)
            more = a.resolve(resolve_params(j.opts, spp + 4, True))
            assert np.array_equal(srv.result(refined[k], hdr=True).view(np.uint32), more.view(np.uint32)), k
        if k == 0:  # not refined (refines follow k = 1, 5, 9, ...): served == alone, checked above
            p_db = _oracle_tiles(j, a.read(), spp, "served", 7)
            print(f"served request 0: PSNR {p_db:.2f} dB on tiles", flush=True)
    srv.close()
