// Serving loop: continuous batching of render and progressive-refine
// requests over the batched launch (oxpt_refine_batch).
//
// * A pool of prepared scenes (host words, or host scenes built on the
//   device) is registered up front; a render request names a pool scene and
//   render options. Scenes are uploaded on first use and stay resident (BVH,
//   tables, atlas) with their accumulators, so a later refine continues the
//   same accumulation — the reference's progressive contract
//   (src/pathtrace.rs:578-583).
// * Admission is by an HBM budget that counts scenes, accumulators, readback
//   staging and the launch scratch reserved up front: a request is admitted
//   when its footprint fits; otherwise unused staging and least-recently-used
//   idle scenes (with their finished accumulators) are evicted, and if that
//   is not enough the request waits (FIFO). A request that cannot fit even in
//   an empty server is rejected at once.
// * Two threads, two streams. The scheduler thread forms one batched launch
//   per tick on the main context; the I/O thread uploads scenes, creates
//   accumulators and reads results back on a peer context, overlapping the
//   running launch. A request that finishes in tick k has its resolve queued
//   on the main stream right after tick k (before tick k + 1's launch), and
//   the I/O thread copies the resolved bytes out while tick k + 1 runs.
// * Tick sizing: with tick_ms > 0 a tick traces about the paths the GPU does
//   in tick_ms (measured rate), requests closest to done first, each active
//   request at least one sample; otherwise spp_per_tick samples each. Each
//   pixel's samples are still added in order, so every image equals its
//   image rendered alone.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <optional>
#include <string>
#include <thread>
#include <vector>

#include "scene.hpp"

namespace {

using Clock = std::chrono::steady_clock;

double now_s() { return std::chrono::duration<double>(Clock::now().time_since_epoch()).count(); }

// Launch scratch held outside the per-request footprints: the NEE queue of
// the early form with punctual lights (lanes x (3 x 31 + 12) words at most)
// and the batched job table.
constexpr uint64_t kQueueReserve = 256ull << 20;
constexpr uint64_t kDefaultTickPaths = 1ull << 27;

struct PoolScene {
    oxpt::Packed pk;
    std::optional<oxpt::EnvMap> env;
    oxpt::CameraSource cam;
    std::vector<uint32_t> tables;  // with the environment appended (host-packed scenes)
    // Device-prepared scenes: the host scene, built on the device at each
    // upload; upload_bytes is an upper bound until the first build.
    std::optional<oxpt::HostScene> host;
    uint64_t upload_bytes = 0;
    oxpt_scene *dev = nullptr;
    bool loading = false;  // an upload is queued or running on the I/O thread
    uint64_t last_use = 0;
    int active = 0;        // requests admitted on this scene and not yet done
};

enum class State { Queued, Loading, Active, Resolving, Done, Rejected };

struct Staging {
    void *dev = nullptr;
    uint64_t bytes = 0;
    bool busy = false;
};

struct Req {
    uint64_t id = 0;
    bool refine = false;
    uint32_t scene = 0;
    uint64_t parent = 0;  // render request a refine continues
    oxpt_render_options opts{};
    uint32_t spp = 0;
    oxpt_params P{};
    oxpt_accum *acc = nullptr;  // owned by the render request
    uint64_t acc_bytes = 0;
    Staging *staging = nullptr;  // resolve target, held from admission to readback
    uint32_t first = 0, remaining = 0, samples_total = 0;
    uint64_t pix64 = 0;  // 8x8-tile padded pixels: paths per sample in a launch
    double t_arrival = 0, t_admit = 0, t_done = 0;
    State state = State::Queued;
    std::string error;
    std::vector<uint8_t> image;  // RGBA8 once done
    std::vector<float> hdr;      // RGBA32F once done
};

struct IoJob {
    bool readback = false;
    std::vector<Req *> reqs;  // readback: the requests resolved before `mark`
    void *mark = nullptr;
    Req *load = nullptr;      // upload / accumulator creation for one request
};

uint32_t req_w(const Req &r) { return std::max(r.opts.width, 1u); }
uint32_t req_h(const Req &r) { return std::max(r.opts.height, 1u); }

}  // namespace

struct oxpt_server {
    oxpt_ctx *ctx = nullptr;  // launches and resolves
    oxpt_ctx *io = nullptr;   // uploads and readbacks (peer of ctx)
    oxpt_server_config cfg{};
    std::string err;
    std::mutex mu;
    std::condition_variable cv_work, cv_done, cv_io;
    std::thread worker, io_thread;
    bool stop = false;
    std::vector<std::unique_ptr<PoolScene>> pool;
    std::map<uint64_t, std::unique_ptr<Req>> reqs;
    std::deque<uint64_t> queue;   // admission order
    std::vector<uint64_t> active; // Loading or Active requests
    std::deque<IoJob> io_jobs;
    std::vector<std::unique_ptr<Staging>> staging;
    uint64_t next_id = 1, tick_no = 0, use_clock = 0;
    uint64_t change_gen = 0;  // bumped by every submission and completion
    uint64_t resident = 0;    // footprints of admitted work + the launch reserve
    uint64_t reserve = 0;
    double rate = 6.0e8;      // measured paths/s (tick sizing)
    oxpt_server_stats stats{};

    void run();
    void run_io();
    bool admit_locked(Req &r);
    bool evict_for_locked(uint64_t need);
    Staging *take_staging_locked(uint64_t bytes);
    void release_locked(Req &r, bool drop_acc);
    uint64_t staging_bytes(const Req &r) const {
        return (uint64_t)req_w(r) * req_h(r) * (cfg.keep_hdr ? 20u : 4u);
    }
    void reject_locked(Req &r, const std::string &why) {
        r.state = State::Rejected;
        r.error = why;
        r.t_done = now_s();
        stats.rejected++;
    }
};

Staging *oxpt_server::take_staging_locked(uint64_t bytes) {
    Staging *best = nullptr;
    for (auto &s : staging)
        if (!s->busy && s->bytes >= bytes && (!best || s->bytes < best->bytes)) best = s.get();
    if (best) {
        best->busy = true;
        return best;
    }
    auto s = std::make_unique<Staging>();
    if (oxpt::device_alloc(ctx, bytes, &s->dev) != OXPT_OK) return nullptr;
    s->bytes = bytes;
    s->busy = true;
    resident += bytes;
    staging.push_back(std::move(s));
    return staging.back().get();
}

// Give back what a request holds: its staging buffer and, with drop_acc, its
// accumulator (a render's own, or a refine's parent's) and the scene slot.
void oxpt_server::release_locked(Req &r, bool drop_acc) {
    if (r.staging) {
        r.staging->busy = false;
        r.staging = nullptr;
    }
    if (!drop_acc) return;
    Req *owner = r.refine ? (reqs.count(r.parent) ? reqs[r.parent].get() : nullptr) : &r;
    if (owner && owner->acc) {
        oxpt_accum_destroy(owner->acc);
        owner->acc = nullptr;
        resident -= std::min(resident, owner->acc_bytes);
    }
}

bool oxpt_server::evict_for_locked(uint64_t need) {
    // Unused staging first, then least-recently-used idle scenes with every
    // finished accumulator on them.
    for (auto it = staging.begin(); it != staging.end() && resident + need > cfg.hbm_budget;) {
        if ((*it)->busy) {
            ++it;
            continue;
        }
        oxpt::device_free(ctx, (*it)->dev, (*it)->bytes);
        resident -= std::min(resident, (*it)->bytes);
        it = staging.erase(it);
    }
    while (resident + need > cfg.hbm_budget) {
        PoolScene *victim = nullptr;
        uint32_t vi = 0;
        for (uint32_t i = 0; i < pool.size(); i++) {
            PoolScene &ps = *pool[i];
            if (ps.dev && !ps.loading && ps.active == 0 && (!victim || ps.last_use < victim->last_use)) {
                victim = &ps;
                vi = i;
            }
        }
        if (!victim) return false;
        for (auto &kv : reqs) {
            Req &r = *kv.second;
            if (r.scene == vi && !r.refine && r.acc && r.state == State::Done) {
                oxpt_accum_destroy(r.acc);
                resident -= std::min(resident, r.acc_bytes);
                r.acc = nullptr;
                stats.evicted_accums++;
            }
        }
        oxpt_scene_destroy(victim->dev);
        victim->dev = nullptr;
        resident -= std::min(resident, victim->upload_bytes);
        stats.evictions++;
    }
    return true;
}

// Decide admission (no device work but staging and evictions); the upload
// and accumulator creation go to the I/O thread. False: keep waiting.
bool oxpt_server::admit_locked(Req &r) {
    PoolScene &ps = *pool[r.scene];
    const uint64_t acc_b = r.refine ? 0 : (uint64_t)req_w(r) * req_h(r) * 16;
    const uint64_t stage_b = staging_bytes(r);
    const uint64_t scene_b = (ps.dev || ps.loading) ? 0 : ps.upload_bytes;
    const uint64_t need = acc_b + stage_b + scene_b;
    r.pix64 = 64ull * ((req_w(r) + 7u) / 8u) * ((req_h(r) + 7u) / 8u);
    if (need + reserve > cfg.hbm_budget) {
        reject_locked(r, "request larger than the server's HBM budget");
        return true;
    }
    if (r.pix64 > cfg.tick_paths) {
        reject_locked(r, "frame larger than one tick's path cap (tick_paths)");
        return true;
    }
    if (resident + need > cfg.hbm_budget && !evict_for_locked(need)) return false;
    r.staging = take_staging_locked(stage_b);
    if (!r.staging) {
        reject_locked(r, std::string("staging: ") + oxpt_last_error(ctx));
        return true;
    }
    if (r.refine) {
        Req &p = *reqs[r.parent];
        r.acc = p.acc;
        r.P = p.P;
        r.opts = p.opts;
        r.first = p.samples_total;
        r.remaining = r.spp;
        r.samples_total = p.samples_total + r.spp;
        p.samples_total = r.samples_total;
    } else {
        r.acc_bytes = acc_b;
        r.first = 0;
        r.remaining = std::max(r.opts.samples_per_pixel, 1u);
        r.samples_total = r.remaining;
    }
    resident += acc_b + scene_b;
    if (!ps.dev && !ps.loading) {
        ps.loading = true;
        stats.uploads++;
    }
    ps.active++;
    ps.last_use = ++use_clock;
    r.t_admit = now_s();
    active.push_back(r.id);
    if (r.refine && ps.dev) {
        r.state = State::Active;  // nothing to load
    } else {
        r.state = State::Loading;
        IoJob j;
        j.load = &r;
        io_jobs.push_back(std::move(j));
        cv_io.notify_one();
    }
    return true;
}

void oxpt_server::run_io() {
    for (;;) {
        IoJob job;
        {
            std::unique_lock<std::mutex> lk(mu);
            cv_io.wait(lk, [&] { return stop || !io_jobs.empty(); });
            if (stop) return;
            job = std::move(io_jobs.front());
            io_jobs.pop_front();
        }
        const double t0 = now_s();
        if (job.readback) {
            int rc = oxpt::mark_wait(io, job.mark);
            std::vector<int> rcs(job.reqs.size(), rc);
            for (size_t k = 0; k < job.reqs.size(); k++) {
                Req &r = *job.reqs[k];
                if (rcs[k] != OXPT_OK || r.state != State::Resolving) continue;
                const size_t px = (size_t)req_w(r) * req_h(r);
                r.image.resize(px * 4);
                rcs[k] = oxpt::copy_to_host(io, r.image.data(), r.staging->dev, px * 4);
                if (rcs[k] == OXPT_OK && cfg.keep_hdr) {
                    r.hdr.resize(px * 4);
                    rcs[k] = oxpt::copy_to_host(io, r.hdr.data(), (const uint8_t *)r.staging->dev + px * 4, px * 16);
                }
            }
            std::lock_guard<std::mutex> lk(mu);
            for (size_t k = 0; k < job.reqs.size(); k++) {
                Req &r = *job.reqs[k];
                if (r.state != State::Resolving) continue;
                PoolScene &ps = *pool[r.scene];
                ps.active--;
                release_locked(r, false);
                if (rcs[k] != OXPT_OK) {
                    reject_locked(r, oxpt_last_error(io));
                    continue;
                }
                stats.completed++;
                r.state = State::Done;
                r.t_done = now_s();
                stats.paths += (uint64_t)req_w(r) * req_h(r) * (r.refine ? r.spp : r.samples_total);
            }
            stats.io_s += now_s() - t0;
            change_gen++;
            cv_done.notify_all();
            cv_work.notify_all();
            continue;
        }
        // Load: the scene (first request on it) and a render's accumulator.
        Req &r = *job.load;
        PoolScene *ps;
        bool need_scene;
        {
            std::lock_guard<std::mutex> lk(mu);
            ps = pool[r.scene].get();
            need_scene = !ps->dev;
        }
        int rc = OXPT_OK;
        oxpt_scene *dev = nullptr;
        if (need_scene) {
            if (ps->host) {
                rc = oxpt::device_prepare(io, *ps->host, ps->pk, &dev, nullptr);
                if (rc == OXPT_OK) {
                    rc = oxpt::upload_tables(io, dev, ps->pk.tables, ps->env ? &*ps->env : nullptr);
                    if (rc != OXPT_OK) {
                        oxpt_scene_destroy(dev);
                        dev = nullptr;
                    }
                }
            } else {
                oxpt_scene_desc d{};
                d.geom = ps->pk.geom.data();
                d.geom_vec4 = ps->pk.geom.size() / 4;
                d.attrs = ps->pk.attrs.data();
                d.attrs_vec4 = ps->pk.attrs.size() / 4;
                d.tables = ps->tables.data();
                d.table_words = ps->tables.size();
                d.atlas = ps->pk.atlas.empty() ? nullptr : ps->pk.atlas.data();
                d.atlas_side = ps->pk.atlas_side;
                d.atlas_layers = ps->pk.atlas_layers;
                rc = oxpt_scene_upload(io, &d, &dev);
            }
        }
        oxpt_accum *acc = nullptr;
        if (rc == OXPT_OK && !r.refine) rc = oxpt_accum_create(io, req_w(r), req_h(r), &acc);
        std::string why = rc != OXPT_OK ? oxpt_last_error(io) : "";
        std::lock_guard<std::mutex> lk(mu);
        if (need_scene) {
            ps->loading = false;
            if (dev) {
                ps->dev = dev;
                const uint64_t actual = oxpt_scene_bytes(dev);
                // Device-built scenes: replace the admission bound by the size.
                resident = resident - std::min(resident, ps->upload_bytes) + actual;
                ps->upload_bytes = actual;
            } else {
                resident -= std::min(resident, ps->upload_bytes);
            }
        }
        if (rc != OXPT_OK || !ps->dev) {
            if (acc) oxpt_accum_destroy(acc);
            if (!r.refine) resident -= std::min(resident, r.acc_bytes);
            release_locked(r, false);
            ps->active--;
            active.erase(std::remove(active.begin(), active.end(), r.id), active.end());
            reject_locked(r, why.empty() ? std::string("scene upload failed") : why);
            cv_done.notify_all();
        } else {
            if (!r.refine) {
                r.acc = acc;
                std::vector<uint32_t> tmp;
                oxpt::CameraFrame cam = oxpt::resolve_camera(ps->cam, req_w(r), req_h(r));
                // Params only: the scene's tables (environment included) are resident.
                oxpt::bind(ps->pk, ps->env ? &*ps->env : nullptr, cam, r.opts, req_w(r), req_h(r), tmp, r.P, false);
            }
            r.state = State::Active;
        }
        stats.io_s += now_s() - t0;
        change_gen++;
        cv_work.notify_all();
    }
}

void oxpt_server::run() {
    std::vector<oxpt_batch_item> items;
    std::vector<Req *> in_tick, finished;
    uint64_t seen_gen = ~0ull;
    for (;;) {
        uint64_t tick_paths = 0;
        {
            std::unique_lock<std::mutex> lk(mu);
            // Sleep until there is something to run or resolve, or something
            // changed that may let a queued request in.
            auto runnable = [&] {
                for (uint64_t id : active)
                    if (reqs[id]->state == State::Active) return true;
                return false;
            };
            cv_work.wait(lk, [&] {
                return stop || !finished.empty() || runnable() || (!queue.empty() && change_gen != seen_gen);
            });
            if (stop) return;
            seen_gen = change_gen;
            // Admit in arrival order while slots and memory allow.
            while (!queue.empty() && (cfg.max_active == 0 || active.size() < cfg.max_active)) {
                Req &r = *reqs[queue.front()];
                if (r.refine) {
                    Req &p = *reqs[r.parent];
                    if (p.state != State::Done && p.state != State::Rejected) break;  // parent still running: keep order
                    if (p.state == State::Rejected || !p.acc) {
                        reject_locked(r, "refine: the render it continues is not resident");
                        queue.pop_front();
                        cv_done.notify_all();
                        continue;
                    }
                    // Only one request may advance an accumulator at a time.
                    bool busy = false;
                    for (uint64_t a : active)
                        if (reqs[a]->acc == p.acc) busy = true;
                    if (busy) break;
                }
                if (!admit_locked(r)) break;
                queue.pop_front();
                if (r.state == State::Rejected) cv_done.notify_all();
            }
            // Form the tick.
            items.clear();
            in_tick.clear();
            std::vector<Req *> ready;
            for (uint64_t id : active)
                if (reqs[id]->state == State::Active) ready.push_back(reqs[id].get());
            const uint64_t cap = cfg.tick_paths;
            if (cfg.tick_ms > 0) {
                // Closest to done first; about tick_ms of work; each request
                // at least one sample while the hard cap allows.
                std::stable_sort(ready.begin(), ready.end(), [](const Req *a, const Req *b) {
                    return a->pix64 * a->remaining < b->pix64 * b->remaining;
                });
                const double budget = std::min<double>((double)cap, rate * cfg.tick_ms * 1e-3);
                for (Req *r : ready) {
                    uint32_t n = r->remaining;
                    if (cfg.spp_per_tick) n = std::min(n, cfg.spp_per_tick);
                    const double left = budget - (double)tick_paths;
                    const uint64_t fit = left > 0 ? (uint64_t)(left / (double)r->pix64) : 0;
                    n = (uint32_t)std::min<uint64_t>(n, std::max<uint64_t>(fit, 1));
                    if (tick_paths + r->pix64 * n > cap && !in_tick.empty()) continue;
                    tick_paths += r->pix64 * n;
                    oxpt_batch_item it{};
                    it.scene = pool[r->scene]->dev;
                    it.accum = r->acc;
                    it.params = r->P;
                    it.first = r->first;
                    it.count = n;
                    items.push_back(it);
                    in_tick.push_back(r);
                }
            } else {
                for (Req *r : ready) {
                    uint32_t n = std::min(r->remaining, std::max(cfg.spp_per_tick, 1u));
                    // Within the cap; a request that alone exceeds it runs with
                    // as few samples as fit (at least one).
                    if (tick_paths + r->pix64 * n > cap) {
                        if (!in_tick.empty()) continue;
                        n = (uint32_t)std::max<uint64_t>(1, cap / r->pix64);
                    }
                    tick_paths += r->pix64 * n;
                    oxpt_batch_item it{};
                    it.scene = pool[r->scene]->dev;
                    it.accum = r->acc;
                    it.params = r->P;
                    it.first = r->first;
                    it.count = n;
                    items.push_back(it);
                    in_tick.push_back(r);
                }
            }
        }
        // The last tick's finished requests: resolve them now, in stream
        // order before the next launch, and hand the readback to the I/O
        // thread, which copies while the next tick runs.
        if (!finished.empty()) {
            IoJob job;
            job.readback = true;
            for (Req *r : finished) {
                oxpt_resolve_params rp{};
                rp.cfg[0] = 1;
                rp.cfg[1] = 3;
                rp.cfg[2] = r->opts.tone_map <= 2 ? r->opts.tone_map : 0u;
                rp.exposure[0] = std::max(r->opts.exposure, 0.0f);
                rp.exposure[1] = (float)r->samples_total;
                rp.exposure[2] = (float)req_w(*r);
                for (int k = 0; k < 4; k++) rp.background[k] = (float)r->opts.background[k] / 255.0f;
                const size_t px = (size_t)req_w(*r) * req_h(*r);
                int rc = oxpt_resolve_device(ctx, r->acc, &rp, r->staging->dev, px * 4);
                if (rc == OXPT_OK && cfg.keep_hdr) {
                    rp.cfg[3] = 1;
                    rc = oxpt_resolve_device(ctx, r->acc, &rp, (uint8_t *)r->staging->dev + px * 4, px * 16);
                }
                if (rc != OXPT_OK) {
                    std::lock_guard<std::mutex> lk(mu);
                    pool[r->scene]->active--;
                    release_locked(*r, false);
                    reject_locked(*r, oxpt_last_error(ctx));
                    cv_done.notify_all();
                    continue;
                }
                job.reqs.push_back(r);
            }
            finished.clear();
            if (!job.reqs.empty()) {
                if (oxpt::ctx_mark(ctx, &job.mark) != OXPT_OK) (void)oxpt_finish(ctx);  // then nothing to wait for
                std::lock_guard<std::mutex> lk(mu);
                io_jobs.push_back(std::move(job));
                cv_io.notify_one();
            }
        }
        if (items.empty()) continue;
        double t0 = now_s();
        int rc;
        std::string why;
        if (cfg.fail_tick && tick_no + 1 == cfg.fail_tick) {
            rc = OXPT_E_BACKEND;
            why = "injected tick failure";
        } else {
            rc = oxpt_refine_batch(ctx, items.data(), items.size());
            if (rc == OXPT_OK) rc = oxpt_finish(ctx);
            if (rc != OXPT_OK) why = oxpt_last_error(ctx);
        }
        double t1 = now_s();
        std::lock_guard<std::mutex> lk(mu);
        tick_no++;
        stats.ticks++;
        stats.busy_s += t1 - t0;
        if (rc == OXPT_OK && t1 - t0 > 1e-3) rate = 0.7 * rate + 0.3 * (double)tick_paths / (t1 - t0);
        if (rc != OXPT_OK) {
            // The accumulators this tick wrote are indeterminate: every request
            // in it fails and gives back what it held (a refine's parent loses
            // its accumulator too; its resolved image stays).
            stats.failed_ticks++;
            for (Req *r : in_tick) {
                pool[r->scene]->active--;
                release_locked(*r, true);
                reject_locked(*r, why);
                active.erase(std::remove(active.begin(), active.end(), r->id), active.end());
            }
            change_gen++;
            cv_done.notify_all();
            continue;
        }
        for (size_t k = 0; k < in_tick.size(); k++) {
            Req &r = *in_tick[k];
            r.first += items[k].count;
            r.remaining -= items[k].count;
            if (r.remaining == 0) {
                r.state = State::Resolving;
                finished.push_back(&r);
                active.erase(std::remove(active.begin(), active.end(), r.id), active.end());
            }
        }
        change_gen++;
    }
}

extern "C" {

void oxpt_server_config_default(oxpt_server_config *c) {
    std::memset(c, 0, sizeof *c);
    c->hbm_budget = 200ull << 30;
    c->spp_per_tick = 16;
    c->max_active = 0;
    c->keep_hdr = 0;
    c->tick_ms = 0;
    c->tick_paths = kDefaultTickPaths;
}

int oxpt_server_create(int device, const oxpt_server_config *cfg, oxpt_server **out) {
    if (!out) return OXPT_E_ARG;
    *out = nullptr;
    oxpt_ctx *c = nullptr;
    int rc = oxpt_ctx_create(device, &c);
    if (rc) return rc;
    oxpt_ctx *io = nullptr;
    if ((rc = oxpt_ctx_create_peer(c, &io))) {
        oxpt_ctx_destroy(c);
        return rc;
    }
    auto *s = new oxpt_server();
    s->ctx = c;
    s->io = io;
    if (cfg) s->cfg = *cfg;
    else oxpt_server_config_default(&s->cfg);
    if (s->cfg.tick_paths == 0) s->cfg.tick_paths = kDefaultTickPaths;
    s->cfg.tick_paths = std::min<uint64_t>(s->cfg.tick_paths, 1ull << 31);
    // The launch scratch is reserved up front: the per-path sample buffer of
    // the largest tick and the NEE queue.
    s->reserve = s->cfg.tick_paths * 16 + kQueueReserve;
    s->resident = s->reserve;
    oxpt_ctx_set_budget(c, s->cfg.hbm_budget);
    oxpt_ctx_set_graph(c, s->cfg.use_graph ? 1 : 0);
    s->worker = std::thread([s] { s->run(); });
    s->io_thread = std::thread([s] { s->run_io(); });
    *out = s;
    return OXPT_OK;
}

void oxpt_server_destroy(oxpt_server *s) {
    if (!s) return;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
        s->queue.clear();
    }
    s->cv_work.notify_all();
    s->cv_io.notify_all();
    if (s->worker.joinable()) s->worker.join();
    if (s->io_thread.joinable()) s->io_thread.join();
    (void)oxpt_finish(s->ctx);
    (void)oxpt_finish(s->io);
    for (auto &j : s->io_jobs)
        if (j.mark) (void)oxpt::mark_wait(s->ctx, j.mark);
    for (auto &kv : s->reqs)
        if (!kv.second->refine && kv.second->acc) oxpt_accum_destroy(kv.second->acc);
    for (auto &ps : s->pool)
        if (ps->dev) oxpt_scene_destroy(ps->dev);
    for (auto &st : s->staging) oxpt::device_free(s->ctx, st->dev, st->bytes);
    oxpt_ctx_destroy(s->io);
    oxpt_ctx_destroy(s->ctx);
    delete s;
}

oxpt_ctx *oxpt_server_ctx(oxpt_server *s) { return s ? s->ctx : nullptr; }
const char *oxpt_server_last_error(const oxpt_server *s) { return s ? s->err.c_str() : "null server"; }

int oxpt_server_add_scene(oxpt_server *s, const oxpt_host_scene *scene, const oxpt_env_map *env, uint32_t *id) {
    if (!s || !scene || !id) return OXPT_E_ARG;
    auto ps = std::make_unique<PoolScene>();
    if (env) ps->env = env->e;
    ps->cam = oxpt::camera_source(scene->s);
    if (s->cfg.device_build && scene->s.soup.count() > 0) {
        // Prepared on the device at first use. Admission bound: BVH nodes
        // (at most 2 per triangle), 4-wide nodes, triangle and attribute
        // words, tables and the environment section, the atlas.
        ps->host = scene->s;
        const oxpt::Packed layout = oxpt::pack_scene(scene->s, false, nullptr, false, false);
        const uint64_t n = scene->s.soup.count();
        const uint64_t env_words = env ? (uint64_t)(env->e.h + 1) + (uint64_t)env->e.h * (env->e.w + 1) +
                                             4ull * env->e.w * env->e.h
                                       : 0;
        ps->upload_bytes = n * (64 + 64 + 48 + 192) + 4 * (layout.tables.size() + 4 * n + env_words) +
                           (layout.chains.empty() ? 8 : (uint64_t)layout.atlas_side * layout.atlas_side *
                                                            layout.atlas_layers * 8);
        std::lock_guard<std::mutex> lk(s->mu);
        *id = (uint32_t)s->pool.size();
        s->pool.push_back(std::move(ps));
        return OXPT_OK;
    }
    ps->pk = oxpt::pack_scene(scene->s);
    oxpt_render_options o;
    oxpt_render_options_default(&o);
    oxpt_params P;
    oxpt::CameraFrame cam = oxpt::resolve_camera(ps->cam, o.width, o.height);
    oxpt::bind(ps->pk, env ? &env->e : nullptr, cam, o, o.width, o.height, ps->tables, P);
    ps->upload_bytes = ps->pk.geom.size() * 4 + ps->pk.attrs.size() * 4 + ps->tables.size() * 4 +
                       (ps->pk.atlas.empty() ? 8 : (uint64_t)ps->pk.atlas_side * ps->pk.atlas_side * ps->pk.atlas_layers * 8);
    std::lock_guard<std::mutex> lk(s->mu);
    *id = (uint32_t)s->pool.size();
    s->pool.push_back(std::move(ps));
    return OXPT_OK;
}

int oxpt_server_submit_render(oxpt_server *s, uint32_t scene_id, const oxpt_render_options *o, uint64_t *req) {
    if (!s || !o || !req) return OXPT_E_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    if (scene_id >= s->pool.size()) {
        s->err = "submit: unknown scene";
        return OXPT_E_ARG;
    }
    if (o->width > 32768 || o->height > 32768) {
        s->err = "submit: dimensions beyond 32768";
        return OXPT_E_INVALID;
    }
    auto r = std::make_unique<Req>();
    r->id = s->next_id++;
    r->scene = scene_id;
    r->opts = *o;
    r->t_arrival = now_s();
    *req = r->id;
    s->queue.push_back(r->id);
    s->reqs[r->id] = std::move(r);
    s->stats.submitted++;
    s->change_gen++;
    s->cv_work.notify_all();
    return OXPT_OK;
}

int oxpt_server_submit_refine(oxpt_server *s, uint64_t render_req, uint32_t spp, uint64_t *req) {
    if (!s || !req || spp == 0) return OXPT_E_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(render_req);
    if (it == s->reqs.end() || it->second->refine) {
        s->err = "refine: unknown render request";
        return OXPT_E_ARG;
    }
    auto r = std::make_unique<Req>();
    r->id = s->next_id++;
    r->refine = true;
    r->parent = render_req;
    r->scene = it->second->scene;
    r->spp = spp;
    r->opts = it->second->opts;
    r->t_arrival = now_s();
    *req = r->id;
    s->queue.push_back(r->id);
    s->reqs[r->id] = std::move(r);
    s->stats.submitted++;
    s->change_gen++;
    s->cv_work.notify_all();
    return OXPT_OK;
}

int oxpt_server_wait(oxpt_server *s, uint64_t req, double timeout_s, oxpt_request_info *info) {
    if (!s) return OXPT_E_ARG;
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->reqs.find(req);
    if (it == s->reqs.end()) return OXPT_E_ARG;
    Req &r = *it->second;
    auto done = [&] { return r.state == State::Done || r.state == State::Rejected; };
    if (timeout_s < 0) s->cv_done.wait(lk, done);
    else if (!s->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), done)) {
        s->err = "wait: timed out";
        return OXPT_E_BACKEND;
    }
    if (info) {
        std::memset(info, 0, sizeof *info);
        info->ok = r.state == State::Done ? 1 : 0;
        info->width = req_w(r);
        info->height = req_h(r);
        info->samples = r.samples_total;
        info->latency_s = r.t_done - r.t_arrival;
        info->queue_s = r.t_admit > 0 ? r.t_admit - r.t_arrival : r.t_done - r.t_arrival;
        info->paths = (uint64_t)info->width * info->height * (r.refine ? r.spp : r.samples_total);
    }
    if (r.state == State::Rejected) {
        s->err = r.error;
        return OXPT_E_BACKEND;
    }
    return OXPT_OK;
}

int oxpt_server_result(oxpt_server *s, uint64_t req, int hdr, void *out, size_t bytes) {
    if (!s || !out) return OXPT_E_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(req);
    if (it == s->reqs.end() || it->second->state != State::Done) return OXPT_E_ARG;
    Req &r = *it->second;
    if (hdr) {
        if (r.hdr.empty() || bytes < r.hdr.size() * 4) return OXPT_E_ARG;
        std::memcpy(out, r.hdr.data(), r.hdr.size() * 4);
    } else {
        if (bytes < r.image.size()) return OXPT_E_ARG;
        std::memcpy(out, r.image.data(), r.image.size());
    }
    return OXPT_OK;
}

int oxpt_server_release(oxpt_server *s, uint64_t req) {
    if (!s) return OXPT_E_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(req);
    if (it == s->reqs.end()) return OXPT_E_ARG;
    Req &r = *it->second;
    if (r.state != State::Done && r.state != State::Rejected) return OXPT_E_ARG;
    r.image.clear();
    r.image.shrink_to_fit();
    r.hdr.clear();
    r.hdr.shrink_to_fit();
    return OXPT_OK;
}

int oxpt_server_stats_get(oxpt_server *s, oxpt_server_stats *out) {
    if (!s || !out) return OXPT_E_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    *out = s->stats;
    out->resident_bytes = s->resident;
    out->active = (uint32_t)s->active.size();
    out->queued = (uint32_t)s->queue.size();
    return OXPT_OK;
}

}  // extern "C"
