// Serving loop: continuous batching of render and progressive-refine
// requests over the batched launch (oxpt_refine_batch).
//
// * A pool of prepared scenes (host words) is registered up front; a render
//   request names a pool scene and render options. Scenes are uploaded on
//   first use and stay resident (BVH, tables, atlas) with their
//   accumulators, so a later refine continues the same accumulation — the
//   reference's progressive contract (src/pathtrace.rs:578-583).
// * Admission is by an HBM budget: a request is admitted when its scene
//   upload and accumulator fit; otherwise least-recently-used idle scenes
//   (and their finished accumulators) are evicted, and if that is not
//   enough the request waits (FIFO).
// * A scheduler thread forms one batched launch per tick from every active
//   request, each advancing at most `spp_per_tick` samples; requests that
//   arrive while a tick runs join the next one. Each pixel's samples are
//   still added in order, so every image equals its image rendered alone.
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
    uint64_t last_use = 0;
    int active = 0;  // active requests on this scene
};

enum class State { Queued, Active, Done, Rejected };

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
    uint32_t first = 0, remaining = 0, samples_total = 0;
    double t_arrival = 0, t_admit = 0, t_done = 0;
    State state = State::Queued;
    std::string error;
    std::vector<uint8_t> image;  // RGBA8 once done
    std::vector<float> hdr;      // RGBA32F once done
};

}  // namespace

struct oxpt_server {
    oxpt_ctx *ctx = nullptr;
    oxpt_server_config cfg{};
    std::string err;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::thread worker;
    bool stop = false;
    std::vector<std::unique_ptr<PoolScene>> pool;
    std::map<uint64_t, std::unique_ptr<Req>> reqs;
    std::deque<uint64_t> queue;   // admission order
    std::vector<uint64_t> active;
    uint64_t next_id = 1, tick_no = 0, use_clock = 0;
    uint64_t change_gen = 0;  // bumped by every submission and completion
    uint64_t resident = 0;
    oxpt_server_stats stats{};

    void run();
    bool admit_locked(Req &r);
    void evict_for_locked(uint64_t need);
    void finish_request_locked(Req &r);
};

bool oxpt_server::admit_locked(Req &r) {
    PoolScene &ps = *pool[r.scene];
    uint64_t need = 0;
    if (!ps.dev) need += ps.upload_bytes;
    if (!r.refine) need += (uint64_t)std::max(r.opts.width, 1u) * std::max(r.opts.height, 1u) * 16;
    if (resident + need > cfg.hbm_budget) evict_for_locked(need);
    if (resident + need > cfg.hbm_budget) return false;
    if (!ps.dev && ps.host) {
        int rc = oxpt::device_prepare(ctx, *ps.host, ps.pk, &ps.dev, nullptr);
        if (rc == OXPT_OK) {
            rc = oxpt::upload_tables(ctx, ps.dev, ps.pk.tables, ps.env ? &*ps.env : nullptr);
            if (rc != OXPT_OK) {
                oxpt_scene_destroy(ps.dev);
                ps.dev = nullptr;
            }
        }
        if (rc != OXPT_OK) {
            r.state = State::Rejected;
            r.error = oxpt_last_error(ctx);
            return true;
        }
        ps.upload_bytes = oxpt_scene_bytes(ps.dev);  // the actual size from now on
        resident += ps.upload_bytes;
        stats.uploads++;
    } else if (!ps.dev) {
        oxpt_scene_desc d{};
        d.geom = ps.pk.geom.data();
        d.geom_vec4 = ps.pk.geom.size() / 4;
        d.attrs = ps.pk.attrs.data();
        d.attrs_vec4 = ps.pk.attrs.size() / 4;
        d.tables = ps.tables.data();
        d.table_words = ps.tables.size();
        d.atlas = ps.pk.atlas.empty() ? nullptr : ps.pk.atlas.data();
        d.atlas_side = ps.pk.atlas_side;
        d.atlas_layers = ps.pk.atlas_layers;
        if (oxpt_scene_upload(ctx, &d, &ps.dev) != OXPT_OK) {
            r.state = State::Rejected;
            r.error = oxpt_last_error(ctx);
            return true;
        }
        resident += ps.upload_bytes;
        stats.uploads++;
    }
    if (r.refine) {
        Req &p = *reqs[r.parent];
        r.acc = p.acc;
        r.P = p.P;
        r.first = p.samples_total;
        r.remaining = r.spp;
        r.samples_total = p.samples_total + r.spp;
        p.samples_total = r.samples_total;
        r.opts = p.opts;
    } else {
        uint32_t w = std::max(r.opts.width, 1u), h = std::max(r.opts.height, 1u);
        if (oxpt_accum_create(ctx, w, h, &r.acc) != OXPT_OK) {
            r.state = State::Rejected;
            r.error = oxpt_last_error(ctx);
            return true;
        }
        r.acc_bytes = (uint64_t)w * h * 16;
        resident += r.acc_bytes;
        // Params only: the scene's tables (environment included) are resident.
        std::vector<uint32_t> tmp;
        oxpt::CameraFrame cam = oxpt::resolve_camera(ps.cam, w, h);
        oxpt::bind(ps.pk, ps.env ? &*ps.env : nullptr, cam, r.opts, w, h, tmp, r.P, false);
        r.first = 0;
        r.remaining = std::max(r.opts.samples_per_pixel, 1u);
        r.samples_total = r.remaining;
    }
    ps.active++;
    ps.last_use = ++use_clock;
    r.state = State::Active;
    r.t_admit = now_s();
    active.push_back(r.id);
    return true;
}

void oxpt_server::evict_for_locked(uint64_t need) {
    while (resident + need > cfg.hbm_budget) {
        PoolScene *victim = nullptr;
        uint32_t vi = 0;
        for (uint32_t i = 0; i < pool.size(); i++) {
            PoolScene &ps = *pool[i];
            if (ps.dev && ps.active == 0 && (!victim || ps.last_use < victim->last_use)) {
                victim = &ps;
                vi = i;
            }
        }
        if (!victim) return;
        // Drop the scene and every finished accumulator on it.
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
}

void oxpt_server::finish_request_locked(Req &r) {
    oxpt_resolve_params rp{};
    rp.cfg[0] = 1;
    rp.cfg[1] = 3;
    rp.cfg[2] = r.opts.tone_map <= 2 ? r.opts.tone_map : 0u;
    rp.exposure[0] = std::max(r.opts.exposure, 0.0f);
    rp.exposure[1] = (float)r.samples_total;
    rp.exposure[2] = (float)std::max(r.opts.width, 1u);
    for (int k = 0; k < 4; k++) rp.background[k] = (float)r.opts.background[k] / 255.0f;
    uint32_t w = std::max(r.opts.width, 1u), h = std::max(r.opts.height, 1u);
    r.image.resize((size_t)w * h * 4);
    int rc = oxpt_resolve(ctx, r.acc, &rp, r.image.data(), r.image.size());
    if (rc == OXPT_OK && cfg.keep_hdr) {
        rp.cfg[3] = 1;
        r.hdr.resize((size_t)w * h * 4);
        rc = oxpt_resolve(ctx, r.acc, &rp, r.hdr.data(), r.hdr.size() * 4);
    }
    if (rc != OXPT_OK) {
        r.state = State::Rejected;
        r.error = oxpt_last_error(ctx);
    } else {
        r.state = State::Done;
    }
    r.t_done = now_s();
    PoolScene &ps = *pool[r.scene];
    ps.active--;
    stats.completed++;
    stats.paths += (uint64_t)w * h * (r.refine ? r.spp : r.samples_total);
}

void oxpt_server::run() {
    std::vector<oxpt_batch_item> items;
    std::vector<Req *> in_tick;
    uint64_t seen_gen = ~0ull;
    for (;;) {
        {
            std::unique_lock<std::mutex> lk(mu);
            // Sleep until there is something to run, or something changed that
            // may let a queued request in (a submission, or a completion that
            // freed its scene or finished the render a refine waits for).
            cv_work.wait(lk, [&] { return stop || !active.empty() || (!queue.empty() && change_gen != seen_gen); });
            if (stop) return;
            seen_gen = change_gen;
            // Admit in arrival order while slots and memory allow.
            while (!queue.empty() && (cfg.max_active == 0 || active.size() < cfg.max_active)) {
                Req &r = *reqs[queue.front()];
                if (r.refine) {
                    Req &p = *reqs[r.parent];
                    if (p.state != State::Done && p.state != State::Rejected) break;  // parent still running: keep order
                    if (p.state == State::Rejected || !p.acc) {
                        r.state = State::Rejected;
                        r.error = "refine: the render it continues is not resident";
                        r.t_done = now_s();
                        stats.rejected++;
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
                if (r.state == State::Rejected) {
                    r.t_done = now_s();
                    stats.rejected++;
                    cv_done.notify_all();
                }
            }
            items.clear();
            in_tick.clear();
            for (uint64_t id : active) {
                Req &r = *reqs[id];
                uint32_t n = std::min(r.remaining, std::max(cfg.spp_per_tick, 1u));
                oxpt_batch_item it{};
                it.scene = pool[r.scene]->dev;
                it.accum = r.acc;
                it.params = r.P;
                it.first = r.first;
                it.count = n;
                items.push_back(it);
                in_tick.push_back(&r);
            }
        }
        if (items.empty()) continue;
        double t0 = now_s();
        int rc = oxpt_refine_batch(ctx, items.data(), items.size());
        if (rc == OXPT_OK) rc = oxpt_finish(ctx);
        double t1 = now_s();
        std::lock_guard<std::mutex> lk(mu);
        stats.ticks++;
        stats.busy_s += t1 - t0;
        std::vector<uint64_t> still;
        for (size_t k = 0; k < in_tick.size(); k++) {
            Req &r = *in_tick[k];
            if (rc != OXPT_OK) {
                r.state = State::Rejected;
                r.error = oxpt_last_error(ctx);
                r.t_done = now_s();
                pool[r.scene]->active--;
                stats.rejected++;
                continue;
            }
            r.first += items[k].count;
            r.remaining -= items[k].count;
            if (r.remaining == 0) finish_request_locked(r);
            else still.push_back(r.id);
        }
        active.swap(still);
        change_gen++;
        cv_done.notify_all();
    }
}

extern "C" {

void oxpt_server_config_default(oxpt_server_config *c) {
    std::memset(c, 0, sizeof *c);
    c->hbm_budget = 200ull << 30;
    c->spp_per_tick = 16;
    c->max_active = 0;
    c->keep_hdr = 0;
}

int oxpt_server_create(int device, const oxpt_server_config *cfg, oxpt_server **out) {
    if (!out) return OXPT_E_ARG;
    *out = nullptr;
    oxpt_ctx *c = nullptr;
    int rc = oxpt_ctx_create(device, &c);
    if (rc) return rc;
    auto *s = new oxpt_server();
    s->ctx = c;
    if (cfg) s->cfg = *cfg;
    else oxpt_server_config_default(&s->cfg);
    oxpt_ctx_set_graph(c, s->cfg.use_graph ? 1 : 0);
    s->worker = std::thread([s] { s->run(); });
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
    if (s->worker.joinable()) s->worker.join();
    for (auto &kv : s->reqs)
        if (!kv.second->refine && kv.second->acc) oxpt_accum_destroy(kv.second->acc);
    for (auto &ps : s->pool)
        if (ps->dev) oxpt_scene_destroy(ps->dev);
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
    else if (!s->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), done)) return OXPT_E_BACKEND;
    if (info) {
        std::memset(info, 0, sizeof *info);
        info->ok = r.state == State::Done ? 1 : 0;
        info->width = std::max(r.opts.width, 1u);
        info->height = std::max(r.opts.height, 1u);
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
