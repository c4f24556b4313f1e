// Device boundary of liboxpt: contexts, resident scenes, accumulators,
// refine launches and the resolve readback (include/oxpt/oxpt.h, layer 1).
// Replaces the reference's wgpu device/queue plumbing
// (oxideav-render-vulkan src/gpu.rs:984-1117, src/pathtrace.rs:584-844).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "oxpt/oxpt.h"
#include "bvh_build.hpp"
#include "pt_kernels.hpp"
#include "wavefront.hpp"

namespace {

// LDS taken by the Sobol' and sheen tables in every trace workgroup, plus
// the light-selection CDF copies of the single-scene kernels.
constexpr uint32_t kTableLds = 2560;
constexpr uint32_t kCdfLds = 4352;
// Persistent kernel tail in 1/16 of the frame (oxpt_tuning.tail), and the
// most tail samples one launch buffers (16 B each: 4 GiB).
constexpr uint32_t kDefaultTail = 8;
// Default path cap per launch (oxpt_tuning.max_paths): as many as 32-bit
// work indices allow, so a batch splits into as few rounds as it can — the
// fewer the rounds, the longer each pixel run (64 uneven terrain scenes:
// 6 rounds at 2^30 paths 952 Mpaths/s, 2 rounds 976).
constexpr uint32_t kDefaultMaxPaths = 0xFFFFFFFFu;
// Batched rounds hold one 16-byte sample slot per tail path; a round's
// samples per item are halved until its slots fit this cap (16 GiB).
constexpr uint64_t kBatchSlotCap = 1ull << 30;
constexpr uint64_t kTailSlots = 1ull << 28;
// Batched rounds of fewer paths than kSmallRound run the 3-waves form by
// default, where the end of the round (one path's latency) weighs more than
// throughput; without punctual lights, rounds from kTinyRound up run the
// 5-waves form (a 256x256 frame at 16 spp, 1 M paths: 3 waves 923, 5 waves
// 848 Mpaths/s; serving ticks of about 8 M paths: 5 waves 794, 3 waves 779,
// 6 waves 756 Mpaths/s of busy time).
constexpr double kSmallRound = 16.0 * 1024 * 1024;
constexpr double kTinyRound = 4.0 * 1024 * 1024;
// Default forms without punctual lights (early continuation): the
// persistent megakernel of a refine and the batched megakernel, both built
// for 6 waves per SIMD (the 1080p frame: 5 waves 1021, 6 waves 1032
// Mpaths/s; the uneven batch of 64 scenes: 5 waves 1063, 6 waves 1073).
constexpr uint32_t kEarlyRefineForm = 5;
constexpr uint32_t kEarlyDefaultForm = 5;
// Default form of a refine with 1..31 punctual lights (early continuation
// with a per-lane NEE queue in device memory).
constexpr uint32_t kQueueDefaultForm = 0;

// Kernel mode of a refine whose scene(s) hold up to n_punct punctual lights
// (0 general, 1 early, 2 early with an NEE queue).
uint32_t trace_mode(const oxpt_tuning &t, uint32_t n_punct) {
    // Early builds exist for the default form and forms 2..5.
    if (t.early == OXPT_EARLY_OFF || !(t.kernel == 0 || (t.kernel >= 2 && t.kernel <= 5))) return 0;
    if (n_punct == 0) return 1;
    return n_punct <= oxd::kQueueMaxPunctual ? 2u : 0u;
}

// Waves per SIMD each megakernel form is built for (oxpt_tuning.kernel;
// kernel 0 is 4 waves for both the single refine and the batched megakernel).
uint32_t form_waves(uint32_t kernel) {
    switch (kernel) {
        case 2: return 3;
        case 3: return 4;
        case 4: return 5;
        case 5: return 6;
        case 6: return 8;
        case 8: return 2;
        default: return 4;
    }
}

// Traversal-stack entries per lane held in LDS. Default: as many as leave
// every resident workgroup room on a CU (160 KiB of LDS), at most 16. A
// workgroup's LDS is allocated in granules, with its dynamic part aligned
// after the static one: 1 KiB per workgroup is kept free for that, since a
// workgroup a few hundred bytes short of its share lowers the waves per
// SIMD the form was built for.
uint32_t stack_lds_entries(const oxpt_tuning &t, uint32_t block, bool batched) {
    if (t.stack_lds == OXPT_STACK_PRIVATE) return 0;
    if (t.stack_lds != 0) return t.stack_lds;
    const uint32_t waves_cu = 4u * form_waves(t.kernel);
    const uint32_t wg_cu = std::max(1u, waves_cu / (block / 64u));
    const uint32_t per_wg = 160u * 1024u / wg_cu;
    const uint32_t fixed = kTableLds + (batched ? 0u : kCdfLds) + 1024u;
    if (per_wg <= fixed) return 0;
    return std::min(16u, (per_wg - fixed) / ((block / 64u) * 512u));
}

// Whether the opaque walk applies to a scene under these Params: every
// material word in the material section [offs0[2], offs0[3]) has alpha mode
// OPAQUE (flags bits 0-1 clear). The section is read from the host copy of
// the tables, at OXPT_MAT_STRIDE words per material; a section that is not
// a whole number of materials, or lies outside the tables, counts as not
// opaque (the general walk, which is right for every scene).
bool opaque_walk(const oxpt_tuning &t, const std::vector<uint32_t> &tables, const oxpt_params &P) {
    if (t.opaque == OXPT_OPAQUE_OFF) return false;
    const uint64_t a = P.offs0[2], b = P.offs0[3];
    if (b < a || b > tables.size() || (b - a) % OXPT_MAT_STRIDE != 0) return false;
    for (uint64_t m = a; m < b; m += OXPT_MAT_STRIDE)
        if (tables[m + OXPT_M_FLAGS] & 3u) return false;
    return true;
}

// Device opening is serialised process-wide (src/gpu.rs:978-992).
std::mutex g_device_init;

struct EventPair {
    hipEvent_t a, b;
};

// Device memory shared by a context and its peers: one budget, one count
// of bytes in use (contexts on different threads allocate concurrently).
struct Domain {
    std::mutex mu;
    uint64_t budget = 0, in_use = 0;
};

}  // namespace

// Scratch of oxpt_refine_batch_noalloc: work counters (4 KiB), the job
// table (256 jobs) and the per-path sample buffer of a round (2^22 paths x
// 16 B), in one device array that every GPU holds from module load to
// process end. Host code never uses the array's name as an address: each
// device's copy is found once with hipGetSymbolAddress (noalloc_scratch).
namespace {
constexpr uint32_t kNaJobs = 256;
constexpr uint64_t kNaSlots = 1ull << 22;
constexpr size_t kNaCounterBytes = 4096;
constexpr size_t kNaJobBytes = (size_t)kNaJobs * 256;
constexpr size_t kNaBytes = kNaCounterBytes + kNaJobBytes + kNaSlots * 16;
constexpr int kNaMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kNaMaxDevices] = {};
}  // namespace
static __device__ __align__(256) unsigned char scratch[kNaBytes];

struct oxpt_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    std::string summary;
    // The context that owns the objects created through this one (itself
    // unless this is a peer), and the memory domain they share.
    oxpt_ctx *base = nullptr;
    std::shared_ptr<Domain> dom;
    uint32_t n_cu = 0;
    oxpt_tuning tuning{256, 32, kDefaultMaxPaths, 0, 0, 0, 0, 0, 0, 0};
    uint32_t *d_work = nullptr;
    oxd::Job *d_jobs = nullptr;      // job table; the count sits in d_counters[2 * job_cap]
    uint32_t *d_counters = nullptr;  // two counters per job (pixels, tail paths) + 1 count word
    size_t job_cap = 0;
    bool use_graph = false;
    hipGraphExec_t graph = nullptr;  // captured batched step (reset + trace)
    uint32_t graph_block = 0, graph_grid = 0, graph_kernel = 0, graph_lds = 0;
    uint32_t graph_mode = 0;
    bool graph_opaque = false;
    float *d_nee_q = nullptr;  // NEE queue of the early form with punctual lights
    size_t nee_q_bytes = 0;
    const char *last_kernel = "";  // name of the last trace kernel launched
    float4 *d_samples = nullptr;  // batched launch: one (rgb, covered) per path of a round
    size_t samples_bytes = 0;
    uint64_t graph_launches = 0;
    void *d_out = nullptr;
    size_t d_out_bytes = 0;
    std::vector<EventPair> events;
    uint64_t launches = 0;
    double total_ms = 0.0;
    // Wavefront buffers (grown on demand).
    struct Wf {
        oxd::PathSlot *slots = nullptr;
        oxd::HitRec *hits = nullptr;
        oxd::ShadowRec *shadows = nullptr;
        uint32_t *occluded = nullptr;
        float4 *samples = nullptr;
        oxd::Job *jobs = nullptr;
        uint32_t *job_base = nullptr;
        uint32_t *counters = nullptr;   // [0] item counter, [1..] live ring
        uint32_t *host_live = nullptr;  // pinned copy of the live ring
        std::vector<hipEvent_t> events;
        size_t slot_bytes = 0, hit_bytes = 0, shadow_bytes = 0, occ_bytes = 0, sample_bytes = 0, job_bytes = 0,
               base_bytes = 0;
        uint32_t n_slots = 0, max_shadow = 0;
        uint64_t iterations = 0;
    } wf;
};

struct oxpt_scene {
    oxpt_ctx *ctx = nullptr;
    std::vector<uint32_t> host_tables;  // estimator constants are checked across batched jobs
    float4 *geom = nullptr;
    float4 *attrs = nullptr;
    uint32_t *tables = nullptr;
    uint16_t *atlas = nullptr;
    uint32_t atlas_side = 1, atlas_layers = 1;
    size_t table_words = 0;
    size_t geom_vec4 = 0;
    size_t attrs_vec4 = 0;
    size_t atlas_words = 0;  // u16 texels held (4 when the scene has no atlas)
    size_t bytes = 0;
};

struct oxpt_accum {
    oxpt_ctx *ctx = nullptr;
    float4 *d = nullptr;
    uint32_t w = 0, h = 0;
    size_t bytes = 0;
};

struct oxpt_image {
    oxpt_ctx *ctx = nullptr;
    void *d = nullptr;
    uint32_t w = 0, h = 0;
    int hdr = 0;
    size_t bytes = 0;
};

namespace {

int fail(oxpt_ctx *c, int code, const std::string &msg) {
    if (c) c->err = msg;
    return code;
}

int hip_fail(oxpt_ctx *c, const char *what, hipError_t e) {
    return fail(c, OXPT_E_BACKEND, std::string("hip: ") + what + ": " + hipGetErrorString(e));
}

int dalloc(oxpt_ctx *c, size_t bytes, void **p, const char *what) {
    *p = nullptr;
    if (bytes == 0) bytes = 16;
    Domain &D = *c->dom;
    {
        std::lock_guard<std::mutex> lk(D.mu);
        if (D.in_use + bytes > D.budget) {
            char buf[256];
            std::snprintf(buf, sizeof buf, "hip: %s needs %zu bytes; the device budget leaves %llu", what, bytes,
                          (unsigned long long)(D.budget > D.in_use ? D.budget - D.in_use : 0));
            return fail(c, OXPT_E_BACKEND, buf);
        }
        D.in_use += bytes;
    }
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
        *p = nullptr;
        std::lock_guard<std::mutex> lk(D.mu);
        D.in_use -= std::min<uint64_t>(D.in_use, bytes);
        return hip_fail(c, what, e);
    }
    return OXPT_OK;
}

void dfree(oxpt_ctx *c, void *p, size_t bytes) {
    if (!p) return;
    (void)hipFree(p);
    if (bytes == 0) bytes = 16;
    std::lock_guard<std::mutex> lk(c->dom->mu);
    c->dom->in_use -= std::min<uint64_t>(c->dom->in_use, bytes);
}

// Objects (scenes, accumulators) of one context may be used with its peers.
inline bool same_domain(const oxpt_ctx *a, const oxpt_ctx *b) { return a && b && a->dom == b->dom; }

int setdev(oxpt_ctx *c) {
    hipError_t e = hipSetDevice(c->device);
    return e == hipSuccess ? OXPT_OK : hip_fail(c, "hipSetDevice", e);
}

void fold_events(oxpt_ctx *c, size_t keep) {
    while (c->events.size() > keep) {
        EventPair ep = c->events.front();
        c->events.erase(c->events.begin());
        (void)hipEventSynchronize(ep.b);
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, ep.a, ep.b) == hipSuccess) c->total_ms += ms;
        (void)hipEventDestroy(ep.a);
        (void)hipEventDestroy(ep.b);
    }
}

oxd::TraceArgs trace_args(const oxpt_ctx *c, const oxpt_scene *s, oxpt_accum *acc, const oxpt_params *P) {
    oxd::TraceArgs a;
    std::memset(&a, 0, sizeof a);
    a.P = *P;
    a.geom = s->geom;
    a.attrs = s->attrs;
    a.tables = s->tables;
    a.atlas = s->atlas;
    a.atlas_side = s->atlas_side;
    if (acc) {
        a.accum = acc->d;
        uint32_t tx = (acc->w + 7u) / 8u, ty = (acc->h + 7u) / 8u;
        a.tiles_x = tx;
        a.n_items = 64u * tx * ty;
        a.split = a.n_items;  // no tail unless the launch sets one
    }
    a.work = c->d_work;
    if (c->tuning.bvh == 2) a.P.offs2[3] = 0;  // walk the binary tree
    return a;
}

// Offsets the kernels dereference must lie inside the uploaded words.
int check_params(oxpt_ctx *c, const oxpt_scene *s, const oxpt_params *P) {
    size_t tw = s->table_words;
    auto in = [&](uint64_t off, uint64_t n) { return off + n <= tw; };
    if (!in(P->offs0[0], OXPT_SOBOL_WORDS) || !in(P->offs0[1], OXPT_SHEEN_WORDS))
        return fail(c, OXPT_E_ARG, "params: Sobol' / sheen offsets outside the tables");
    if (P->cfg[2] && !in(P->offs0[3], 16ull * P->cfg[2])) return fail(c, OXPT_E_ARG, "params: punctual lights outside the tables");
    if (P->cfg[3] && !in(P->offs1[0], 4ull * P->cfg[3])) return fail(c, OXPT_E_ARG, "params: emissive lights outside the tables");
    if (P->env[2]) {
        uint64_t w = P->env[0], h = P->env[1];
        if (w == 0 || h == 0 || !in(P->offs1[2], h + 1) || !in(P->offs1[3], h * (w + 1)) || !in(P->offs2[0], w * h) ||
            !in(P->offs2[1], 3 * w * h))
            return fail(c, OXPT_E_ARG, "params: environment sections outside the tables");
    }
    if (P->offs2[3] != 0 && (P->offs2[3] % 8u != 0 || (uint64_t)P->offs2[3] + 8u > s->geom_vec4))
        return fail(c, OXPT_E_ARG, "params: 4-wide nodes outside the geometry");
    return OXPT_OK;
}

}  // namespace

namespace {

// (Re)allocate a device buffer to hold at least `need` bytes.
template <typename T>
int wf_grow(oxpt_ctx *c, T *&p, size_t &have, size_t need, const char *what) {
    if (have >= need && p) return OXPT_OK;
    dfree(c, p, have);
    p = nullptr;
    have = 0;
    void *q = nullptr;
    if (int rc = dalloc(c, need, &q, what)) return rc;
    p = static_cast<T *>(q);
    have = need;
    return OXPT_OK;
}

// Wavefront engine: traces every job's [first, first + count) samples in
// passes of at most 2^27 (pixel, sample) items (max_paths caps it further),
// then adds each pass's samples into the accumulators in sample order.
int wf_run(oxpt_ctx *c, const std::vector<oxd::Job> &jobs) {
    auto &W = c->wf;
    uint32_t max_shadow = 0, max_count = 0;
    uint64_t pixels = 0;
    for (auto &J : jobs) {
        max_shadow = std::max(max_shadow, J.P.cfg[2] + 1u);
        pixels += (uint64_t)J.n_items;  // 64 * tiles (incl. ragged edges)
        max_count = std::max(max_count, J.count);
    }
    const uint64_t cap = std::min<uint64_t>(c->tuning.max_paths, 1ull << 27);
    uint32_t spd = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(max_count, cap / std::max<uint64_t>(pixels, 1)));
    uint64_t items_max = 0;
    for (auto &J : jobs) items_max += (uint64_t)J.n_items * std::min(spd, J.count);
    if (items_max == 0) return OXPT_OK;
    if (items_max >= 0xffffffffull) return fail(c, OXPT_E_ARG, "wavefront: pass too large");
    // Slots x shadow rays per slot stays below 2^31 (32-bit ray ids).
    const uint64_t slot_cap = std::min<uint64_t>(1ull << 21, ((1ull << 31) / max_shadow) / 256 * 256);
    if (slot_cap == 0) return fail(c, OXPT_E_ARG, "wavefront: too many punctual lights");
    uint32_t want_slots = (uint32_t)std::min<uint64_t>(slot_cap, (items_max + 255) / 256 * 256);
    if (W.n_slots < want_slots || W.max_shadow < max_shadow) {
        uint32_t ns = std::max(want_slots, W.n_slots), ms = std::max(max_shadow, W.max_shadow);
        if (int rc = wf_grow(c, W.slots, W.slot_bytes, (size_t)ns * sizeof(oxd::PathSlot), "wavefront slots")) return rc;
        if (int rc = wf_grow(c, W.hits, W.hit_bytes, (size_t)ns * sizeof(oxd::HitRec), "wavefront hits")) return rc;
        if (int rc = wf_grow(c, W.shadows, W.shadow_bytes, (size_t)ns * ms * sizeof(oxd::ShadowRec), "wavefront shadow rays"))
            return rc;
        if (int rc = wf_grow(c, W.occluded, W.occ_bytes, (size_t)ns * ms * 4, "wavefront occlusion")) return rc;
        W.n_slots = ns;
        W.max_shadow = ms;
    }
    if (int rc = wf_grow(c, W.samples, W.sample_bytes, items_max * 16, "wavefront samples")) return rc;
    if (int rc = wf_grow(c, W.jobs, W.job_bytes, jobs.size() * sizeof(oxd::Job), "wavefront jobs")) return rc;
    if (int rc = wf_grow(c, W.job_base, W.base_bytes, jobs.size() * 4, "wavefront job bases")) return rc;
    if (!W.counters) {
        void *q = nullptr;
        if (int rc = dalloc(c, 4 * (1 + oxd::WF_LIVE_RING), &q, "wavefront counters")) return rc;
        W.counters = (uint32_t *)q;
    }
    if (!W.host_live && hipHostMalloc((void **)&W.host_live, 4 * oxd::WF_LIVE_RING, hipHostMallocDefault) != hipSuccess) {
        W.host_live = nullptr;
        return fail(c, OXPT_E_BACKEND, "hip: pinned wavefront counters");
    }
    if (W.events.empty()) {
        W.events.assign(oxd::WF_LIVE_RING, nullptr);
        for (auto &e : W.events)
            if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) return fail(c, OXPT_E_BACKEND, "hip: events");
    }
    for (uint32_t done = 0; done < max_count; done += spd) {
        std::vector<oxd::Job> pass;
        std::vector<uint32_t> base;
        uint64_t n_items = 0;
        for (auto &J0 : jobs) {
            if (J0.count <= done) continue;
            oxd::Job J = J0;
            J.first = J0.first + done;
            J.count = std::min(spd, J0.count - done);
            base.push_back((uint32_t)n_items);
            n_items += (uint64_t)J.n_items * J.count;
            pass.push_back(J);
        }
        hipError_t e = hipMemcpyAsync(W.jobs, pass.data(), pass.size() * sizeof(oxd::Job), hipMemcpyHostToDevice, c->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(W.job_base, base.data(), base.size() * 4, hipMemcpyHostToDevice, c->stream);
        oxd::WfArgs A{};
        A.jobs = W.jobs;
        A.job_base = W.job_base;
        A.n_jobs = (uint32_t)pass.size();
        A.slots = W.slots;
        A.hits = W.hits;
        A.shadows = W.shadows;
        A.occluded = W.occluded;
        A.samples = W.samples;
        A.item_counter = W.counters;
        A.live = W.counters + 1;
        A.n_slots = (uint32_t)std::min<uint64_t>(W.n_slots, (n_items + 255) / 256 * 256);
        A.max_shadow = max_shadow;
        A.n_items = (uint32_t)n_items;
        if (e == hipSuccess) e = oxd::wf_launch_clear(A, c->stream);
        // Iterate until a shade pass finds no live slot. The check lags by a
        // few iterations so the host never waits on the newest launch.
        const uint32_t LAG = 6;
        for (uint32_t it = 0; e == hipSuccess; it++) {
            A.iter = it;
            e = oxd::wf_launch_iteration(A, c->tuning.wf_variant, c->stream);
            if (e != hipSuccess) break;
            uint32_t r = it & (oxd::WF_LIVE_RING - 1);
            e = hipMemcpyAsync(W.host_live + r, A.live + r, 4, hipMemcpyDeviceToHost, c->stream);
            if (e == hipSuccess) e = hipEventRecord(W.events[r], c->stream);
            W.iterations++;
            if (e == hipSuccess && it >= LAG) {
                uint32_t q = (it - LAG) & (oxd::WF_LIVE_RING - 1);
                if ((e = hipEventSynchronize(W.events[q])) != hipSuccess) break;
                if (W.host_live[q] == 0) break;
            }
        }
        for (uint32_t k = 0; e == hipSuccess && k < pass.size(); k++) e = oxd::wf_launch_reduce(A, k, pass[k].n_items, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) return hip_fail(c, "wavefront", e);
    }
    return OXPT_OK;
}

}  // namespace

namespace {

// Checks every chain of an atlas build against the tables' texture
// descriptors and the atlas shape, uploads the textures and the decode
// table, and runs the device atlas builder into `atlas`.
int build_atlas_on_device(oxpt_ctx *c, const oxpt_atlas_desc &ad, const std::vector<uint32_t> &tables,
                          uint16_t *atlas, uint32_t side, uint32_t layers) {
    if (!ad.decode_lut || (ad.n_chains && (!ad.chains || !ad.textures)))
        return fail(c, OXPT_E_ARG, "atlas build: chains, textures and the decode table are required");
    std::vector<oxd::AtlasChainDev> chains(ad.n_chains);
    std::vector<int> used(ad.n_textures, 0);
    for (size_t i = 0; i < ad.n_chains; i++) {
        const oxpt_atlas_chain &ch = ad.chains[i];
        if (ch.texture >= ad.n_textures) return fail(c, OXPT_E_ARG, "atlas build: texture index out of range");
        const oxpt_texture_desc &tx = ad.textures[ch.texture];
        if (!tx.rgba8 || tx.w == 0 || tx.h == 0) return fail(c, OXPT_E_ARG, "atlas build: empty texture");
        if ((size_t)ch.desc_offset + 8 > tables.size()) return fail(c, OXPT_E_ARG, "atlas build: descriptor outside the tables");
        const uint32_t *dw = &tables[ch.desc_offset];
        uint32_t levels = 1;
        for (uint32_t w = tx.w, h = tx.h; w > 1 || h > 1; levels++) {
            w = std::max(1u, w / 2);
            h = std::max(1u, h / 2);
        }
        if (dw[0] != levels || dw[2] != tx.w || dw[3] != tx.h || dw[4] > levels || dw[5] != levels - dw[4] ||
            (size_t)ch.desc_offset + 8 + 4ull * dw[5] > tables.size())
            return fail(c, OXPT_E_ARG, "atlas build: descriptor does not match its texture");
        oxd::AtlasChainDev &cd = chains[i];
        cd.w = tx.w;
        cd.h = tx.h;
        cd.srgb = ch.srgb ? 1u : 0u;
        cd.levels = levels;
        cd.skip = dw[4];
        cd.rects.assign(dw + 8, dw + 8 + 4 * dw[5]);
        uint32_t w = tx.w, h = tx.h;
        for (uint32_t li = 0; li < levels; li++) {
            if (li >= cd.skip) {
                const uint32_t *r = &cd.rects[4 * (li - cd.skip)];
                if (r[0] >= layers || r[3] != (w | h << 16) || (uint64_t)r[1] + w > side || (uint64_t)r[2] + h > side)
                    return fail(c, OXPT_E_ARG, "atlas build: level rectangle outside the atlas");
            }
            w = std::max(1u, w / 2);
            h = std::max(1u, h / 2);
        }
        used[ch.texture] = 1;
    }
    std::vector<void *> held;
    std::vector<size_t> held_bytes;
    auto release = [&]() {
        (void)hipStreamSynchronize(c->stream);
        for (size_t i = 0; i < held.size(); i++) dfree(c, held[i], held_bytes[i]);
    };
    std::vector<const uint8_t *> dev_tex(ad.n_textures, nullptr);
    int rc = OXPT_OK;
    for (size_t t = 0; t < ad.n_textures && rc == OXPT_OK; t++) {
        if (!used[t]) continue;
        const size_t bytes = (size_t)ad.textures[t].w * ad.textures[t].h * 4;
        void *p = nullptr;
        if ((rc = dalloc(c, bytes, &p, "atlas build: textures"))) break;
        held.push_back(p);
        held_bytes.push_back(bytes);
        dev_tex[t] = (const uint8_t *)p;
        hipError_t e = hipMemcpyAsync(p, ad.textures[t].rgba8, bytes, hipMemcpyHostToDevice, c->stream);
        if (e != hipSuccess) rc = hip_fail(c, "atlas build: texture upload", e);
    }
    void *lut = nullptr;
    if (rc == OXPT_OK && (rc = dalloc(c, 512 * 4, &lut, "atlas build: decode table")) == OXPT_OK) {
        held.push_back(lut);
        held_bytes.push_back(512 * 4);
        hipError_t e = hipMemcpyAsync(lut, ad.decode_lut, 512 * 4, hipMemcpyHostToDevice, c->stream);
        if (e != hipSuccess) rc = hip_fail(c, "atlas build: decode table upload", e);
    }
    if (rc == OXPT_OK) {
        for (size_t i = 0; i < chains.size(); i++) chains[i].rgba8 = dev_tex[ad.chains[i].texture];
        oxd::DevAlloc al;
        al.alloc = [c](size_t bytes, void **p, const char *what) { return dalloc(c, bytes, p, what); };
        al.release = [c](void *p, size_t bytes) { dfree(c, p, bytes); };
        std::string err;
        if ((rc = oxd::build_atlas(chains, (const float *)lut, atlas, side, layers, c->stream, al, err)))
            fail(c, rc, err);
    }
    release();
    return rc;
}

}  // namespace

extern "C" {

int oxpt_ctx_create(int ordinal, oxpt_ctx **out) {
    if (!out) return OXPT_E_ARG;
    *out = nullptr;
    std::lock_guard<std::mutex> lock(g_device_init);
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n <= 0) return OXPT_E_BACKEND;
    if (ordinal < 0 || ordinal >= n) return OXPT_E_BACKEND;
    auto *c = new oxpt_ctx();
    c->device = ordinal;
    if (hipSetDevice(ordinal) != hipSuccess) {
        delete c;
        return OXPT_E_BACKEND;
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, ordinal) != hipSuccess) {
        delete c;
        return OXPT_E_BACKEND;
    }
    c->n_cu = (uint32_t)prop.multiProcessorCount;
    c->base = c;
    c->dom = std::make_shared<Domain>();
    c->dom->budget = (uint64_t)prop.totalGlobalMem;
    std::string name = prop.name;
    if (name.empty()) {
        char buf[256] = {0};
        if (hipDeviceGetName(buf, sizeof buf, ordinal) == hipSuccess) name = buf;
    }
    if (name.empty()) name = "AMD GPU";
    c->summary = name + " (" + prop.gcnArchName + ", " + std::to_string(c->n_cu) + " CUs, HIP)";
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return OXPT_E_BACKEND;
    }
    void *w = nullptr;
    if (dalloc(c, 256, &w, "work counter") != OXPT_OK) {
        (void)hipStreamDestroy(c->stream);
        delete c;
        return OXPT_E_BACKEND;
    }
    c->d_work = static_cast<uint32_t *>(w);
    // This device's copy of the no-allocation scratch (once per device,
    // with the device current).
    if (ordinal < kNaMaxDevices) {
        std::lock_guard<std::mutex> lk(g_scratch_mu);
        if (!g_scratch_ptr[ordinal]) {
            void *ptr = nullptr;
            if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) == hipSuccess) g_scratch_ptr[ordinal] = ptr;
            else (void)hipGetLastError();
        }
    }
    *out = c;
    return OXPT_OK;
}

int oxpt_ctx_create_peer(oxpt_ctx *base, oxpt_ctx **out) {
    if (!base || !out) return OXPT_E_ARG;
    *out = nullptr;
    std::lock_guard<std::mutex> lock(g_device_init);
    if (setdev(base)) return OXPT_E_BACKEND;
    auto *c = new oxpt_ctx();
    c->device = base->device;
    c->n_cu = base->n_cu;
    c->summary = base->summary;
    c->tuning = base->tuning;
    c->base = base->base;
    c->dom = base->dom;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return fail(base, OXPT_E_BACKEND, "hip: peer stream");
    }
    void *w = nullptr;
    if (int rc = dalloc(c, 256, &w, "work counter")) {
        (void)hipStreamDestroy(c->stream);
        base->err = c->err;
        delete c;
        return rc;
    }
    c->d_work = static_cast<uint32_t *>(w);
    *out = c;
    return OXPT_OK;
}

void oxpt_ctx_destroy(oxpt_ctx *c) {
    if (!c) return;
    setdev(c);
    (void)hipStreamSynchronize(c->stream);
    fold_events(c, 0);
    dfree(c, c->d_work, 256);
    if (c->graph) (void)hipGraphExecDestroy(c->graph);
    {
        auto &W = c->wf;
        dfree(c, W.slots, W.slot_bytes);
        dfree(c, W.hits, W.hit_bytes);
        dfree(c, W.shadows, W.shadow_bytes);
        dfree(c, W.occluded, W.occ_bytes);
        dfree(c, W.samples, W.sample_bytes);
        dfree(c, W.jobs, W.job_bytes);
        dfree(c, W.job_base, W.base_bytes);
        dfree(c, W.counters, 4 * (1 + oxd::WF_LIVE_RING));
        if (W.host_live) (void)hipHostFree(W.host_live);
        for (auto &e : W.events)
            if (e) (void)hipEventDestroy(e);
    }
    dfree(c, c->d_jobs, c->job_cap * sizeof(oxd::Job));
    dfree(c, c->d_counters, c->job_cap * 8 + 4);
    dfree(c, c->d_samples, c->samples_bytes);
    dfree(c, c->d_nee_q, c->nee_q_bytes);
    dfree(c, c->d_out, c->d_out_bytes);
    (void)hipStreamDestroy(c->stream);
    delete c;
}

const char *oxpt_last_error(const oxpt_ctx *c) { return c ? c->err.c_str() : "null context"; }
const char *oxpt_ctx_summary(const oxpt_ctx *c) { return c ? c->summary.c_str() : ""; }
const char *oxpt_ctx_last_kernel(const oxpt_ctx *c) { return c ? c->last_kernel : ""; }

int oxpt_ctx_set_budget(oxpt_ctx *c, uint64_t bytes) {
    if (!c) return OXPT_E_ARG;
    std::lock_guard<std::mutex> lk(c->dom->mu);
    c->dom->budget = bytes;
    return OXPT_OK;
}
uint64_t oxpt_ctx_bytes_in_use(const oxpt_ctx *c) {
    if (!c) return 0;
    std::lock_guard<std::mutex> lk(c->dom->mu);
    return c->dom->in_use;
}

int oxpt_ctx_set_tuning(oxpt_ctx *c, const oxpt_tuning *t) {
    if (!c || !t) return OXPT_E_ARG;
    if (t->block) {
        if (t->block != 64 && t->block != 128 && t->block != 256) return fail(c, OXPT_E_ARG, "tuning: block must be 64, 128 or 256");
        c->tuning.block = t->block;
    }
    if (t->waves_per_cu) c->tuning.waves_per_cu = std::min(t->waves_per_cu, 64u);
    if (t->max_paths) c->tuning.max_paths = t->max_paths;
    if (t->kernel > 8) return fail(c, OXPT_E_ARG, "tuning: kernel must be 0..8");
    if (t->stack_lds != OXPT_STACK_PRIVATE && t->stack_lds > 64)
        return fail(c, OXPT_E_ARG, "tuning: stack_lds must be 0 (default), 1..64 or OXPT_STACK_PRIVATE");
    if (t->stack_lds != OXPT_STACK_PRIVATE && t->stack_lds != 0 &&
        (c->tuning.block / 64u) * t->stack_lds * 512u + kTableLds + kCdfLds > 65536u)
        return fail(c, OXPT_E_ARG, "tuning: stack_lds does not fit in 64 KiB of LDS per workgroup");
    if (t->bvh != 0 && t->bvh != 2) return fail(c, OXPT_E_ARG, "tuning: bvh must be 0 (4-wide) or 2 (binary)");
    if (t->tail > 16 && t->tail != OXPT_TAIL_NONE)
        return fail(c, OXPT_E_ARG, "tuning: tail must be 0 (default), 1..16 or OXPT_TAIL_NONE");
    if (t->early != 0 && t->early != OXPT_EARLY_OFF)
        return fail(c, OXPT_E_ARG, "tuning: early must be 0 (default) or OXPT_EARLY_OFF");
    if (t->opaque != 0 && t->opaque != OXPT_OPAQUE_OFF)
        return fail(c, OXPT_E_ARG, "tuning: opaque must be 0 (default) or OXPT_OPAQUE_OFF");
    c->tuning.kernel = t->kernel;
    c->tuning.wf_variant = t->wf_variant;
    c->tuning.stack_lds = t->stack_lds;
    c->tuning.bvh = t->bvh;
    c->tuning.tail = t->tail;
    c->tuning.early = t->early;
    c->tuning.opaque = t->opaque;
    return OXPT_OK;
}
int oxpt_ctx_set_graph(oxpt_ctx *c, int enable) {
    if (!c) return OXPT_E_ARG;
    c->use_graph = enable != 0;
    return OXPT_OK;
}

uint64_t oxpt_ctx_graph_launches(const oxpt_ctx *c) { return c ? c->graph_launches : 0; }

int oxpt_ctx_get_tuning(const oxpt_ctx *c, oxpt_tuning *t) {
    if (!c || !t) return OXPT_E_ARG;
    *t = c->tuning;
    return OXPT_OK;
}

int oxpt_scene_upload(oxpt_ctx *c, const oxpt_scene_desc *d, oxpt_scene **out) {
    if (!c || !d || !out) return OXPT_E_ARG;
    *out = nullptr;
    if (!d->geom || !d->attrs || !d->tables || d->geom_vec4 == 0 || d->attrs_vec4 == 0 || d->table_words == 0)
        return fail(c, OXPT_E_ARG, "scene: geom, attrs and tables are required");
    if (setdev(c)) return OXPT_E_BACKEND;
    auto *s = new oxpt_scene();
    s->ctx = c->base;
    s->table_words = d->table_words;
    s->geom_vec4 = d->geom_vec4;
    s->attrs_vec4 = d->attrs_vec4;
    s->atlas_side = std::max(d->atlas_side, 1u);
    s->atlas_layers = std::max(d->atlas_layers, 1u);
    size_t gb = d->geom_vec4 * 16, ab = d->attrs_vec4 * 16, tb = d->table_words * 4;
    size_t xb = d->atlas ? (size_t)s->atlas_side * s->atlas_side * s->atlas_layers * 8 : 8;
    s->atlas_words = xb / 2;
    void *p;
    int rc;
    if ((rc = dalloc(c, gb, &p, "BVH + triangle buffer"))) goto err;
    s->geom = (float4 *)p;
    if ((rc = dalloc(c, ab, &p, "vertex attribute buffer"))) goto err;
    s->attrs = (float4 *)p;
    if ((rc = dalloc(c, tb, &p, "scene tables"))) goto err;
    s->tables = (uint32_t *)p;
    if ((rc = dalloc(c, xb, &p, "texture atlas"))) goto err;
    s->atlas = (uint16_t *)p;
    s->bytes = gb + ab + tb + xb;
    s->host_tables.assign(d->tables, d->tables + d->table_words);
    {
        hipError_t e;
        if ((e = hipMemcpyAsync(s->geom, d->geom, gb, hipMemcpyHostToDevice, c->stream)) != hipSuccess ||
            (e = hipMemcpyAsync(s->attrs, d->attrs, ab, hipMemcpyHostToDevice, c->stream)) != hipSuccess ||
            (e = hipMemcpyAsync(s->tables, d->tables, tb, hipMemcpyHostToDevice, c->stream)) != hipSuccess) {
            rc = hip_fail(c, "scene upload", e);
            goto err;
        }
        if (d->atlas) {
            if ((e = hipMemcpyAsync(s->atlas, d->atlas, xb, hipMemcpyHostToDevice, c->stream)) != hipSuccess) {
                rc = hip_fail(c, "atlas upload", e);
                goto err;
            }
        } else if ((e = hipMemsetAsync(s->atlas, 0, xb, c->stream)) != hipSuccess) {
            rc = hip_fail(c, "atlas clear", e);
            goto err;
        }
        if ((e = hipStreamSynchronize(c->stream)) != hipSuccess) {
            rc = hip_fail(c, "scene upload", e);
            goto err;
        }
    }
    *out = s;
    return OXPT_OK;
err:
    oxpt_scene_destroy(s);
    return rc;
}

int oxpt_scene_build(oxpt_ctx *c, const oxpt_soup_desc *d, uint32_t *emissive_slots, oxpt_scene **out,
                     oxpt_build_info *info) {
    if (!c || !d || !out) return OXPT_E_ARG;
    *out = nullptr;
    if (!d->pos || !d->material || !d->flags || !d->tables || d->table_words == 0)
        return fail(c, OXPT_E_ARG, "scene build: positions, materials, flags and tables are required");
    if (d->n_tris == 0 || d->n_tris > oxd::kBuildMaxTris)
        return fail(c, OXPT_E_INVALID, "scene build: the device builder takes 1 to 16777216 triangles");
    if (d->n_emissive > 0 && (!d->light_of || (size_t)d->elight_offset + 4ull * d->n_emissive > d->table_words))
        return fail(c, OXPT_E_ARG, "scene build: emissive records need light_of and fit the tables");
    if (d->atlas && d->atlas_build) return fail(c, OXPT_E_ARG, "scene build: atlas texels or atlas_build, not both");
    if (d->material_emission && (d->light_of || d->n_emissive || d->n_materials == 0))
        return fail(c, OXPT_E_ARG, "scene build: lights are either found on the device or given, not both");
    const size_t n = d->n_tris;
    uint32_t any = 0;
    for (size_t i = 0; i < n; i++) any |= d->flags[i];
    struct Arr {
        const void *src;
        size_t bytes;
        uint32_t flag;
        const char *what;
    };
    const Arr arrs[9] = {{d->pos, 36 * n, 0, "positions"},
                         {d->normals, 36 * n, OXPT_F_NORMALS, "normals"},
                         {d->tangents, 48 * n, OXPT_F_TANGENTS, "tangents"},
                         {d->uv0, 24 * n, OXPT_F_UV0, "uv0"},
                         {d->uv1, 24 * n, OXPT_F_UV1, "uv1"},
                         {d->colors, 48 * n, OXPT_F_COLORS, "colors"},
                         {d->material, 4 * n, 0, "materials"},
                         {d->flags, n, 0, "flags"},
                         {d->light_of, 4 * n, 0, "light ids"}};
    for (const Arr &a : arrs)
        if (a.flag && (any & a.flag) && !a.src)
            return fail(c, OXPT_E_ARG, std::string("scene build: triangles select ") + a.what + " but none given");
    if (setdev(c)) return OXPT_E_BACKEND;
    const auto t0 = std::chrono::steady_clock::now();
    // The soup and the emissive slot words live on the device for the build only.
    void *dev[11] = {};
    size_t dev_bytes[11] = {};
    auto release = [&]() {
        (void)hipStreamSynchronize(c->stream);
        for (int i = 0; i < 11; i++) dfree(c, dev[i], dev_bytes[i]);
    };
    int rc = OXPT_OK;
    for (int i = 0; i < 9 && rc == OXPT_OK; i++) {
        const Arr &a = arrs[i];
        if (!a.src || (a.flag && !(any & a.flag))) continue;
        dev_bytes[i] = a.bytes;
        if ((rc = dalloc(c, a.bytes, &dev[i], "scene build: soup"))) break;
        hipError_t e = hipMemcpyAsync(dev[i], a.src, a.bytes, hipMemcpyHostToDevice, c->stream);
        if (e != hipSuccess) rc = hip_fail(c, "scene build: soup upload", e);
    }
    if (rc == OXPT_OK && d->n_emissive > 0) {
        dev_bytes[9] = 4ull * d->n_emissive;
        rc = dalloc(c, dev_bytes[9], &dev[9], "scene build: emissive slots");
    }
    if (rc == OXPT_OK && d->material_emission) {
        dev_bytes[10] = 4ull * d->n_materials;
        if ((rc = dalloc(c, dev_bytes[10], &dev[10], "scene build: material emission")) == OXPT_OK) {
            hipError_t e = hipMemcpyAsync(dev[10], d->material_emission, dev_bytes[10], hipMemcpyHostToDevice,
                                          c->stream);
            if (e != hipSuccess) rc = hip_fail(c, "scene build: material emission upload", e);
        }
    }
    if (rc == OXPT_OK) {
        hipError_t e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) rc = hip_fail(c, "scene build: soup upload", e);
    }
    if (rc) {
        release();
        return rc;
    }
    const double upload_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    oxd::SoupDev sd;
    sd.n = (uint32_t)n;
    sd.pos = (const float *)dev[0];
    sd.normals = (const float *)dev[1];
    sd.tangents = (const float *)dev[2];
    sd.uv0 = (const float *)dev[3];
    sd.uv1 = (const float *)dev[4];
    sd.colors = (const float *)dev[5];
    sd.material = (const uint32_t *)dev[6];
    sd.flags = (const uint8_t *)dev[7];
    sd.light_of = (const uint32_t *)dev[8];
    sd.mat_lum = (const float *)dev[10];
    sd.n_mat = d->material_emission ? d->n_materials : 0u;
    oxd::DevAlloc al;
    al.alloc = [c](size_t bytes, void **p, const char *what) { return dalloc(c, bytes, p, what); };
    al.release = [c](void *p, size_t bytes) { dfree(c, p, bytes); };
    oxd::BuiltGeometry bg;
    std::string err;
    rc = oxd::build_geometry(sd, d->n_emissive, (uint32_t *)dev[9], c->stream, al, &bg, err);
    if (rc) {
        release();
        return fail(c, rc, err);
    }
    // Tables with the emissive slot words filled in, then the atlas.
    std::vector<uint32_t> tables(d->tables, d->tables + d->table_words);
    // Lights found on the device: their records follow the tables.
    const uint32_t elight_offset = d->material_emission ? (uint32_t)d->table_words : d->elight_offset;
    tables.insert(tables.end(), bg.emissive_records.begin(), bg.emissive_records.end());
    std::vector<uint32_t> slots(d->n_emissive);
    if (d->n_emissive > 0) {
        hipError_t e = hipMemcpy(slots.data(), dev[9], 4ull * d->n_emissive, hipMemcpyDeviceToHost);
        if (e != hipSuccess) rc = hip_fail(c, "scene build: emissive slots", e);
        for (uint32_t i = 0; i < d->n_emissive; i++) tables[d->elight_offset + 4ull * i + 3] = slots[i];
    }
    release();
    auto *s = new oxpt_scene();
    s->ctx = c->base;
    s->geom = bg.geom;
    s->attrs = bg.attrs;
    s->geom_vec4 = bg.geom_vec4;
    s->attrs_vec4 = bg.attrs_vec4;
    s->table_words = tables.size();
    s->atlas_side = std::max(d->atlas_side, 1u);
    s->atlas_layers = std::max(d->atlas_layers, 1u);
    const size_t tb = tables.size() * 4;
    const size_t xb =
        (d->atlas || d->atlas_build) ? (size_t)s->atlas_side * s->atlas_side * s->atlas_layers * 8 : 8;
    s->atlas_words = xb / 2;
    s->bytes = bg.geom_vec4 * 16 + bg.attrs_vec4 * 16 + tb + xb;
    void *p = nullptr;
    if (rc == OXPT_OK && (rc = dalloc(c, tb, &p, "scene tables")) == OXPT_OK) s->tables = (uint32_t *)p;
    if (rc == OXPT_OK && (rc = dalloc(c, xb, &p, "texture atlas")) == OXPT_OK) s->atlas = (uint16_t *)p;
    double atlas_ms = 0.0;
    if (rc == OXPT_OK) {
        s->host_tables = tables;
        hipError_t e = hipMemcpyAsync(s->tables, tables.data(), tb, hipMemcpyHostToDevice, c->stream);
        if (e == hipSuccess && !d->atlas_build)
            e = d->atlas ? hipMemcpyAsync(s->atlas, d->atlas, xb, hipMemcpyHostToDevice, c->stream)
                         : hipMemsetAsync(s->atlas, 0, xb, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) rc = hip_fail(c, "scene build: tables / atlas upload", e);
    }
    if (rc == OXPT_OK && d->atlas_build) {
        const auto ta = std::chrono::steady_clock::now();
        rc = build_atlas_on_device(c, *d->atlas_build, tables, s->atlas, s->atlas_side, s->atlas_layers);
        atlas_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - ta).count();
    }
    if (rc) {
        // Account the build's buffers to the scene so destroy releases them.
        if (!s->tables) s->bytes -= tb;
        if (!s->atlas) s->bytes -= xb;
        oxpt_scene_destroy(s);
        return rc;
    }
    if (emissive_slots && d->n_emissive > 0) std::memcpy(emissive_slots, slots.data(), 4ull * d->n_emissive);
    if (info) {
        info->node_count = bg.node_count;
        info->tri_base = bg.tri_base;
        info->wide_base = bg.wide_base;
        info->bvh_depth = bg.bvh_depth;
        info->wide_depth = bg.wide_depth;
        info->pad = 0;
        info->upload_ms = upload_ms;
        info->build_ms = bg.ms;
        info->atlas_ms = atlas_ms;
        info->n_emissive = d->material_emission ? bg.n_emissive : d->n_emissive;
        info->elight_offset = elight_offset;
    }
    *out = s;
    return OXPT_OK;
}

int oxpt_scene_read(oxpt_ctx *c, const oxpt_scene *s, int which, void *host_out, size_t bytes, size_t *words) {
    if (!c || !s || !same_domain(s->ctx, c)) return OXPT_E_ARG;
    const void *src;
    size_t count, elem;
    switch (which) {
        case 0: src = s->geom; count = s->geom_vec4; elem = 16; break;
        case 1: src = s->attrs; count = s->attrs_vec4; elem = 16; break;
        case 2: src = s->tables; count = s->table_words; elem = 4; break;
        case 3: src = s->atlas; count = s->atlas_words; elem = 2; break;
        default: return OXPT_E_ARG;
    }
    if (words) *words = count;
    if (!host_out) return OXPT_OK;
    if (bytes < count * elem) return fail(c, OXPT_E_ARG, "scene read: buffer too small");
    if (setdev(c)) return OXPT_E_BACKEND;
    hipError_t e = hipStreamSynchronize(c->stream);
    if (e == hipSuccess) e = hipMemcpy(host_out, src, count * elem, hipMemcpyDeviceToHost);
    return e == hipSuccess ? OXPT_OK : hip_fail(c, "scene read", e);
}

int oxpt_scene_set_tables(oxpt_ctx *c, oxpt_scene *s, const uint32_t *tables, size_t words) {
    if (!c || !s || !tables || words == 0 || !same_domain(s->ctx, c)) return OXPT_E_ARG;
    if (setdev(c)) return OXPT_E_BACKEND;
    hipError_t e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return hip_fail(c, "tables update", e);
    if (words != s->table_words) {
        void *p = nullptr;
        if (int rc = dalloc(c, words * 4, &p, "scene tables")) return rc;
        dfree(c, s->tables, s->table_words * 4);
        s->bytes = s->bytes - s->table_words * 4 + words * 4;
        s->tables = (uint32_t *)p;
        s->table_words = words;
    }
    e = hipMemcpyAsync(s->tables, tables, words * 4, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e == hipSuccess) s->host_tables.assign(tables, tables + words);
    return e == hipSuccess ? OXPT_OK : hip_fail(c, "tables update", e);
}

int oxpt_scene_set_tables_env(oxpt_ctx *c, oxpt_scene *s, const uint32_t *tables, size_t words,
                              const oxpt_env_desc *env) {
    if (!c || !s || !tables || words == 0 || !env || !same_domain(s->ctx, c)) return OXPT_E_ARG;
    if (!env->rgb || !env->row_weight || env->w == 0 || env->h == 0 || (uint64_t)env->w * env->h > (1ull << 28))
        return fail(c, OXPT_E_ARG, "environment: rgb, row weights and 1 .. 2^28 texels are required");
    const size_t w = env->w, h = env->h, n = w * h;
    const size_t env_words = (h + 1) + h * (w + 1) + n + 3 * n;
    const size_t total = words + env_words;
    if (setdev(c)) return OXPT_E_BACKEND;
    hipError_t e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return hip_fail(c, "tables update", e);
    if (total != s->table_words) {
        void *p = nullptr;
        if (int rc = dalloc(c, total * 4, &p, "scene tables")) return rc;
        dfree(c, s->tables, s->table_words * 4);
        s->bytes = s->bytes - s->table_words * 4 + total * 4;
        s->tables = (uint32_t *)p;
        s->table_words = total;
    }
    void *rgb = nullptr, *roww = nullptr;
    int rc = dalloc(c, n * 12, &rgb, "environment: texels");
    if (rc == OXPT_OK) rc = dalloc(c, h * 8, &roww, "environment: row weights");
    if (rc == OXPT_OK) {
        e = hipMemcpyAsync(s->tables, tables, words * 4, hipMemcpyHostToDevice, c->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(rgb, env->rgb, n * 12, hipMemcpyHostToDevice, c->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(roww, env->row_weight, h * 8, hipMemcpyHostToDevice, c->stream);
        if (e != hipSuccess) rc = hip_fail(c, "environment upload", e);
    }
    if (rc == OXPT_OK) {
        oxd::DevAlloc al;
        al.alloc = [c](size_t bytes, void **p, const char *what) { return dalloc(c, bytes, p, what); };
        al.release = [c](void *p, size_t bytes) { dfree(c, p, bytes); };
        std::string err;
        if ((rc = oxd::build_env((const float *)rgb, env->w, env->h, env->intensity, (const double *)roww,
                                 s->tables + words, c->stream, al, err)))
            fail(c, rc, err);
    }
    (void)hipStreamSynchronize(c->stream);
    dfree(c, rgb, n * 12);
    dfree(c, roww, h * 8);
    if (rc == OXPT_OK) s->host_tables.assign(tables, tables + words);
    return rc;
}

uint64_t oxpt_scene_bytes(const oxpt_scene *s) { return s ? s->bytes : 0; }

void oxpt_scene_destroy(oxpt_scene *s) {
    if (!s) return;
    oxpt_ctx *c = s->ctx;
    setdev(c);
    (void)hipStreamSynchronize(c->stream);
    (void)hipFree(s->geom);
    (void)hipFree(s->attrs);
    (void)hipFree(s->tables);
    (void)hipFree(s->atlas);
    {
        std::lock_guard<std::mutex> lk(c->dom->mu);
        c->dom->in_use -= std::min<uint64_t>(c->dom->in_use, s->bytes);
    }
    delete s;
}

int oxpt_accum_create(oxpt_ctx *c, uint32_t w, uint32_t h, oxpt_accum **out) {
    if (!c || !out) return OXPT_E_ARG;
    *out = nullptr;
    if (w == 0 || h == 0) return fail(c, OXPT_E_INVALID, "accumulator: zero dimension");
    if (setdev(c)) return OXPT_E_BACKEND;
    auto *a = new oxpt_accum();
    a->ctx = c->base;
    a->w = w;
    a->h = h;
    a->bytes = (size_t)w * h * 16;
    void *p;
    int rc = dalloc(c, a->bytes, &p, "accumulation buffer");
    if (rc) {
        delete a;
        return rc;
    }
    a->d = (float4 *)p;
    hipError_t e = hipMemsetAsync(a->d, 0, a->bytes, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) {
        oxpt_accum_destroy(a);
        return hip_fail(c, "accumulator clear", e);
    }
    *out = a;
    return OXPT_OK;
}

void oxpt_accum_destroy(oxpt_accum *a) {
    if (!a) return;
    setdev(a->ctx);
    (void)hipStreamSynchronize(a->ctx->stream);
    dfree(a->ctx, a->d, a->bytes);
    delete a;
}

int oxpt_accum_clear(oxpt_ctx *c, oxpt_accum *a) {
    if (!c || !a || !same_domain(a->ctx, c)) return OXPT_E_ARG;
    if (setdev(c)) return OXPT_E_BACKEND;
    hipError_t e = hipMemsetAsync(a->d, 0, a->bytes, c->stream);
    return e == hipSuccess ? OXPT_OK : hip_fail(c, "accumulator clear", e);
}

int oxpt_accum_read(oxpt_ctx *c, const oxpt_accum *a, float *host, size_t bytes) {
    if (!c || !a || !host || !same_domain(a->ctx, c)) return OXPT_E_ARG;
    if (bytes < a->bytes) return fail(c, OXPT_E_ARG, "accumulator read: buffer too small");
    if (setdev(c)) return OXPT_E_BACKEND;
    hipError_t e = hipMemcpyAsync(host, a->d, a->bytes, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    return e == hipSuccess ? OXPT_OK : hip_fail(c, "accumulator read", e);
}

int oxpt_accum_write(oxpt_ctx *c, oxpt_accum *a, const float *host, size_t bytes) {
    if (!c || !a || !host || !same_domain(a->ctx, c)) return OXPT_E_ARG;
    if (bytes < a->bytes) return fail(c, OXPT_E_ARG, "accumulator write: buffer too small");
    if (setdev(c)) return OXPT_E_BACKEND;
    hipError_t e = hipMemcpyAsync(a->d, host, a->bytes, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    return e == hipSuccess ? OXPT_OK : hip_fail(c, "accumulator write", e);
}

// Grows the per-path sample buffer (batched launches and the persistent
// kernel's tail) to at least `slots` float4s. A captured graph refers to the
// old buffer, so it is dropped.
// NEE queue of the early form with punctual lights: nee_queue_words(n_punct)
// words for each lane of a launch of `lanes` threads.
static int ensure_queue(oxpt_ctx *c, uint64_t lanes, uint32_t n_punct) {
    const uint64_t bytes = lanes * oxd::nee_queue_words(n_punct) * 4u;
    if (c->nee_q_bytes >= bytes) return OXPT_OK;
    (void)hipStreamSynchronize(c->stream);
    if (c->graph) {
        (void)hipGraphExecDestroy(c->graph);
        c->graph = nullptr;
    }
    dfree(c, c->d_nee_q, c->nee_q_bytes);
    c->d_nee_q = nullptr;
    c->nee_q_bytes = 0;
    void *p = nullptr;
    if (int rc = dalloc(c, bytes, &p, "NEE queue")) return rc;
    c->d_nee_q = (float *)p;
    c->nee_q_bytes = bytes;
    return OXPT_OK;
}

// Work split of one batched round whose jobs are sorted largest first:
// each job's first tiles are traced as whole-pixel runs and its last
// tail16/16 of tiles (rounded up) one path per claim, as the persistent
// megakernel traces a frame. A job keeps pixel runs only when
//  - its tail has at least one item per two resident lanes: runs end at
//    times spread over one run's length after the last pixel is claimed,
//    and a wave stays on the job until its last run ends, so its lanes done
//    first need about half a run's worth of tail paths each; and
//  - its tail and the jobs after it hold at least one run's worth of paths
//    per resident lane: otherwise the launch would end waiting on the runs.
// Other jobs are traced path by path throughout (split = 0), jobs of one
// sample as one-sample runs. Places each job's tail samples in the sample
// buffer and returns the slots the round needs.
static uint64_t split_round(oxd::Job *jobs, size_t nj, uint32_t tail16, uint64_t lanes) {
    uint64_t after = 0;  // paths of the jobs after job k
    for (size_t k = nj; k-- > 0;) {
        oxd::Job &J = jobs[k];
        const uint32_t tiles = J.n_items / 64u;
        uint32_t tail_tiles = tiles;
        if (J.count == 1 || tail16 == 0) {
            tail_tiles = 0;  // one-sample runs, or every pixel a run (OXPT_TAIL_NONE)
        } else if (tail16 < 16) {
            const uint64_t t = ((uint64_t)tiles * tail16 + 15u) / 16u;
            if (t * 128u >= lanes && t * 64u * J.count + after >= (uint64_t)J.count * lanes) tail_tiles = (uint32_t)t;
        }
        J.split = (tiles - tail_tiles) * 64u;
        J.tail_work = tail_tiles * 64u * J.count;
        after += (uint64_t)J.n_items * J.count;
    }
    uint64_t slots = 0;
    for (size_t k = 0; k < nj; k++) {
        jobs[k].sample_base = (uint32_t)slots;
        slots += jobs[k].tail_work;
    }
    return slots;
}

// Lanes resident at once in a batched launch of the context's form.
static uint64_t resident_lanes(const oxpt_ctx *c) {
    const uint32_t k = c->tuning.kernel == 0 ? kEarlyDefaultForm : c->tuning.kernel;
    return (uint64_t)c->n_cu * 4u * 64u * form_waves(k);
}

// Job fields of one batched item for a round adding `count` samples from
// `first` (the split is set by split_round).
static oxd::Job make_job(const oxpt_ctx *c, const oxpt_batch_item &it, uint32_t first, uint32_t count) {
    oxd::Job J;
    std::memset(&J, 0, sizeof J);
    J.P = it.params;
    if (c->tuning.bvh == 2) J.P.offs2[3] = 0;
    J.geom = it.scene->geom;
    J.attrs = it.scene->attrs;
    J.tables = it.scene->tables;
    J.atlas = it.scene->atlas;
    J.accum = it.accum->d;
    J.atlas_side = it.scene->atlas_side;
    J.first = first;
    J.count = count;
    const uint32_t tx = (it.accum->w + 7u) / 8u, ty = (it.accum->h + 7u) / 8u;
    J.n_items = 64u * tx * ty;
    J.split = J.n_items;
    const oxd::UDiv dt = oxd::udiv_magic(64u * count), dx = oxd::udiv_magic(tx);
    J.tile_m = dt.m;
    J.tx_m = dx.m;
    J.shifts = dt.l | dx.l << 8;
    return J;
}

static int ensure_samples(oxpt_ctx *c, uint64_t slots) {
    if (c->samples_bytes >= slots * 16) return OXPT_OK;
    (void)hipStreamSynchronize(c->stream);
    if (c->graph) {
        (void)hipGraphExecDestroy(c->graph);
        c->graph = nullptr;
    }
    dfree(c, c->d_samples, c->samples_bytes);
    c->d_samples = nullptr;
    c->samples_bytes = 0;
    void *p = nullptr;
    if (int rc = dalloc(c, slots * 16, &p, "per-path sample buffer")) return rc;
    c->d_samples = (float4 *)p;
    c->samples_bytes = slots * 16;
    return OXPT_OK;
}

int oxpt_refine(oxpt_ctx *c, const oxpt_scene *s, oxpt_accum *a, const oxpt_params *P, uint32_t first, uint32_t count) {
    if (!c || !s || !a || !P || !same_domain(s->ctx, c) || !same_domain(a->ctx, c)) return OXPT_E_ARG;
    if (P->dims[0] != a->w || P->dims[1] != a->h) return fail(c, OXPT_E_ARG, "refine: params size differs from the accumulator");
    if (count == 0) return OXPT_OK;
    if (int rc = check_params(c, s, P)) return rc;
    if (setdev(c)) return OXPT_E_BACKEND;
    // A frame with fewer pixels than the GPU holds resident lanes would
    // leave most lanes idle with one pixel per lane: the default form then
    // runs as a one-job batch, whose work item is a single path (each
    // pixel's samples are still added in order, so the image is the same).
    const uint64_t resident_lanes = (uint64_t)c->n_cu * 4u * form_waves(0) * 64u;
    if (c->tuning.kernel == 0 && (uint64_t)a->w * a->h < resident_lanes && count > 1) {
        oxpt_batch_item it;
        it.scene = s;
        it.accum = a;
        it.params = *P;
        it.first = first;
        it.count = count;
        return oxpt_refine_batch(c, &it, 1);
    }
    if (c->tuning.kernel == 7) {
        oxd::Job J;
        std::memset(&J, 0, sizeof J);
        J.P = *P;
        if (c->tuning.bvh == 2) J.P.offs2[3] = 0;
        J.geom = s->geom;
        J.attrs = s->attrs;
        J.tables = s->tables;
        J.atlas = s->atlas;
        J.accum = a->d;
        J.atlas_side = s->atlas_side;
        J.first = first;
        J.count = count;
        J.n_items = 64u * J.tiles_x() * ((a->h + 7u) / 8u);
        EventPair ep{};
        bool timed = hipEventCreate(&ep.a) == hipSuccess && hipEventCreate(&ep.b) == hipSuccess;
        if (timed) (void)hipEventRecord(ep.a, c->stream);
        c->last_kernel = "wavefront";
        int rc = wf_run(c, {J});
        if (timed) {
            (void)hipEventRecord(ep.b, c->stream);
            c->events.push_back(ep);
            c->launches++;
        }
        return rc;
    }
    uint64_t pixels = (uint64_t)a->w * a->h;
    uint64_t per = std::max<uint64_t>(1, c->tuning.max_paths / std::max<uint64_t>(pixels, 1));
    uint32_t block = c->tuning.block;
    uint32_t grid = std::max(1u, c->n_cu * c->tuning.waves_per_cu / (block / 64u));
    // The early-continuation forms apply to scenes with up to 31 punctual
    // lights (beyond none, through the NEE queue); without punctual lights
    // their default runs at 5 waves per SIMD (fewer live registers than the
    // general form, which keeps 4).
    const uint32_t mode = trace_mode(c->tuning, P->cfg[2]);
    const bool opaque = mode != 0 && opaque_walk(c->tuning, s->host_tables, *P);
    oxpt_tuning form = c->tuning;
    if (form.kernel == 0 && mode == 1) form.kernel = kEarlyRefineForm;
    if (form.kernel == 0 && mode == 2) form.kernel = kQueueDefaultForm;
    if (mode == 2)
        if (int rc = ensure_queue(c, (uint64_t)grid * block, P->cfg[2])) return rc;
    const uint32_t lds = stack_lds_entries(form, block, false);
    // Persistent forms: the last tail/16 of the frame's tiles are traced one
    // path per claim (a wave's lanes on one tile at one sample, so the wave
    // stays coherent), and lanes that run out of pixels take tail paths
    // instead of idling while the rest of their wave finishes whole runs.
    const bool persistent = c->tuning.kernel != 1;
    const uint32_t tail16 = c->tuning.tail == 0 ? kDefaultTail : (c->tuning.tail == OXPT_TAIL_NONE ? 0u : c->tuning.tail);
    uint64_t end = (uint64_t)first + count;
    for (uint64_t f = first; f < end;) {
        uint32_t n = (uint32_t)std::min<uint64_t>(per, end - f);
        oxd::TraceArgs args = trace_args(c, s, a, P);
        args.first = (uint32_t)f;
        args.count = n;
        if (persistent && n > 1 && tail16 > 0) {
            const uint64_t tiles = args.n_items / 64u;
            uint64_t tail_tiles = (tiles * tail16 + 15u) / 16u;
            tail_tiles = std::min<uint64_t>(tail_tiles, kTailSlots / 64u / n);
            const uint64_t tail_items = tail_tiles * 64u;
            if (tail_items > 0) {
                if (int rc = ensure_samples(c, tail_items * n)) return rc;
                args.split = args.n_items - (uint32_t)tail_items;
                args.tail_work = (uint32_t)(tail_items * n);
                args.samples = c->d_samples;
            }
        }
        EventPair ep{};
        bool timed = hipEventCreate(&ep.a) == hipSuccess && hipEventCreate(&ep.b) == hipSuccess;
        if (timed) (void)hipEventRecord(ep.a, c->stream);
        args.nee_q = c->d_nee_q;
        c->last_kernel = c->tuning.kernel == 1 ? "trace_tiles"
                         : mode == 2           ? (opaque ? "trace_persistent(early, queue, opaque)" : "trace_persistent(early, queue)")
                         : mode == 1           ? (opaque ? "trace_persistent(early, opaque)" : "trace_persistent(early)")
                                               : "trace_persistent";
        hipError_t e = oxd::launch_trace(args, form.kernel, block, grid, lds, mode, opaque, c->stream);
        if (e != hipSuccess) return hip_fail(c, "trace launch", e);
        if (timed) {
            (void)hipEventRecord(ep.b, c->stream);
            c->events.push_back(ep);
            c->launches++;
        }
        if (c->events.size() > 4096) fold_events(c, 2048);
        f += n;
    }
    return OXPT_OK;
}

int oxpt_refine_batch(oxpt_ctx *c, const oxpt_batch_item *items, size_t n) {
    if (!c || (!items && n)) return OXPT_E_ARG;
    if (n == 0) return OXPT_OK;
    // Validate every item; the Sobol' matrices and sheen LUT (staged once
    // per workgroup) must be the same words in every job.
    const oxpt_batch_item &i0 = items[0];
    for (size_t i = 0; i < n; i++) {
        const oxpt_batch_item &it = items[i];
        if (!it.scene || !it.accum || !same_domain(it.scene->ctx, c) || !same_domain(it.accum->ctx, c)) return fail(c, OXPT_E_ARG, "batch: item belongs to another context");
        if (it.params.dims[0] != it.accum->w || it.params.dims[1] != it.accum->h)
            return fail(c, OXPT_E_ARG, "batch: params size differs from the accumulator");
        if (int rc = check_params(c, it.scene, &it.params)) return rc;
        for (size_t k = 0; k < i; k++)
            if (items[k].accum == it.accum) return fail(c, OXPT_E_ARG, "batch: two items share an accumulator");
        const uint32_t *a = it.scene->host_tables.data(), *b = i0.scene->host_tables.data();
        if (std::memcmp(a + it.params.offs0[0], b + i0.params.offs0[0], OXPT_SOBOL_WORDS * 4) != 0 ||
            std::memcmp(a + it.params.offs0[1], b + i0.params.offs0[1], OXPT_SHEEN_WORDS * 4) != 0)
            return fail(c, OXPT_E_ARG, "batch: jobs carry different Sobol' / sheen tables");
    }
    if (setdev(c)) return OXPT_E_BACKEND;
    if (c->job_cap < n) {
        (void)hipStreamSynchronize(c->stream);
        if (c->graph) {
            (void)hipGraphExecDestroy(c->graph);
            c->graph = nullptr;
        }
        dfree(c, c->d_jobs, c->job_cap * sizeof(oxd::Job));
        dfree(c, c->d_counters, c->job_cap * 8 + 4);
        c->d_jobs = nullptr;
        c->d_counters = nullptr;
        c->job_cap = 0;
        size_t cap = std::max<size_t>(n, 256);
        void *p = nullptr;
        if (int rc = dalloc(c, cap * sizeof(oxd::Job), &p, "batch job table")) return rc;
        c->d_jobs = (oxd::Job *)p;
        if (int rc = dalloc(c, cap * 8 + 4, &p, "batch counters")) {
            dfree(c, c->d_jobs, cap * sizeof(oxd::Job));
            c->d_jobs = nullptr;
            return rc;
        }
        c->d_counters = (uint32_t *)p;
        c->job_cap = cap;
    }
    if (c->tuning.kernel == 7) {
        std::vector<oxd::Job> jobs;
        for (size_t i = 0; i < n; i++) {
            const oxpt_batch_item &it = items[i];
            if (it.count == 0) continue;
            oxd::Job J;
            std::memset(&J, 0, sizeof J);
            J.P = it.params;
            if (c->tuning.bvh == 2) J.P.offs2[3] = 0;
            J.geom = it.scene->geom;
            J.attrs = it.scene->attrs;
            J.tables = it.scene->tables;
            J.atlas = it.scene->atlas;
            J.accum = it.accum->d;
            J.atlas_side = it.scene->atlas_side;
            J.first = it.first;
            J.count = it.count;
            J.n_items = 64u * J.tiles_x() * ((it.accum->h + 7u) / 8u);
            jobs.push_back(J);
        }
        if (jobs.empty()) return OXPT_OK;
        EventPair ep{};
        bool timed = hipEventCreate(&ep.a) == hipSuccess && hipEventCreate(&ep.b) == hipSuccess;
        if (timed) (void)hipEventRecord(ep.a, c->stream);
        c->last_kernel = "wavefront";
        int rc = wf_run(c, jobs);
        if (timed) {
            (void)hipEventRecord(ep.b, c->stream);
            c->events.push_back(ep);
            c->launches++;
        }
        return rc;
    }
    // Rounds: each round adds at most `spd` samples to every item that still
    // has samples left, so one launch stays under max_paths paths.
    uint32_t max_count = 0;
    for (size_t i = 0; i < n; i++) max_count = std::max(max_count, items[i].count);
    // Samples per round: the most that keep the first (largest) round within
    // max_paths, counting each item's own samples (an item of fewer samples
    // than the round takes adds only its own).
    auto round_paths = [&](uint32_t spr) {
        uint64_t t = 0;
        for (size_t i = 0; i < n; i++) t += (uint64_t)items[i].accum->w * items[i].accum->h * std::min(spr, items[i].count);
        return t;
    };
    uint32_t spd = 1;
    for (uint32_t lo = 1, hi = max_count; lo <= hi;) {
        const uint32_t mid = lo + (hi - lo) / 2;
        if (round_paths(mid) <= c->tuning.max_paths) {
            spd = mid;
            lo = mid + 1;
        } else {
            hi = mid - 1;
        }
    }
    uint32_t block = c->tuning.block;
    uint32_t grid = std::max(1u, c->n_cu * c->tuning.waves_per_cu / (block / 64u));
    const uint32_t tail16 = c->tuning.tail == 0 ? kDefaultTail : (c->tuning.tail == OXPT_TAIL_NONE ? 0u : c->tuning.tail);
    // The jobs of round `done` (items with samples left), largest first,
    // split into pixel runs and path tails; returns the round's sample slots.
    std::vector<oxd::Job> sorted;
    std::vector<double> work;
    auto build_round = [&](uint32_t done) -> uint64_t {
        std::vector<oxd::Job> jobs;
        work.clear();
        for (size_t i = 0; i < n; i++) {
            const oxpt_batch_item &it = items[i];
            if (it.count <= done) continue;
            jobs.push_back(make_job(c, it, it.first + done, std::min(spd, it.count - done)));
            work.push_back((double)it.accum->w * it.accum->h * jobs.back().count);
        }
        // Largest job first: the small ones fill the end of the launch.
        std::vector<size_t> order(jobs.size());
        for (size_t k = 0; k < order.size(); k++) order[k] = k;
        std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) { return work[a] > work[b]; });
        sorted.resize(jobs.size());
        for (size_t k = 0; k < order.size(); k++) sorted[k] = jobs[order[k]];
        return split_round(sorted.data(), sorted.size(), tail16, resident_lanes(c));
    };
    // One sample slot per tail path, sized for the largest round; fewer
    // samples per round while the slots would pass the cap.
    uint64_t slots = 0;
    for (;;) {
        slots = 0;
        for (uint32_t done = 0; done < max_count; done += spd) slots = std::max(slots, build_round(done));
        if (slots <= kBatchSlotCap || spd == 1) break;
        spd = (spd + 1) / 2;
    }
    if (slots >= (1ull << 32)) return fail(c, OXPT_E_ARG, "batch: round too large (lower max_paths)");
    if (int rc = ensure_samples(c, std::max<uint64_t>(slots, 1))) return rc;
    for (uint32_t done = 0; done < max_count; done += spd) {
        build_round(done);
        uint32_t nj = (uint32_t)sorted.size();
        // Default form: 4 waves per SIMD, but 3 for a small round (a few
        // paths per resident lane), where the end of the round — one path's
        // latency — weighs more than throughput.
        double round_paths = 0.0;
        for (double wk : work) round_paths += wk;
        uint32_t max_punct = 0;
        for (const oxd::Job &J : sorted) max_punct = std::max(max_punct, J.P.cfg[2]);
        const uint32_t mode = trace_mode(c->tuning, max_punct);
        bool opaque = mode != 0;
        for (size_t i = 0; i < n && opaque; i++)
            if (items[i].count > done) opaque = opaque_walk(c->tuning, items[i].scene->host_tables, items[i].params);
        oxpt_tuning form = c->tuning;
        if (form.kernel == 0 && round_paths < kSmallRound) form.kernel = mode == 1 && round_paths >= kTinyRound ? 4 : 2;
        else if (form.kernel == 0 && mode == 1) form.kernel = kEarlyDefaultForm;
        else if (form.kernel == 0 && mode == 2) form.kernel = kQueueDefaultForm;
        if (mode == 2)
            if (int rc = ensure_queue(c, (uint64_t)grid * block, max_punct)) return rc;
        const uint32_t lds = stack_lds_entries(form, block, true);
        hipError_t e = hipMemcpyAsync(c->d_jobs, sorted.data(), sorted.size() * sizeof(oxd::Job), hipMemcpyHostToDevice,
                                      c->stream);
        if (e == hipSuccess)
            e = hipMemcpyAsync(c->d_counters + 2 * c->job_cap, &nj, 4, hipMemcpyHostToDevice, c->stream);
        if (e != hipSuccess) return hip_fail(c, "batch job upload", e);
        oxd::BatchArgs args{c->d_jobs, c->d_counters + 2 * c->job_cap, c->d_counters, (uint32_t)c->job_cap, c->d_samples,
                            c->d_nee_q};
        c->last_kernel = mode == 2 ? (opaque ? "trace_batch(early, queue, opaque)" : "trace_batch(early, queue)")
                         : mode == 1 ? (opaque ? "trace_batch(early, opaque)" : "trace_batch(early)")
                                     : "trace_batch";
        EventPair ep{};
        bool timed = hipEventCreate(&ep.a) == hipSuccess && hipEventCreate(&ep.b) == hipSuccess;
        if (timed) (void)hipEventRecord(ep.a, c->stream);
        if (c->use_graph) {
            if (c->graph && (c->graph_block != block || c->graph_grid != grid || c->graph_kernel != form.kernel ||
                             c->graph_lds != lds || c->graph_mode != mode || c->graph_opaque != opaque)) {
                (void)hipGraphExecDestroy(c->graph);
                c->graph = nullptr;
            }
            if (!c->graph) {
                // Capture the step once: counter reset + trace, both reading
                // the job count from device memory.
                hipGraph_t g = nullptr;
                e = hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal);
                if (e == hipSuccess) {
                    hipError_t le = oxd::launch_trace_batch(args, form.kernel, block, grid, lds, mode, opaque, c->stream);
                    e = hipStreamEndCapture(c->stream, &g);
                    if (e == hipSuccess) e = le;
                }
                if (e == hipSuccess) e = hipGraphInstantiate(&c->graph, g, nullptr, nullptr, 0);
                if (g) (void)hipGraphDestroy(g);
                if (e != hipSuccess) return hip_fail(c, "batch graph capture", e);
                c->graph_block = block;
                c->graph_grid = grid;
                c->graph_kernel = form.kernel;
                c->graph_lds = lds;
                c->graph_mode = mode;
                c->graph_opaque = opaque;
            }
            e = hipGraphLaunch(c->graph, c->stream);
            c->graph_launches++;
        } else {
            e = oxd::launch_trace_batch(args, form.kernel, block, grid, lds, mode, opaque, c->stream);
        }
        if (e != hipSuccess) return hip_fail(c, "batch launch", e);
        if (timed) {
            (void)hipEventRecord(ep.b, c->stream);
            c->events.push_back(ep);
            c->launches++;
        }
        if (c->events.size() > 4096) fold_events(c, 2048);
        // The host staging of the next round's table must not race this copy.
        e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) return hip_fail(c, "batch", e);
    }
    return OXPT_OK;
}

int oxpt_refine_batch_noalloc(oxpt_ctx *c, const oxpt_batch_item *items, size_t n) {
    if (!c || (!items && n)) return OXPT_E_ARG;
    if (n == 0) return OXPT_OK;
    if (n > kNaJobs) return fail(c, OXPT_E_ARG, "batch (no allocation): more than 256 items");
    const oxpt_batch_item &i0 = items[0];
    uint64_t pix64 = 0;
    uint32_t max_count = 0;
    for (size_t i = 0; i < n; i++) {
        const oxpt_batch_item &it = items[i];
        if (!it.scene || !it.accum || !same_domain(it.scene->ctx, c) || !same_domain(it.accum->ctx, c))
            return fail(c, OXPT_E_ARG, "batch: item belongs to another context");
        if (it.params.dims[0] != it.accum->w || it.params.dims[1] != it.accum->h)
            return fail(c, OXPT_E_ARG, "batch: params size differs from the accumulator");
        if (int rc = check_params(c, it.scene, &it.params)) return rc;
        for (size_t k = 0; k < i; k++)
            if (items[k].accum == it.accum) return fail(c, OXPT_E_ARG, "batch: two items share an accumulator");
        const uint32_t *a = it.scene->host_tables.data(), *b = i0.scene->host_tables.data();
        if (std::memcmp(a + it.params.offs0[0], b + i0.params.offs0[0], OXPT_SOBOL_WORDS * 4) != 0 ||
            std::memcmp(a + it.params.offs0[1], b + i0.params.offs0[1], OXPT_SHEEN_WORDS * 4) != 0)
            return fail(c, OXPT_E_ARG, "batch: jobs carry different Sobol' / sheen tables");
        if (it.count == 0) continue;
        pix64 += 64ull * ((it.accum->w + 7u) / 8u) * ((it.accum->h + 7u) / 8u);
        max_count = std::max(max_count, it.count);
    }
    if (max_count == 0) return OXPT_OK;
    if (pix64 > kNaSlots)
        return fail(c, OXPT_E_ARG, "batch (no allocation): one sample of every item is more than the 4194304 paths "
                                   "(8x8-tile padded pixels) its scratch holds");
    if (c->tuning.kernel == 7) return fail(c, OXPT_E_ARG, "batch (no allocation): the wavefront engine allocates");
    if (setdev(c)) return OXPT_E_BACKEND;
    uint8_t *base = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_scratch_mu);
        if (c->device < kNaMaxDevices) base = static_cast<uint8_t *>(g_scratch_ptr[c->device]);
    }
    if (!base) return fail(c, OXPT_E_BACKEND, "batch (no allocation): no scratch address for this device");
    uint32_t *counters = reinterpret_cast<uint32_t *>(base);
    oxd::Job *djobs = reinterpret_cast<oxd::Job *>(base + kNaCounterBytes);
    float4 *samples = reinterpret_cast<float4 *>(base + kNaCounterBytes + kNaJobBytes);
    // Rounds of spd samples per item, each round's paths within the scratch
    // (and within max_paths).
    const uint64_t cap = std::min<uint64_t>(kNaSlots, std::max<uint64_t>(c->tuning.max_paths, 1));
    const uint32_t spd = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(max_count, cap / pix64));
    const uint32_t block = c->tuning.block;
    const uint32_t grid = std::max(1u, c->n_cu * c->tuning.waves_per_cu / (block / 64u));
    const uint32_t tail16 = c->tuning.tail == 0 ? kDefaultTail : (c->tuning.tail == OXPT_TAIL_NONE ? 0u : c->tuning.tail);
    // Host staging on the stack (about 67 KiB): no heap allocation either.
    oxd::Job jobs[kNaJobs];
    double work[kNaJobs];
    uint32_t order[kNaJobs];
    for (uint32_t done = 0; done < max_count; done += spd) {
        // Live items of this round, largest first (a stable insertion sort:
        // std::stable_sort may allocate).
        uint32_t nj = 0;
        double round_paths = 0.0;
        for (size_t i = 0; i < n; i++) {
            const oxpt_batch_item &it = items[i];
            if (it.count <= done) continue;
            const double wk = (double)it.accum->w * it.accum->h * std::min(spd, it.count - done);
            uint32_t k = nj++;
            while (k > 0 && work[k - 1] < wk) {
                work[k] = work[k - 1];
                order[k] = order[k - 1];
                k--;
            }
            work[k] = wk;
            order[k] = (uint32_t)i;
            round_paths += wk;
        }
        uint32_t max_punct = 0;
        for (uint32_t k = 0; k < nj; k++) {
            const oxpt_batch_item &it = items[order[k]];
            jobs[k] = make_job(c, it, it.first + done, std::min(spd, it.count - done));
            max_punct = std::max(max_punct, it.params.cfg[2]);
        }
        // Tail paths are at most every path of the round: within the scratch.
        split_round(jobs, nj, tail16, resident_lanes(c));
        // The NEE queue of the early form with punctual lights would need
        // memory: those rounds run the general form (the same images).
        uint32_t mode = trace_mode(c->tuning, max_punct);
        if (mode == 2) mode = 0;
        bool opaque = mode != 0;
        for (size_t i = 0; i < n && opaque; i++)
            if (items[i].count > done) opaque = opaque_walk(c->tuning, items[i].scene->host_tables, items[i].params);
        oxpt_tuning form = c->tuning;
        if (form.kernel == 0 && round_paths < kSmallRound) form.kernel = mode == 1 && round_paths >= kTinyRound ? 4 : 2;
        else if (form.kernel == 0 && mode == 1) form.kernel = kEarlyDefaultForm;
        const uint32_t lds = stack_lds_entries(form, block, true);
        hipError_t e = hipMemcpyAsync(djobs, jobs, nj * sizeof(oxd::Job), hipMemcpyHostToDevice, c->stream);
        if (e == hipSuccess)
            e = hipMemcpyAsync(counters + 2 * kNaJobs, &nj, 4, hipMemcpyHostToDevice, c->stream);
        if (e != hipSuccess) return hip_fail(c, "batch (no allocation) job upload", e);
        oxd::BatchArgs args{djobs, counters + 2 * kNaJobs, counters, kNaJobs, samples, nullptr};
        c->last_kernel = mode == 1 ? (opaque ? "trace_batch(early, opaque)" : "trace_batch(early)") : "trace_batch";
        e = oxd::launch_trace_batch(args, form.kernel, block, grid, lds, mode, opaque, c->stream);
        if (e != hipSuccess) return hip_fail(c, "batch (no allocation) launch", e);
        // The stack staging of the next round must not race this copy.
        e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) return hip_fail(c, "batch (no allocation)", e);
    }
    return OXPT_OK;
}

int oxpt_finish(oxpt_ctx *c) {
    if (!c) return OXPT_E_ARG;
    if (setdev(c)) return OXPT_E_BACKEND;
    hipError_t e = hipStreamSynchronize(c->stream);
    return e == hipSuccess ? OXPT_OK : hip_fail(c, "finish", e);
}

int oxpt_autotune(oxpt_ctx *c, const oxpt_scene *s, const oxpt_params *P, uint32_t spp_per_trial,
                  oxpt_tuning *best_out, double *best_ms) {
    if (!c || !s || !P || !same_domain(s->ctx, c)) return OXPT_E_ARG;
    if (int rc = check_params(c, s, P)) return rc;
    if (setdev(c)) return OXPT_E_BACKEND;
    const uint32_t spp = std::max(1u, std::min(spp_per_trial, 64u));
    // Candidates: megakernel occupancy forms x traversal-stack LDS share x
    // path-granular tail share, and the wavefront engine's default form.
    // Every one of them gives the same images, so the choice is purely a
    // matter of speed.
    struct Cand {
        uint32_t kernel, stack_lds, wf_variant, tail;
    };
    const Cand cands[] = {{0, 0, 0, 0}, {0, 0, 0, 16}, {0, 0, 0, 4}, {2, 0, 0, 0}, {3, 0, 0, 0},
                          {4, 0, 0, 0}, {5, 0, 0, 0}, {0, 8, 0, 0}, {0, OXPT_STACK_PRIVATE, 0, 0},
                          {7, 0, 0, 0}};
    oxpt_accum *tmp = nullptr;
    if (int rc = oxpt_accum_create(c, P->dims[0], P->dims[1], &tmp)) return rc;
    const oxpt_tuning saved = c->tuning;
    oxpt_tuning best = saved;
    double best_t = 1e300;
    int rc = OXPT_OK;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) rc = fail(c, OXPT_E_BACKEND, "hip: events");
    for (const Cand &cd : cands) {
        if (rc) break;
        oxpt_tuning t = saved;
        t.kernel = cd.kernel;
        t.stack_lds = cd.stack_lds;
        t.wf_variant = cd.wf_variant;
        t.tail = cd.tail;
        if ((rc = oxpt_ctx_set_tuning(c, &t))) break;
        // One warm trial (first-launch costs, caches), then the fastest of
        // two timed ones.
        if ((rc = oxpt_refine(c, s, tmp, P, 0, spp))) break;
        float cand_ms = 1e30f;
        hipError_t e = hipSuccess;
        for (uint32_t rep = 0; rep < 2 && e == hipSuccess && rc == OXPT_OK; rep++) {
            e = hipEventRecord(e0, c->stream);
            if (e == hipSuccess && (rc = oxpt_refine(c, s, tmp, P, spp * (1 + rep), spp))) break;
            if (e == hipSuccess) e = hipEventRecord(e1, c->stream);
            if (e == hipSuccess) e = hipEventSynchronize(e1);
            float ms = 0.0f;
            if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
            if (e == hipSuccess) cand_ms = std::min(cand_ms, ms);
        }
        if (rc) break;
        if (e != hipSuccess) {
            rc = hip_fail(c, "autotune", e);
            break;
        }
        if (cand_ms < best_t) {
            best_t = cand_ms;
            best = c->tuning;
        }
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    oxpt_accum_destroy(tmp);
    c->tuning = rc ? saved : best;
    if (rc) return rc;
    if (best_out) *best_out = best;
    if (best_ms) *best_ms = best_t;
    return OXPT_OK;
}

int oxpt_autotune_batch(oxpt_ctx *c, const oxpt_batch_item *items, size_t n, uint32_t spp_per_trial,
                        oxpt_tuning *best_out, double *best_ms) {
    if (!c || !items || n == 0) return OXPT_E_ARG;
    if (setdev(c)) return OXPT_E_BACKEND;
    // Trials run the caller's items with at most spp_per_trial samples each
    // into scratch accumulators (the caller's stay untouched). Candidates:
    // megakernel occupancy forms x samples per dispatch (the path cap that
    // splits the batch into rounds) x traversal-stack LDS share. Every
    // candidate gives the same images.
    const uint32_t spp = std::max(1u, std::min(spp_per_trial, 256u));
    std::vector<oxpt_accum *> tmp(n, nullptr);
    std::vector<oxpt_batch_item> trial(items, items + n);
    int rc = OXPT_OK;
    for (size_t i = 0; i < n && rc == OXPT_OK; i++) {
        if (!items[i].accum || !same_domain(items[i].accum->ctx, c)) rc = fail(c, OXPT_E_ARG, "autotune: bad item");
        else rc = oxpt_accum_create(c, items[i].accum->w, items[i].accum->h, &tmp[i]);
        if (rc == OXPT_OK) {
            trial[i].accum = tmp[i];
            trial[i].first = 0;
            trial[i].count = std::min(items[i].count, spp);
        }
    }
    struct Cand {
        uint32_t kernel, max_paths, stack_lds;
    };
    std::vector<Cand> cands;
    for (uint32_t k : {0u, 2u, 3u, 4u, 5u})
        for (uint32_t mp : {kDefaultMaxPaths, 1u << 28, 1u << 26}) cands.push_back({k, mp, 0u});
    cands.push_back({0u, kDefaultMaxPaths, 8u});
    cands.push_back({0u, kDefaultMaxPaths, OXPT_STACK_PRIVATE});
    const oxpt_tuning saved = c->tuning;
    oxpt_tuning best = saved;
    double best_t = 1e300;
    for (const Cand &cd : cands) {
        if (rc) break;
        oxpt_tuning t = saved;
        t.kernel = cd.kernel;
        t.max_paths = cd.max_paths;
        t.stack_lds = cd.stack_lds;
        if ((rc = oxpt_ctx_set_tuning(c, &t))) break;
        // One warm trial, then the faster of two timed ones (wall time of
        // the whole call: rounds, uploads and launches).
        if ((rc = oxpt_refine_batch(c, trial.data(), n))) break;
        double cand = 1e300;
        for (int rep = 0; rep < 2 && rc == OXPT_OK; rep++) {
            auto t0 = std::chrono::steady_clock::now();
            rc = oxpt_refine_batch(c, trial.data(), n);
            if (rc == OXPT_OK) rc = oxpt_finish(c);
            cand = std::min(cand, std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count());
        }
        if (rc) break;
        if (cand < best_t) {
            best_t = cand;
            best = c->tuning;
        }
    }
    for (oxpt_accum *a : tmp) oxpt_accum_destroy(a);
    c->tuning = rc ? saved : best;
    if (rc) return rc;
    if (best_out) *best_out = best;
    if (best_ms) *best_ms = best_t;
    return OXPT_OK;
}

int oxpt_ctx_timing(oxpt_ctx *c, int reset, uint64_t *launches, double *total_ms) {
    if (!c) return OXPT_E_ARG;
    if (setdev(c)) return OXPT_E_BACKEND;
    fold_events(c, 0);
    if (launches) *launches = c->launches;
    if (total_ms) *total_ms = c->total_ms;
    if (reset) {
        c->launches = 0;
        c->total_ms = 0.0;
    }
    return OXPT_OK;
}

int oxpt_resolve(oxpt_ctx *c, const oxpt_accum *a, const oxpt_resolve_params *rp, void *host, size_t bytes) {
    if (!c || !a || !rp || !host || !same_domain(a->ctx, c)) return OXPT_E_ARG;
    size_t need = (size_t)a->w * a->h * (rp->cfg[3] == 1 ? 16 : 4);
    if (bytes < need) return fail(c, OXPT_E_ARG, "resolve: output buffer too small");
    if (setdev(c)) return OXPT_E_BACKEND;
    if (c->d_out_bytes < need) {
        dfree(c, c->d_out, c->d_out_bytes);
        c->d_out = nullptr;
        c->d_out_bytes = 0;
        if (int rc = dalloc(c, need, &c->d_out, "resolve target")) return rc;
        c->d_out_bytes = need;
    }
    hipError_t e = oxd::launch_resolve(a->d, a->w * a->h, *rp, c->d_out, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(host, c->d_out, need, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    return e == hipSuccess ? OXPT_OK : hip_fail(c, "resolve", e);
}

int oxpt_resolve_device(oxpt_ctx *c, const oxpt_accum *a, const oxpt_resolve_params *rp, void *dev, size_t bytes) {
    if (!c || !a || !rp || !dev || !same_domain(a->ctx, c)) return OXPT_E_ARG;
    const size_t need = (size_t)a->w * a->h * (rp->cfg[3] == 1 ? 16 : 4);
    if (bytes < need) return fail(c, OXPT_E_ARG, "resolve: output buffer too small");
    if (setdev(c)) return OXPT_E_BACKEND;
    // Only device memory of this context's GPU (or managed memory) may be
    // written by the kernel; anything else is refused here.
    hipPointerAttribute_t at;
    std::memset(&at, 0, sizeof at);
    hipError_t e = hipPointerGetAttributes(&at, dev);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return fail(c, OXPT_E_ARG, "resolve: output is not device memory");
    }
    const bool device_mem = at.type == hipMemoryTypeDevice || at.type == hipMemoryTypeManaged;
    if (!device_mem || (at.type == hipMemoryTypeDevice && at.device != c->device))
        return fail(c, OXPT_E_ARG, "resolve: output is not device memory of this context's GPU");
    e = oxd::launch_resolve(a->d, a->w * a->h, *rp, dev, c->stream);
    return e == hipSuccess ? OXPT_OK : hip_fail(c, "resolve", e);
}

}  // extern "C"

// Stream-ordered helpers of the serving loop (declared in scene.hpp).
namespace oxpt {

int ctx_mark(oxpt_ctx *c, void **mark) {
    *mark = nullptr;
    if (setdev(c)) return OXPT_E_BACKEND;
    hipEvent_t ev = nullptr;
    hipError_t e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventRecord(ev, c->stream);
    if (e != hipSuccess) {
        if (ev) (void)hipEventDestroy(ev);
        return hip_fail(c, "mark", e);
    }
    *mark = ev;
    return OXPT_OK;
}

int mark_wait(oxpt_ctx *c, void *mark) {
    if (!mark) return OXPT_OK;
    if (setdev(c)) return OXPT_E_BACKEND;
    hipEvent_t ev = static_cast<hipEvent_t>(mark);
    hipError_t e = hipEventSynchronize(ev);
    (void)hipEventDestroy(ev);
    return e == hipSuccess ? OXPT_OK : hip_fail(c, "mark wait", e);
}

int device_alloc(oxpt_ctx *c, size_t bytes, void **p) {
    if (setdev(c)) return OXPT_E_BACKEND;
    return dalloc(c, bytes, p, "staging buffer");
}

void device_free(oxpt_ctx *c, void *p, size_t bytes) {
    if (!p) return;
    setdev(c);
    dfree(c, p, bytes);
}

int copy_to_host(oxpt_ctx *c, void *host, const void *dev, size_t bytes) {
    if (setdev(c)) return OXPT_E_BACKEND;
    hipError_t e = hipMemcpyAsync(host, dev, bytes, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    return e == hipSuccess ? OXPT_OK : hip_fail(c, "readback", e);
}

}  // namespace oxpt

extern "C" {

int oxpt_image_create(oxpt_ctx *c, uint32_t w, uint32_t h, int hdr, oxpt_image **out) {
    if (!c || !out) return OXPT_E_ARG;
    *out = nullptr;
    if (w == 0 || h == 0) return fail(c, OXPT_E_INVALID, "image: zero dimension");
    if (setdev(c)) return OXPT_E_BACKEND;
    auto *im = new oxpt_image();
    im->ctx = c->base;
    im->w = w;
    im->h = h;
    im->hdr = hdr ? 1 : 0;
    im->bytes = (size_t)w * h * (hdr ? 16u : 4u);
    if (int rc = dalloc(c, im->bytes, &im->d, "display image")) {
        delete im;
        return rc;
    }
    *out = im;
    return OXPT_OK;
}

void oxpt_image_destroy(oxpt_image *im) {
    if (!im) return;
    setdev(im->ctx);
    (void)hipStreamSynchronize(im->ctx->stream);
    dfree(im->ctx, im->d, im->bytes);
    delete im;
}

void *oxpt_image_device_ptr(const oxpt_image *im) { return im ? im->d : nullptr; }
size_t oxpt_image_bytes(const oxpt_image *im) { return im ? im->bytes : 0; }

int oxpt_image_resolve(oxpt_ctx *c, const oxpt_accum *a, const oxpt_resolve_params *rp, oxpt_image *im) {
    if (!c || !a || !rp || !im || !same_domain(a->ctx, c) || !same_domain(im->ctx, c)) return OXPT_E_ARG;
    if (im->w != a->w || im->h != a->h || im->hdr != (rp->cfg[3] == 1 ? 1 : 0))
        return fail(c, OXPT_E_ARG, "image: size or format differs from the resolve");
    return oxpt_resolve_device(c, a, rp, im->d, im->bytes);
}

int oxpt_image_export(oxpt_ctx *c, const oxpt_image *im, void *handle_out, size_t handle_bytes) {
    if (!c || !im || !handle_out || !same_domain(im->ctx, c)) return OXPT_E_ARG;
    if (handle_bytes < sizeof(hipIpcMemHandle_t)) return fail(c, OXPT_E_ARG, "image export: handle buffer too small");
    // Shared allocations are kept under 2 GiB (a larger one is refused).
    if (im->bytes >= (2ull << 30)) return fail(c, OXPT_E_ARG, "image export: images of 2 GiB or more are not shared");
    if (setdev(c)) return OXPT_E_BACKEND;
    hipIpcMemHandle_t h;
    hipError_t e = hipIpcGetMemHandle(&h, im->d);
    if (e != hipSuccess) return hip_fail(c, "image export", e);
    std::memcpy(handle_out, &h, sizeof h);
    return OXPT_OK;
}

int oxpt_debug_pattern_words(oxpt_ctx *c, const oxpt_scene *s, uint32_t seed, size_t n, const uint32_t *x,
                             const uint32_t *y, const uint32_t *index, const uint32_t *p, uint32_t *out4) {
    if (!c || !s || !x || !y || !index || !p || !out4 || !same_domain(s->ctx, c)) return OXPT_E_ARG;
    if (n == 0) return OXPT_OK;
    if (n > (1u << 28)) return fail(c, OXPT_E_ARG, "debug: too many tuples");
    if (setdev(c)) return OXPT_E_BACKEND;
    void *d = nullptr;
    size_t in_b = n * 4, bytes = 4 * in_b + n * 16;
    if (int rc = dalloc(c, bytes, &d, "debug buffers")) return rc;
    uint32_t *dx = (uint32_t *)d, *dy = dx + n, *di = dy + n, *dp = di + n, *dout = dp + n;
    hipError_t e = hipMemcpyAsync(dx, x, in_b, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(dy, y, in_b, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(di, index, in_b, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(dp, p, in_b, hipMemcpyHostToDevice, c->stream);
    // The Sobol' table sits at the start of the tables (offset 0) unless the
    // caller packed it elsewhere; the tracer always packs it first.
    if (e == hipSuccess) e = oxd::launch_debug_pattern(s->tables, 0, seed, (uint32_t)n, dx, dy, di, dp, dout, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(out4, dout, n * 16, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    dfree(c, d, bytes);
    return e == hipSuccess ? OXPT_OK : hip_fail(c, "debug pattern words", e);
}

int oxpt_debug_light_picks(oxpt_ctx *c, const oxpt_scene *s, const oxpt_params *P, size_t n, const float *u,
                           uint32_t *out3) {
    if (!c || !s || !P || !u || !out3 || !same_domain(s->ctx, c)) return OXPT_E_ARG;
    if (n == 0) return OXPT_OK;
    if (n > (1u << 28)) return fail(c, OXPT_E_ARG, "debug: too many uniforms");
    if (int rc = check_params(c, s, P)) return rc;
    if (setdev(c)) return OXPT_E_BACKEND;
    void *d = nullptr;
    size_t bytes = n * 4 + n * 12;
    if (int rc = dalloc(c, bytes, &d, "debug buffers")) return rc;
    float *du = (float *)d;
    uint32_t *dout = (uint32_t *)(du + n);
    oxd::TraceArgs args = trace_args(c, s, nullptr, P);
    hipError_t e = hipMemcpyAsync(du, u, n * 4, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = oxd::launch_debug_picks(args, (uint32_t)n, du, dout, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(out3, dout, n * 12, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    dfree(c, d, bytes);
    return e == hipSuccess ? OXPT_OK : hip_fail(c, "debug light picks", e);
}

void oxpt_udiv_magic(uint32_t d, uint32_t *m, uint32_t *l) {
    const oxd::UDiv u = oxd::udiv_magic(d);
    if (m) *m = u.m;
    if (l) *l = u.l;
}

int oxpt_debug_udiv(oxpt_ctx *c, size_t n, const uint32_t *num, const uint32_t *den, uint32_t *out) {
    if (!c || !num || !den || !out) return OXPT_E_ARG;
    if (n == 0) return OXPT_OK;
    if (n > (1u << 26)) return fail(c, OXPT_E_ARG, "debug: too many divisions");
    for (size_t i = 0; i < n; i++)
        if (den[i] == 0) return fail(c, OXPT_E_ARG, "debug: division by zero");
    if (setdev(c)) return OXPT_E_BACKEND;
    std::vector<uint32_t> m(n), l(n);
    for (size_t i = 0; i < n; i++) oxpt_udiv_magic(den[i], &m[i], &l[i]);
    void *d = nullptr;
    size_t bytes = n * 16;
    if (int rc = dalloc(c, bytes, &d, "debug buffers")) return rc;
    uint32_t *dn = (uint32_t *)d, *dm = dn + n, *dl = dm + n, *dout = dl + n;
    hipError_t e = hipMemcpyAsync(dn, num, n * 4, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(dm, m.data(), n * 4, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(dl, l.data(), n * 4, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = oxd::launch_debug_udiv((uint32_t)n, dn, dm, dl, dout, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(out, dout, n * 4, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    dfree(c, d, bytes);
    return e == hipSuccess ? OXPT_OK : hip_fail(c, "debug udiv", e);
}

int oxpt_debug_texel_index(oxpt_ctx *c, size_t n, const float *u, const int32_t *size, const uint32_t *mode,
                           int32_t *out3) {
    if (!c || !u || !size || !mode || !out3) return OXPT_E_ARG;
    if (n == 0) return OXPT_OK;
    if (n > (1u << 26)) return fail(c, OXPT_E_ARG, "debug: too many coordinates");
    // Level sizes are 16-bit fields of the descriptors; modes 0..2.
    for (size_t i = 0; i < n; i++)
        if (size[i] < 1 || size[i] > 65535 || mode[i] > 2) return fail(c, OXPT_E_ARG, "debug: level size or wrap mode out of range");
    if (setdev(c)) return OXPT_E_BACKEND;
    void *d = nullptr;
    size_t bytes = n * 24;
    if (int rc = dalloc(c, bytes, &d, "debug buffers")) return rc;
    float *du = (float *)d;
    int32_t *ds = (int32_t *)(du + n);
    uint32_t *dm = (uint32_t *)(ds + n);
    int32_t *dout = (int32_t *)(dm + n);
    hipError_t e = hipMemcpyAsync(du, u, n * 4, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(ds, size, n * 4, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(dm, mode, n * 4, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = oxd::launch_debug_texel_index((uint32_t)n, du, ds, dm, dout, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(out3, dout, n * 12, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    dfree(c, d, bytes);
    return e == hipSuccess ? OXPT_OK : hip_fail(c, "debug texel index", e);
}

int oxpt_debug_closest_hit(oxpt_ctx *c, const oxpt_scene *s, const oxpt_params *P, size_t n, const float *o3,
                           const float *d3, uint32_t *out2) {
    if (!c || !s || !P || !o3 || !d3 || !out2 || !same_domain(s->ctx, c)) return OXPT_E_ARG;
    if (n == 0) return OXPT_OK;
    if (n > (1u << 26)) return fail(c, OXPT_E_ARG, "debug: too many rays");
    if (setdev(c)) return OXPT_E_BACKEND;
    void *d = nullptr;
    size_t bytes = n * 12 * 2 + n * 8;
    if (int rc = dalloc(c, bytes, &d, "debug buffers")) return rc;
    float *dorig = (float *)d, *ddir = dorig + 3 * n;
    uint32_t *dout = (uint32_t *)(ddir + 3 * n);
    oxd::TraceArgs args = trace_args(c, s, nullptr, P);
    hipError_t e = hipMemcpyAsync(dorig, o3, n * 12, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(ddir, d3, n * 12, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = oxd::launch_debug_hits(args, (uint32_t)n, dorig, ddir, dout, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(out2, dout, n * 8, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    dfree(c, d, bytes);
    return e == hipSuccess ? OXPT_OK : hip_fail(c, "debug closest hit", e);
}

}  // extern "C"
