// Wavefront path tracing: per-slot path state in HBM between kernels.
#pragma once

#include "pt_kernels.hpp"

namespace oxd {

constexpr uint32_t WF_LIVE_RING = 64;

// One in-flight path (112 bytes, read and written as float4s).
struct alignas(16) PathSlot {
    uint32_t item, job, pixel_seed, path_seed;
    uint32_t index, k, flags, n_shadow;
    float prev_pdf, beta_x, beta_y, beta_z;
    float rad_x, rad_y, rad_z, prev_x;
    float prev_y, prev_z, sigma_x, sigma_y;
    float sigma_z, d_x, d_y, d_z;
    float o_x, o_y, o_z, pad;
};
static_assert(sizeof(PathSlot) == 112, "PathSlot is 7 float4");

// Closest hit of a path ray: leaf slot (bit 31 = front face, NONE = miss),
// distance and the barycentric (u, v).
struct alignas(16) HitRec {
    uint32_t slot;
    float t, b1, b2;
};

// One NEE shadow ray and the contribution it adds when unoccluded.
struct alignas(16) ShadowRec {
    float4 o_tmax;       // origin, t_max
    float4 d_skip;       // direction, skipped global triangle id (bits)
    float4 pending_rid;  // contribution, ray id (bits)
};

struct WfArgs {
    const Job *jobs;
    const uint32_t *job_base;  // first item of each job (prefix sums)
    uint32_t n_jobs;
    PathSlot *slots;
    HitRec *hits;
    ShadowRec *shadows;   // [max_shadow][n_slots]
    uint32_t *occluded;   // [max_shadow][n_slots]
    float4 *samples;      // per item: rgb + covered
    uint32_t *item_counter;
    uint32_t *live;       // WF_LIVE_RING counters of live slots per iteration
    uint32_t n_slots, max_shadow, n_items, iter;
};

hipError_t wf_launch_clear(const WfArgs &A, hipStream_t s);
// variant: see oxpt_tuning.wf_variant.
hipError_t wf_launch_iteration(const WfArgs &A, uint32_t variant, hipStream_t s);
hipError_t wf_launch_reduce(const WfArgs &A, uint32_t job, uint32_t n_items, hipStream_t s);

}  // namespace oxd
