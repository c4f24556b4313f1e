// Kernel argument block and the launchers the runtime calls.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "oxpt/layout.h"

namespace oxd {

struct TraceArgs {
    oxpt_params P;
    const float4 *geom;
    const float4 *attrs;
    const uint32_t *tables;
    const uint16_t *atlas;
    uint32_t atlas_side;
    float4 *accum;
    uint32_t first, count;  // sample run of this launch
    uint32_t n_items;       // 64 * number of 8x8 tiles
    uint32_t tiles_x;
    uint32_t *work;         // work-queue counters (persistent kernel): [0] pixels, [1] tail paths
    // Persistent kernel tail: items [split, n_items) (whole tiles) are traced
    // one path per claim, each path's (rgb, covered) into samples[tile' * 64
    // * count + sample * 64 + in] (tile' counted from split), and tail_reduce
    // adds them to the accumulator in sample order. split == n_items: no tail.
    uint32_t split;
    uint32_t tail_work;     // (n_items - split) * count
    float4 *samples;
    // Early form with punctual lights: each lane's NEE queue, word-major
    // across the launch's lanes (word w of lane g at nee_q[w * lanes + g]).
    float *nee_q;
};

// One job of a batched launch: a resident scene, its Params, an accumulator
// and a sample run. 256 bytes, read with scalar loads by the wave that owns it.
// Items [0, split) (whole 8x8 tiles) are traced as whole-pixel runs added
// straight into the accumulator; items [split, n_items) are traced one path
// per claim, tail_work = (n_items - split) * count paths, each path's (rgb,
// covered) at sample_base + tile' * 64 * count + sample * 64 + in of the
// launch's sample buffer (tile' counted from split), added in sample order
// by batch_reduce.
struct alignas(16) Job {
    oxpt_params P;  // 176 B
    const float4 *geom;
    const float4 *attrs;
    const uint32_t *tables;
    const uint16_t *atlas;
    float4 *accum;
    uint32_t atlas_side, first, count, n_items;
    uint32_t split, tail_work, sample_base;
    // The claim step's divisors as multiply-shift pairs (udiv_magic): the
    // paths per tile (64 * count: tile_m, shift in bits 0-7 of `shifts`)
    // and the tiles per row (tx_m, shift in bits 8-15).
    uint32_t tile_m, tx_m, shifts;
    __host__ __device__ uint32_t tiles_x() const { return (P.dims[0] + 7u) >> 3; }
    __host__ __device__ uint32_t tile_l() const { return shifts & 0xffu; }
    __host__ __device__ uint32_t tx_l() const { return (shifts >> 8) & 0xffu; }
};
static_assert(sizeof(Job) == 256, "Job is 256 bytes");

// Unsigned division by a divisor known ahead, exact for every 32-bit n
// (the round-up multiply-shift method with an add step, Granlund and
// Montgomery 1994, as in Hacker's Delight 10-8): l = ceil(log2 d),
// m = floor(2^32 (2^l - d) / d) + 1; q = (t + ((n - t) >> 1)) >> (l - 1)
// with t = the high word of n * m. d = 1 is l = 0 (q = n).
struct UDiv {
    uint32_t m, l;
};
inline UDiv udiv_magic(uint32_t d) {
    uint32_t l = 0;
    while (l < 32 && (1ull << l) < d) l++;
    if (l == 0) return {0u, 0u};
    return {(uint32_t)(((((1ull << l) - d) << 32) / d) + 1ull), l};
}
__device__ __forceinline__ uint32_t udiv(uint32_t n, uint32_t m, uint32_t l) {
    if (l == 0) return n;
    const uint32_t t = __umulhi(n, m);
    return (t + ((n - t) >> 1)) >> (l - 1u);
}

// The job count lives in device memory (written with the job table), so
// the launch arguments never change and the step can be captured once in a
// HIP graph and replayed.
struct BatchArgs {
    const Job *jobs;
    const uint32_t *n_jobs;
    uint32_t *counters;  // two work counters per job: [2j] pixels, [2j + 1] tail paths
    uint32_t max_jobs;   // capacity of jobs (counters: 2 * max_jobs)
    float4 *samples;     // one (rgb, covered) per tail path of the launch
    float *nee_q;        // NEE queue of the early form with punctual lights (as TraceArgs)
};

// lds_depth: traversal-stack entries per lane held in LDS (dynamic shared
// memory, block / 64 * lds_depth * 512 bytes per workgroup).
// mode: 0 the general forms; 1 the early forms (no punctual lights: NEE
// and BSDF continuation made together at a hit); 2 the early forms with a
// per-lane NEE queue (1..31 punctual lights; nee_q holds
// nee_queue_words(max punctual) words per lane). Same images in every mode.
constexpr uint32_t kQueueMaxPunctual = 31;
inline uint32_t nee_queue_words(uint32_t n_punct) { return 3u * n_punct + 12u; }
// opaque: every job's scene has only alpha-mode OPAQUE materials (the
// early modes then run the walk built without the alpha test).
hipError_t launch_trace_batch(const BatchArgs &args, uint32_t kernel, uint32_t block, uint32_t grid,
                              uint32_t lds_depth, uint32_t mode, bool opaque, hipStream_t s);
hipError_t launch_trace(const TraceArgs &args, uint32_t kernel, uint32_t block, uint32_t grid, uint32_t lds_depth,
                        uint32_t mode, bool opaque, hipStream_t s);
hipError_t launch_resolve(const float4 *accum, uint32_t n, const oxpt_resolve_params &rp, void *out, hipStream_t s);
hipError_t launch_debug_pattern(const uint32_t *tables, uint32_t off_sobol, uint32_t seed, uint32_t n,
                                const uint32_t *x, const uint32_t *y, const uint32_t *index, const uint32_t *p,
                                uint32_t *out4, hipStream_t s);
hipError_t launch_debug_picks(const TraceArgs &args, uint32_t n, const float *u, uint32_t *out3, hipStream_t s);
hipError_t launch_debug_udiv(uint32_t n, const uint32_t *num, const uint32_t *m, const uint32_t *l, uint32_t *out,
                             hipStream_t s);
hipError_t launch_debug_texel_index(uint32_t n, const float *u, const int32_t *size, const uint32_t *mode,
                                    int32_t *out3, hipStream_t s);
hipError_t launch_debug_hits(const TraceArgs &args, uint32_t n, const float *o3, const float *d3, uint32_t *out2,
                             hipStream_t s);

}  // namespace oxd
