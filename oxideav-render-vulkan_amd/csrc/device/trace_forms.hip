// Instantiations of the megakernel forms. The Makefile compiles this file
// once per (kernel, mode) — OXD_FORMS_p0 / p1 / p2 for trace_persistent in
// modes 0 / 1 / 2, OXD_FORMS_b0 / b1 / b2 for trace_batch — so the forms
// (each a whole path tracer inlined into one kernel) build in parallel.
//
// Form numbers (oxpt_tuning.kernel): 2, 4, 5 = built for 3, 5, 6 waves per
// SIMD, anything else the 4-waves build; the general mode also has 6 (8
// waves) and 8 (2 waves). The early modes (1, 2) are built twice, with the
// general walk and with the opaque walk (accept_hit): same images.
#include "pt_trace.hpp"

namespace oxd {

#if defined(OXD_FORMS_p1)
#define OXD_MODE 1
#define OXD_PERSISTENT launch_persistent_m1
#elif defined(OXD_FORMS_p2)
#define OXD_MODE 2
#define OXD_PERSISTENT launch_persistent_m2
#elif defined(OXD_FORMS_b1)
#define OXD_MODE 1
#define OXD_BATCH launch_batch_m1
#elif defined(OXD_FORMS_b2)
#define OXD_MODE 2
#define OXD_BATCH launch_batch_m2
#endif

#if defined(OXD_FORMS_p0)
void launch_persistent_m0(const TraceArgs &a, uint32_t kernel, uint32_t block, uint32_t grid, uint32_t lds_depth,
                          size_t shm, bool, hipStream_t s) {
    switch (kernel) {
        case 2: hipLaunchKernelGGL(trace_persistent<3>, dim3(grid), dim3(block), shm, s, a, lds_depth); break;
        case 4: hipLaunchKernelGGL(trace_persistent<5>, dim3(grid), dim3(block), shm, s, a, lds_depth); break;
        case 5: hipLaunchKernelGGL(trace_persistent<6>, dim3(grid), dim3(block), shm, s, a, lds_depth); break;
        case 6: hipLaunchKernelGGL(trace_persistent<8>, dim3(grid), dim3(block), shm, s, a, lds_depth); break;
        case 8: hipLaunchKernelGGL(trace_persistent<2>, dim3(grid), dim3(block), shm, s, a, lds_depth); break;
        default: hipLaunchKernelGGL(trace_persistent<4>, dim3(grid), dim3(block), shm, s, a, lds_depth); break;
    }
}
#endif

#if defined(OXD_FORMS_b0)
void launch_batch_m0(const BatchArgs &a, uint32_t kernel, uint32_t block, uint32_t grid, uint32_t lds_depth,
                     size_t shm, bool, hipStream_t s) {
    switch (kernel) {
        case 2: hipLaunchKernelGGL(trace_batch<3>, dim3(grid), dim3(block), shm, s, a, a.jobs, lds_depth); break;
        case 4: hipLaunchKernelGGL(trace_batch<5>, dim3(grid), dim3(block), shm, s, a, a.jobs, lds_depth); break;
        case 5: hipLaunchKernelGGL(trace_batch<6>, dim3(grid), dim3(block), shm, s, a, a.jobs, lds_depth); break;
        case 6: hipLaunchKernelGGL(trace_batch<8>, dim3(grid), dim3(block), shm, s, a, a.jobs, lds_depth); break;
        case 8: hipLaunchKernelGGL(trace_batch<2>, dim3(grid), dim3(block), shm, s, a, a.jobs, lds_depth); break;
        default: hipLaunchKernelGGL(trace_batch<4>, dim3(grid), dim3(block), shm, s, a, a.jobs, lds_depth); break;
    }
}
#endif

#if defined(OXD_PERSISTENT)
template <bool Opaque>
static void persistent_forms(const TraceArgs &a, uint32_t kernel, uint32_t block, uint32_t grid, uint32_t lds_depth,
                             size_t shm, hipStream_t s) {
    switch (kernel) {
        case 2:
            hipLaunchKernelGGL((trace_persistent<3, OXD_MODE, Opaque>), dim3(grid), dim3(block), shm, s, a, lds_depth);
            break;
        case 4:
            hipLaunchKernelGGL((trace_persistent<5, OXD_MODE, Opaque>), dim3(grid), dim3(block), shm, s, a, lds_depth);
            break;
        case 5:
            hipLaunchKernelGGL((trace_persistent<6, OXD_MODE, Opaque>), dim3(grid), dim3(block), shm, s, a, lds_depth);
            break;
        default:
            hipLaunchKernelGGL((trace_persistent<4, OXD_MODE, Opaque>), dim3(grid), dim3(block), shm, s, a, lds_depth);
            break;
    }
}
void OXD_PERSISTENT(const TraceArgs &a, uint32_t kernel, uint32_t block, uint32_t grid, uint32_t lds_depth, size_t shm,
                    bool opaque, hipStream_t s) {
    if (opaque) persistent_forms<true>(a, kernel, block, grid, lds_depth, shm, s);
    else persistent_forms<false>(a, kernel, block, grid, lds_depth, shm, s);
}
#endif

#if defined(OXD_BATCH)
template <bool Opaque>
static void batch_forms(const BatchArgs &a, uint32_t kernel, uint32_t block, uint32_t grid, uint32_t lds_depth,
                        size_t shm, hipStream_t s) {
    switch (kernel) {
        case 2:
            hipLaunchKernelGGL((trace_batch<3, OXD_MODE, Opaque>), dim3(grid), dim3(block), shm, s, a, a.jobs, lds_depth);
            break;
        case 4:
            hipLaunchKernelGGL((trace_batch<5, OXD_MODE, Opaque>), dim3(grid), dim3(block), shm, s, a, a.jobs, lds_depth);
            break;
        case 5:
            hipLaunchKernelGGL((trace_batch<6, OXD_MODE, Opaque>), dim3(grid), dim3(block), shm, s, a, a.jobs, lds_depth);
            break;
        default:
            hipLaunchKernelGGL((trace_batch<4, OXD_MODE, Opaque>), dim3(grid), dim3(block), shm, s, a, a.jobs, lds_depth);
            break;
    }
}
void OXD_BATCH(const BatchArgs &a, uint32_t kernel, uint32_t block, uint32_t grid, uint32_t lds_depth, size_t shm,
               bool opaque, hipStream_t s) {
    if (opaque) batch_forms<true>(a, kernel, block, grid, lds_depth, shm, s);
    else batch_forms<false>(a, kernel, block, grid, lds_depth, shm, s);
}
#endif

}  // namespace oxd
