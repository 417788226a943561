// Batched compress kernel instantiations (4 or 8 waves per workgroup, lookback
// windows of 2 or 4 words per lane) and launchers.
#include "launch.hpp"

namespace cuszp {

// Advances the scan epoch of a batched plan; runs (one thread) before every
// batched launch, also inside captured graphs.
__global__ void bump_epoch_kernel(unsigned long long* epoch) {
    if (threadIdx.x == 0) epoch[0] = epoch[0] + 1;
}

hipError_t launch_compress_batch(int waves, int window, const Item* items, uint32_t nitems, uint64_t ntiles,
                                 const ScanCtx& sc, const unsigned long long* epoch, hipStream_t st) {
    if (!ntiles) return hipSuccess;
#define CUSZP_LAUNCH_C(W, K)                                                                            \
    if (waves == W && window == K)                                                                    \
        return launch_tiles(compress_batch_kernel<W, K>, W * kWaveSize, ntiles, st, items, nitems, sc, epoch);
    CUSZP_LAUNCH_C(4, 2)
    CUSZP_LAUNCH_C(4, 4)
    CUSZP_LAUNCH_C(8, 2)
    CUSZP_LAUNCH_C(8, 4)
#undef CUSZP_LAUNCH_C
    return hipErrorInvalidValue;
}

hipError_t launch_bump_epoch(unsigned long long* epoch, hipStream_t st) {
    hipLaunchKernelGGL(bump_epoch_kernel, dim3(1), dim3(64), 0, st, epoch);
    return hipGetLastError();
}

}  // namespace cuszp
