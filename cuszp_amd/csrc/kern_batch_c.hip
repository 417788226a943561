// Batched compress kernel instantiations (4 and 8 waves per workgroup) and launchers.
#include "launch.hpp"

namespace cuszp {

// Advances the scan epoch of a batched plan; runs (one thread) before every
// batched launch, also inside captured graphs.
__global__ void bump_epoch_kernel(unsigned long long* epoch) {
    if (threadIdx.x == 0) epoch[0] = epoch[0] + 1;
}

hipError_t launch_compress_batch(int waves, const Item* items, uint32_t nitems, uint64_t ntiles,
                                 const ScanCtx& sc, const unsigned long long* epoch, hipStream_t st) {
    if (!ntiles) return hipSuccess;
    if (waves == 4) return launch_tiles(compress_batch_kernel<4>, 256, ntiles, st, items, nitems, sc, epoch);
    if (waves == 8) return launch_tiles(compress_batch_kernel<8>, 512, ntiles, st, items, nitems, sc, epoch);
    return hipErrorInvalidValue;
}

hipError_t launch_bump_epoch(unsigned long long* epoch, hipStream_t st) {
    hipLaunchKernelGGL(bump_epoch_kernel, dim3(1), dim3(64), 0, st, epoch);
    return hipGetLastError();
}

}  // namespace cuszp
