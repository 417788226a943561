// Batched decompress kernel instantiations (4 or 8 waves per workgroup, lookback
// windows of 2 or 4 words per lane) and launchers.
#include "launch.hpp"

namespace cuszp {

hipError_t launch_decompress_batch(int waves, int window, const Item* items, uint32_t nitems, uint64_t ntiles,
                                   const ScanCtx& sc, const unsigned long long* epoch, hipStream_t st) {
    if (!ntiles) return hipSuccess;
#define CUSZP_LAUNCH_D(W, K)                                                                            \
    if (waves == W && window == K)                                                                    \
        return launch_tiles(decompress_batch_kernel<W, K>, W * kWaveSize, ntiles, st, items, nitems, sc, epoch);
    CUSZP_LAUNCH_D(4, 2)
    CUSZP_LAUNCH_D(4, 4)
    CUSZP_LAUNCH_D(8, 2)
    CUSZP_LAUNCH_D(8, 4)
#undef CUSZP_LAUNCH_D
    return hipErrorInvalidValue;
}

}  // namespace cuszp
