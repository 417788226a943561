// Batched decompress kernel instantiations (4 and 8 waves per workgroup) and launchers.
#include "launch.hpp"

namespace cuszp {

hipError_t launch_decompress_batch(int waves, const Item* items, uint32_t nitems, uint64_t ntiles,
                                   const ScanCtx& sc, const unsigned long long* epoch, hipStream_t st) {
    if (!ntiles) return hipSuccess;
    if (waves == 4) return launch_tiles(decompress_batch_kernel<4>, 256, ntiles, st, items, nitems, sc, epoch);
    if (waves == 8) return launch_tiles(decompress_batch_kernel<8>, 512, ntiles, st, items, nitems, sc, epoch);
    return hipErrorInvalidValue;
}

}  // namespace cuszp
