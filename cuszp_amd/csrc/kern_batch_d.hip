// Batched decompress kernel instantiations (4 and 8 waves per workgroup) and launchers.
#include "launch.hpp"

namespace cuszp {

namespace {
template <class Kernel>
unsigned batch_grid(Kernel kernel, int threads, unsigned grid, uint64_t ntiles) {
    if (!grid) {
        static int max_blocks[64][2] = {};
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess) return 0;
        int& mb = max_blocks[dev & 63][threads == 256 ? 0 : 1];
        if (!mb) mb = device_resident(kernel, threads, dev);
        grid = (unsigned)((int64_t)mb * grid_sixteenths(dev) / 16);
        if (!grid) grid = 1;
    }
    return (uint64_t)grid < ntiles ? grid : (unsigned)ntiles;
}
}  // namespace

int batch_resident_decompress(int waves) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    return waves == 4 ? device_resident(decompress_batch_kernel<4>, 256, dev)
                      : device_resident(decompress_batch_kernel<8>, 512, dev);
}

hipError_t launch_decompress_batch(int waves, unsigned grid, const Item* items, uint32_t nitems, uint64_t ntiles,
                                   const ScanCtx& sc, const unsigned long long* epoch, hipStream_t st) {
    if (!ntiles) return hipSuccess;
    if (waves == 4) {
        const unsigned g = batch_grid(decompress_batch_kernel<4>, 256, grid, ntiles);
        hipLaunchKernelGGL(decompress_batch_kernel<4>, dim3(g), dim3(256), 0, st, items, nitems, ntiles, sc, epoch);
    } else if (waves == 8) {
        const unsigned g = batch_grid(decompress_batch_kernel<8>, 512, grid, ntiles);
        hipLaunchKernelGGL(decompress_batch_kernel<8>, dim3(g), dim3(512), 0, st, items, nitems, ntiles, sc, epoch);
    } else {
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace cuszp
