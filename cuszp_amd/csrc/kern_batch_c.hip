// Batched compress kernel instantiations (4 and 8 waves per workgroup) and launchers.
#include "launch.hpp"

namespace cuszp {

// Advances the scan epoch of a batched plan; runs (one thread) before every
// batched launch, also inside captured graphs.
__global__ void bump_epoch_kernel(unsigned long long* epoch) {
    if (threadIdx.x == 0) epoch[0] = epoch[0] + 1;
}

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

int batch_resident_decompress(int waves);   // kern_batch_d.hip

int batch_resident(bool compress, int waves) {
    if (!compress) return batch_resident_decompress(waves);
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    return waves == 4 ? device_resident(compress_batch_kernel<4>, 256, dev)
                      : device_resident(compress_batch_kernel<8>, 512, dev);
}

hipError_t launch_compress_batch(int waves, unsigned grid, const Item* items, uint32_t nitems, uint64_t ntiles,
                                 const ScanCtx& sc, const unsigned long long* epoch, hipStream_t st) {
    if (!ntiles) return hipSuccess;
    if (waves == 4) {
        const unsigned g = batch_grid(compress_batch_kernel<4>, 256, grid, ntiles);
        hipLaunchKernelGGL(compress_batch_kernel<4>, dim3(g), dim3(256), 0, st, items, nitems, ntiles, sc, epoch);
    } else if (waves == 8) {
        const unsigned g = batch_grid(compress_batch_kernel<8>, 512, grid, ntiles);
        hipLaunchKernelGGL(compress_batch_kernel<8>, dim3(g), dim3(512), 0, st, items, nitems, ntiles, sc, epoch);
    } else {
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_bump_epoch(unsigned long long* epoch, hipStream_t st) {
    hipLaunchKernelGGL(bump_epoch_kernel, dim3(1), dim3(64), 0, st, epoch);
    return hipGetLastError();
}

}  // namespace cuszp
