// Host-side launch interface between the C ABI (cuszp_api.cpp, cuszp_batch.cpp)
// and the kernel instantiations (kern_f32.hip, kern_f64.hip, kern_batch_*.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "cuszp_kernels.hpp"

namespace cuszp {

// Groups (64-element units) per wave; a tile is waves * groups_per_wave(waves)
// groups = 4096 elements, so the double-buffered stream words of a compress
// tile take 32 KB of LDS whatever the workgroup size.
constexpr int groups_per_wave(int waves) { return waves >= 8 ? 8 : 16; }
// Batched kernels: groups per wave by element type (f64 tiles hold the same
// bytes as f32 tiles, and keep the f64 variants' registers in check).
constexpr int kBatchGroupsF32 = 16;
constexpr int kBatchGroupsF64 = 8;

// Tile plan: waves per workgroup (4 or 8) and the number of tiles.
// CUSZP_WAVES in the environment overrides the choice (tuning experiments).
struct TilePlan {
    int waves;
    uint64_t ntiles;
};
TilePlan plan_tiles(int dim, const Geom& g);

// Single-array launches; item.param is 0.5/eb (compress) or eb (decompress).
hipError_t launch_compress_f32(int dim, int mode, int waves, const Item& item, const ScanCtx& sc, hipStream_t st);
hipError_t launch_compress_f64(int dim, int mode, int waves, const Item& item, const ScanCtx& sc, hipStream_t st);
hipError_t launch_decompress_f32(int dim, int mode, int waves, const Item& item, const ScanCtx& sc, hipStream_t st);
hipError_t launch_decompress_f64(int dim, int mode, int waves, const Item& item, const ScanCtx& sc, hipStream_t st);

// Batched launches over a device table of items (ordered by tile0); the scan
// tag is read from *epoch when the kernel starts (graph-replay safe).
// `grid` = 0 picks the persistent grid from the occupancy model.
hipError_t launch_compress_batch(int waves, unsigned grid, const Item* items, uint32_t nitems, uint64_t ntiles,
                                 const ScanCtx& sc, const unsigned long long* epoch, hipStream_t st);
hipError_t launch_decompress_batch(int waves, unsigned grid, const Item* items, uint32_t nitems, uint64_t ntiles,
                                   const ScanCtx& sc, const unsigned long long* epoch, hipStream_t st);
// Increments *epoch (one thread); runs before each batched launch.
hipError_t launch_bump_epoch(unsigned long long* epoch, hipStream_t st);
// Resident workgroups of a batched kernel on the whole device.
int batch_resident(bool compress, int waves);

// Workgroups of `kernel` that can be resident on one CU: the smaller of HIP's
// occupancy query and a register model (the HIP query only sees VGPRs; these
// kernels also use ~106 SGPRs, which caps them at 7 waves per SIMD).
template <class Kernel>
int resident_per_cu(Kernel kernel, int threads) {
    int hip_est = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&hip_est, kernel, threads, 0) != hipSuccess) return 0;
    hipFuncAttributes attr;
    if (hipFuncGetAttributes(&attr, (const void*)kernel) != hipSuccess) return 0;
    const int vgpr = attr.numRegs > 0 ? (attr.numRegs + 7) / 8 * 8 : 512;
    int waves_simd = 512 / vgpr;
    if (waves_simd > 7) waves_simd = 7;
    const int own = waves_simd * 4 / (threads / kWaveSize);
    return hip_est < own ? hip_est : own;
}

// Persistent grids are this many sixteenths of the occupancy estimate. The C
// ABI lowers it by one sixteenth (down to a quarter) each time a launch reports
// kInfoStolen, i.e. some workgroup waited on a tile whose owner was not running.
int& grid_sixteenths(int dev);

template <class Kernel>
int device_resident(Kernel kernel, int threads, int dev) {
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
    const int per_cu = resident_per_cu(kernel, threads);
    return per_cu > 0 && cus > 0 ? per_cu * cus : 0;
}

// Launch of a persistent grid: at most as many workgroups as can be resident
// at once (never more than ntiles).
template <class Kernel, class... Args>
hipError_t launch_persistent(Kernel kernel, int threads, uint64_t ntiles, hipStream_t st, Args... args) {
    static int max_blocks[64] = {0};   // per device
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    int& mb = max_blocks[dev & 63];
    if (!mb) {
        mb = device_resident(kernel, threads, dev);
        if (mb <= 0) return hipErrorInvalidConfiguration;
    }
    int lim = (int)((int64_t)mb * grid_sixteenths(dev) / 16);
    if (lim < 1) lim = 1;
    const unsigned grid = (unsigned)(ntiles < (uint64_t)lim ? ntiles : (uint64_t)lim);
    hipLaunchKernelGGL(kernel, dim3(grid), dim3(threads), 0, st, args...);
    return hipGetLastError();
}

}  // namespace cuszp
