// Host-side launch interface between the C ABI (cuszp_api.cpp) and the kernel
// instantiations (kern_f32.hip, kern_f64.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "cuszp_kernels.hpp"

namespace cuszp {

// Groups (64-element units) per wave; a tile is waves * kGroupsPerWave groups.
constexpr int kGroupsPerWave = 16;

// Tile plan: waves per workgroup (4, 8 or 16) and the number of tiles.
// Larger tiles cut the number of decoupled-lookback steps on big arrays;
// CUSZP_WAVES in the environment overrides the choice (tuning experiments).
struct TilePlan {
    int waves;
    uint64_t ntiles;
};
TilePlan plan_tiles(int dim, const Geom& g);

hipError_t launch_compress_f32(int dim, int mode, int waves, const float* in, uint8_t* out, const Geom& g,
                               float eb, const ScanArgs& sa, uint64_t cap, hipStream_t st);
hipError_t launch_compress_f64(int dim, int mode, int waves, const double* in, uint8_t* out, const Geom& g,
                               double eb, const ScanArgs& sa, uint64_t cap, hipStream_t st);
hipError_t launch_decompress_f32(int dim, int mode, int waves, float* out, const uint8_t* cmp, const Geom& g,
                                 float eb, const ScanArgs& sa, uint64_t cmp_size, hipStream_t st);
hipError_t launch_decompress_f64(int dim, int mode, int waves, double* out, const uint8_t* cmp, const Geom& g,
                                 double eb, const ScanArgs& sa, uint64_t cmp_size, hipStream_t st);

// Workgroups of `kernel` that can be resident on one CU: the smaller of HIP's
// occupancy query and a register model. The HIP query only sees VGPRs; these
// kernels also use ~106 SGPRs, which caps them at 7 waves per SIMD (800 SGPRs
// per SIMD in 16-register granules).
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

// Divisor applied to every persistent grid after a scan stall (see kErrStall);
// the C ABI doubles it and retries, so a wrong residency estimate costs one
// retry instead of a hang.
int& grid_divisor(int dev);

// Launch of a persistent grid: as many workgroups as can be resident at once
// (never more than ntiles), so every tile a workgroup waits on belongs to a
// running workgroup.
template <class Kernel, class... Args>
hipError_t launch_persistent(Kernel kernel, int threads, uint64_t ntiles, hipStream_t st, Args... args) {
    static int max_blocks[64] = {0};   // per device
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    int& mb = max_blocks[dev & 63];
    if (!mb) {
        int cus = 0;
        e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        if (e != hipSuccess) return e;
        const int per_cu = resident_per_cu(kernel, threads);
        if (per_cu <= 0 || cus <= 0) return hipErrorInvalidConfiguration;
        mb = per_cu * cus;
    }
    int lim = mb / grid_divisor(dev);
    if (lim < 1) lim = 1;
    const unsigned grid = (unsigned)(ntiles < (uint64_t)lim ? ntiles : (uint64_t)lim);
    hipLaunchKernelGGL(kernel, dim3(grid), dim3(threads), 0, st, args...);
    return hipGetLastError();
}

}  // namespace cuszp
