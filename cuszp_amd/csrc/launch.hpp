// Host-side launch interface between the C ABI (cuszp_api.cpp) and the kernel
// instantiations (kern_f32.hip, kern_f64.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "cuszp_kernels.hpp"

namespace cuszp {

// Groups (64-element units) per wave; a tile is kWaves * kGroupsPerWave groups.
constexpr int kGroupsPerWave = 16;
constexpr uint64_t kGroupsPerTile = (uint64_t)kWaves * kGroupsPerWave;

hipError_t launch_compress_f32(int dim, int mode, const float* in, uint8_t* out, const Geom& g,
                               float eb, const ScanArgs& sa, uint64_t cap, hipStream_t st);
hipError_t launch_compress_f64(int dim, int mode, const double* in, uint8_t* out, const Geom& g,
                               double eb, const ScanArgs& sa, uint64_t cap, hipStream_t st);
hipError_t launch_decompress_f32(int dim, int mode, float* out, const uint8_t* cmp, const Geom& g,
                                 float eb, const ScanArgs& sa, uint64_t cmp_size, hipStream_t st);
hipError_t launch_decompress_f64(int dim, int mode, double* out, const uint8_t* cmp, const Geom& g,
                                 double eb, const ScanArgs& sa, uint64_t cmp_size, hipStream_t st);

}  // namespace cuszp
