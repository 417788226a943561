// Host-side launch interface between the C ABI (cuszp_api.cpp, cuszp_batch.cpp)
// and the kernel instantiations (kern_f32.hip, kern_f64.hip, kern_batch_*.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "cuszp_kernels.hpp"

namespace cuszp {

// Tile plan: waves per workgroup (4 or 8) and the number of tiles (tile
// shape: tile_groups).
// CUSZP_WAVES in the environment overrides the choice (tuning experiments).
struct TilePlan {
    int waves;
    uint64_t ntiles;
};
TilePlan plan_tiles(int dim, int type, bool compress, const Geom& g);

// Single-array launches; item.param is 0.5/eb (compress) or eb (decompress).
hipError_t launch_compress_f32(int dim, int mode, int waves, const Item& item, const ScanCtx& sc, hipStream_t st);
hipError_t launch_compress_f64(int dim, int mode, int waves, const Item& item, const ScanCtx& sc, hipStream_t st);
// Single-array decompression reads each wave's payload offset from `offs`,
// which launch_unit_offsets_* fills first on the same stream.
hipError_t launch_decompress_f32(int dim, int mode, int waves, const Item& item, const ScanCtx& sc,
                                 const unsigned long long* offs, hipStream_t st);
hipError_t launch_decompress_f64(int dim, int mode, int waves, const Item& item, const ScanCtx& sc,
                                 const unsigned long long* offs, hipStream_t st);
// Offset table of a single-array decompression (unit_offsets_kernel): offs[u]
// for the nunits decode waves of `item`'s plan, from the header bytes of the
// stream at item.src; sc's status words serve the pass's own scan (one per
// segment of kPrefixSegBytes header bytes).
hipError_t launch_unit_offsets_f32(int dim, int mode, const Item& item, uint64_t nunits, unsigned long long* offs,
                                   const ScanCtx& sc, hipStream_t st);
hipError_t launch_unit_offsets_f64(int dim, int mode, const Item& item, uint64_t nunits, unsigned long long* offs,
                                   const ScanCtx& sc, hipStream_t st);
// Blocks one decode wave handles (kChunks chunks of G groups).
inline uint64_t unit_blocks(int dim, bool f64) { return (uint64_t)kChunks * (f64 ? 4 : 8) * (dim == 1 ? 2 : 1); }
// Segments of the offset pass over `nunits` units.
inline uint64_t unit_segments(int dim, bool f64, uint64_t nunits) {
    return (nunits * unit_blocks(dim, f64) + kPrefixSegBytes - 1) / kPrefixSegBytes;
}

// Batched launches over a device table of items (ordered by tile0), one
// workgroup per tile; the scan tag is read from *epoch when the kernel starts
// (graph-replay safe). waves: 4 or 8 per workgroup; window: lookback words per
// lane, 2 or 4 (128 or 256 predecessors per round trip).
hipError_t launch_compress_batch(int waves, int window, const Item* items, uint32_t nitems, uint64_t ntiles,
                                 const ScanCtx& sc, const unsigned long long* epoch, hipStream_t st);
hipError_t launch_decompress_batch(int waves, int window, const Item* items, uint32_t nitems, uint64_t ntiles,
                                   const ScanCtx& sc, const unsigned long long* epoch, hipStream_t st);
// Increments *epoch (one thread); runs before each batched launch.
hipError_t launch_bump_epoch(unsigned long long* epoch, hipStream_t st);
// Largest grid of a tile launch; arrays with more tiles take several launches.
constexpr uint64_t kMaxGrid = 1ull << 22;

// Launches with one workgroup per tile: tile = base + blockIdx.x, in launches
// of up to kMaxGrid tiles. Tiles are numbered in launch order, so the
// predecessors a tile waits on normally belong to workgroups that started
// earlier; the lookback's stealing covers any order.
template <class Kernel, class... Args>
hipError_t launch_tiles(Kernel kernel, int threads, uint64_t ntiles, hipStream_t st, Args... args) {
    for (uint64_t base = 0; base < ntiles; base += kMaxGrid) {
        const uint64_t n = ntiles - base < kMaxGrid ? ntiles - base : kMaxGrid;
        hipLaunchKernelGGL(kernel, dim3((unsigned)n), dim3(threads), 0, st, args..., base);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

}  // namespace cuszp
