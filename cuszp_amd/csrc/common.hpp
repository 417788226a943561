// Argument validation, geometry and item construction shared by the
// single-array C ABI (cuszp_api.cpp) and the batched one (cuszp_batch.cpp).
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <cmath>
#include <cstring>

#include "../../include/cuSZp.h"
#include "launch.hpp"

namespace cuszp {

constexpr uint64_t kMaxStream = (1ull << 38) - 1;   // status-word value field

// Status of a launch's (or an array's) sticky bits: errors first, then the
// lossy-block warning; kInfoStolen is information only.
inline int status_from_bits(unsigned long long bits) {
    if (bits & kErrStall) return CUSZP_ERROR_INTERNAL;
    if (bits & kErrOverflow) return CUSZP_ERROR_OUTPUT_TOO_SMALL;
    if (bits & kErrTruncated) return CUSZP_ERROR_CORRUPT_STREAM;
    if (bits & kWarnLossy) return CUSZP_WARNING_BOUND_EXCEEDED;
    return CUSZP_SUCCESS;
}
// A status after which the output is complete (a warning still delivers it).
inline bool status_delivers(int st) { return st == CUSZP_SUCCESS || st == CUSZP_WARNING_BOUND_EXCEEDED; }

// Scan settings of later launches (cuSZp_scan_config), shared by every entry.
struct ScanSettings {
    std::atomic<uint32_t> steal_ticks{(uint32_t)kStealTicks};
    std::atomic<uint32_t> hold_every{0};
    std::atomic<uint32_t> hold_ticks{0};
    std::atomic<uint32_t> stats{0};
};
inline ScanSettings& scan_settings() {
    static ScanSettings s;
    return s;
}

// Allocation of a scan-state status array (the arrays that grow with the
// work: single-array slots, batch plans, serving lanes). A test can make the
// next `n` of these fail as if the device were out of memory
// (cuSZp_debug_fail_allocs), to check that a failed growth leaves the state
// usable by the next call.
inline std::atomic<uint32_t>& failing_allocs() {
    static std::atomic<uint32_t> n{0};
    return n;
}
inline hipError_t scan_state_alloc(unsigned long long** p, size_t bytes) {
    uint32_t n = failing_allocs().load(std::memory_order_relaxed);
    while (n && !failing_allocs().compare_exchange_weak(n, n - 1, std::memory_order_relaxed)) {
    }
    if (n) {
        *p = nullptr;
        return hipErrorOutOfMemory;
    }
    return hipMalloc((void**)p, bytes);
}

// Scan context of a launch over `status` (kStatusGuard guard words in front
// are the caller's), with the launch's error word and the diagnostic words
// diag[0..7].
inline ScanCtx make_scan_ctx(unsigned long long* status, unsigned int* err, unsigned long long* diag,
                             unsigned long long tag) {
    const ScanSettings& k = scan_settings();
    ScanCtx c;
    c.status = status;
    c.err = err;
    c.diag = diag;
    c.tag = tag;
    c.steal_ticks = k.steal_ticks.load(std::memory_order_relaxed);
    c.hold_every = k.hold_every.load(std::memory_order_relaxed);
    c.hold_ticks = k.hold_ticks.load(std::memory_order_relaxed);
    c.stats = k.stats.load(std::memory_order_relaxed);
    if (!c.hold_ticks) c.hold_every = 0;
    return c;
}

// Device words of a launch's scan state besides the status array:
// [0] epoch (batched launches), [1] stream length (blocking entries),
// [2] sticky bits, [4..7] watchdog diagnostics, [8..11] lookback counters.
constexpr int kMiscWords = 16;
constexpr int kMiscErr = 2, kMiscDiag = 4, kMiscCounters = 8;

// Geometry of a (dim, nbEle, dims) triple; mirrors the reference's block counts.
inline int make_geom(int dim, size_t n, cuszp_dims_t d, Geom& g) {
    std::memset(&g, 0, sizeof(g));
    if (dim == CUSZP_DIM_1D) {
        g.n = n;
        g.nblocks = (n + 31) / 32;
        g.H = (n + 32767) / 32768 * 1024;          // src/cuSZp_entry_1D_f32.cu:43
        return CUSZP_SUCCESS;
    }
    if (dim != CUSZP_DIM_2D && dim != CUSZP_DIM_3D) return CUSZP_ERROR_INVALID_ARGUMENT;
    if (!d.x || !d.y || !d.z || (d.x % 4)) return CUSZP_ERROR_INVALID_DIMS;   // float4 loads in the reference
    const uint64_t xyz = (uint64_t)d.x * d.y * d.z;
    if ((uint64_t)n != xyz) return CUSZP_ERROR_INVALID_DIMS;
    g.n = n;
    g.X = d.x; g.Y = d.y; g.Z = d.z;
    g.XY = (uint64_t)d.x * d.y;
    uint64_t bz;
    if (dim == CUSZP_DIM_2D) {                      // src/cuSZp_entry_2D_f32.cu:21-23
        g.Bx = (d.x + 7) / 8; g.By = (d.y + 7) / 8; bz = d.z;
    } else {                                        // src/cuSZp_entry_3D_f32.cu:21-24
        g.Bx = (d.x + 3) / 4; g.By = (d.y + 3) / 4; bz = (d.z + 3) / 4;
    }
    g.nblocks = (uint64_t)g.Bx * g.By * bz;
    g.aligned = (dim == CUSZP_DIM_2D) ? (d.x % 8 == 0 && d.y % 8 == 0)
                                      : (d.x % 4 == 0 && d.y % 4 == 0 && d.z % 4 == 0);
    if (g.nblocks >= (1ull << 40)) return CUSZP_ERROR_TOO_LARGE;
    g.H = (g.nblocks + 3) / 4 * 4;                  // src/cuSZp_kernels_2D_f32.cu:46
    return CUSZP_SUCCESS;
}

inline uint64_t max_stream_bytes(int dim, const Geom& g) {
    const uint64_t S = (dim == CUSZP_DIM_1D) ? 4 : 8;
    return g.H + g.nblocks * (S + 32 * S + 4);
}

inline bool bad_enums(int dim, int type, int mode) {
    return (dim < 1 || dim > 3) || (type != CUSZP_TYPE_FLOAT && type != CUSZP_TYPE_DOUBLE) ||
           (mode < CUSZP_MODE_FIXED || mode > CUSZP_MODE_OUTLIER);
}

inline bool bad_eb(double eb, int type) {
    if (type == CUSZP_TYPE_FLOAT) {
        const float f = (float)eb;
        return !(f > 0.0f) || !std::isfinite(f);
    }
    return !(eb > 0.0) || !std::isfinite(eb);
}

inline double decompress_param(double eb, int type) { return type == CUSZP_TYPE_FLOAT ? (double)(float)eb : eb; }
inline double compress_param(double eb, int type) {
    // 0.5f/eb rounded in the array's type (src/cuSZp_kernels_1D_f32.cu:48, _1D_f64.cu:33)
    if (type == CUSZP_TYPE_FLOAT) return (double)(0.5f / (float)eb);
    return 0.5 / eb;
}

inline Item make_item(const void* src, void* dst, uint64_t limit, unsigned long long* total, unsigned int* err,
               double param, int dim, int type, int mode, const Geom& g) {
    Item it;
    std::memset(&it, 0, sizeof(it));
    it.src = src;
    it.dst = dst;
    it.limit = limit;
    it.total = total;
    it.err = err;
    it.param = param;
    it.tile0 = 0;
    it.variant = (uint32_t)(type * 9 + (dim - 1) * 3 + mode);
    it.flags = g.n * (type == CUSZP_TYPE_DOUBLE ? 8u : 4u) >= kStreamNtBytes ? kItemStreamNt : 0u;
    it.geo = g;
    return it;
}

}  // namespace cuszp
