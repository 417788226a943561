// cuSZp_compress_noalloc: single-array compression that allocates nothing.
//
// The other entries keep their scan state (tile status words, sticky error
// bits, watchdog words) in a per-(device, stream) pool that grows with
// hipMalloc on demand. Callers that cannot allocate, such as code under HIP
// graph capture or serving loops that forbid allocation, use this entry
// instead: its scan state lives in `scratch`, a fixed device array of this
// library, and the launches are kernels only (no hipMemsetAsync, no copies).
//
// Limits and sharing:
//   * at most kNoAllocTiles tiles (2^31 f32 or 2^30 f64 values, 8 GiB of
//     either); larger arrays get CUSZP_ERROR_TOO_LARGE before anything runs;
//   * each GPU holds one copy of `scratch` (about 514 KiB) until the process
//     ends, and every call on that GPU uses it: calls that overlap on one GPU
//     (other streams, other host threads) race and must be serialised by the
//     caller. Calls on different GPUs are independent.
//   * a small kernel clears the words a call uses (sticky bits, watchdog
//     words and the status words of its tiles) before its compress kernel, so
//     every call, also one replayed from a captured graph, starts from the
//     same state and the launch tag can stay fixed.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <mutex>

#include "../../include/cuSZp.h"
#include "common.hpp"

namespace cuszp {
namespace {

// Tiles of the largest array (a tile holds 32,768 f32 or 16,384 f64 values).
constexpr uint64_t kNoAllocTiles = 1ull << 16;
// Layout of `scratch` (64-bit words): [0, kMiscWords) sticky bits and
// watchdog words as in the pooled entries, then kStatusGuard guard words and
// kNoAllocTiles status words.
constexpr uint64_t kScratchWords = kMiscWords + kStatusGuard + kNoAllocTiles;

// Zeroes words [0, n) of p (grid-stride; any grid).
__global__ void clear_words(unsigned long long* __restrict__ p, uint64_t n) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = 0;
}

// result[1] = the call's status code from its sticky bits (result[0], the
// stream length, is written by the compress kernel's last tile); with
// `empty`, the call had no values: result = {0, success}.
__global__ void finish_result(const unsigned int* __restrict__ err, unsigned long long* __restrict__ result,
                              int empty) {
    if (threadIdx.x != 0) return;
    if (empty) {
        result[0] = 0;
        result[1] = CUSZP_SUCCESS;
        return;
    }
    // the precedence of status_from_bits: errors, then the lossy-block warning
    const unsigned int bits = *err;
    int st = CUSZP_SUCCESS;
    if (bits & kErrStall) st = CUSZP_ERROR_INTERNAL;
    else if (bits & kErrOverflow) st = CUSZP_ERROR_OUTPUT_TOO_SMALL;
    else if (bits & kWarnLossy) st = CUSZP_WARNING_BOUND_EXCEEDED;
    result[1] = (unsigned long long)st;
}

// Launch tag of every call (its status words are cleared before each launch).
constexpr unsigned long long kNoAllocTag = 1ull << kEpochShift;

// Per device: the device address of `scratch` (hipGetSymbolAddress, taken once
// with the device current), or null before the first call on the device.
constexpr int kMaxDevices = 64;
std::mutex g_mu;
unsigned long long* g_scratch[kMaxDevices];

}  // namespace

// The scan state of cuSZp_compress_noalloc, one copy per GPU. In host code
// this name is only a stand-in: kernels get the address from g_scratch.
static __device__ __align__(256) unsigned long long scratch[kScratchWords];

}  // namespace cuszp

using namespace cuszp;

extern "C" int cuSZp_compress_noalloc(const void* d_oriData, unsigned char* d_cmpBytes, size_t cmpCapacity,
                                      size_t nbEle, unsigned long long* d_result, double errorBound,
                                      cuszp_dim_t dim, cuszp_dims_t dims, cuszp_type_t type, cuszp_mode_t mode,
                                      hipStream_t stream) {
    // everything that can be checked without the device first
    if (bad_enums(dim, type, mode) || bad_eb(errorBound, type) || !d_result) return CUSZP_ERROR_INVALID_ARGUMENT;
    Geom g;
    int rc = make_geom(dim, nbEle, dims, g);
    if (rc) return rc;
    TilePlan plan{4, 0};
    if (nbEle) {
        if (!d_oriData || !d_cmpBytes) return CUSZP_ERROR_INVALID_ARGUMENT;
        if (max_stream_bytes(dim, g) > kMaxStream) return CUSZP_ERROR_TOO_LARGE;
        plan = plan_tiles(dim, type, true, g);
        if (plan.ntiles > kNoAllocTiles) {
            std::fprintf(stderr,
                         "cuSZp_compress_noalloc: %zu values need %llu tiles; the fixed scratch holds %llu "
                         "(2^31 f32 or 2^30 f64 values)\n",
                         nbEle, (unsigned long long)plan.ntiles, (unsigned long long)kNoAllocTiles);
            return CUSZP_ERROR_TOO_LARGE;
        }
        if (cmpCapacity < g.H) return CUSZP_ERROR_OUTPUT_TOO_SMALL;
        g.ntiles = plan.ntiles;
    }
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return CUSZP_ERROR_HIP;
    unsigned long long* base;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        if (!g_scratch[dev]) {
            // this device's copy of the array; the device is current here
            void* p = nullptr;
            if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return CUSZP_ERROR_HIP;
            g_scratch[dev] = (unsigned long long*)p;
        }
        base = g_scratch[dev];
    }
    unsigned long long* const misc = base;
    unsigned long long* const status = base + kMiscWords;   // kStatusGuard guard words, then the tiles'
    unsigned int* const err = (unsigned int*)(misc + kMiscErr);
    // the words this call reads start at zero: sticky bits, watchdog words and
    // the status words of its tiles (a zero word reads as "not published")
    const uint64_t nclear = kMiscWords + kStatusGuard + g.ntiles;
    const uint64_t cblocks = (nclear + 255) / 256;
    hipLaunchKernelGGL(clear_words, dim3((unsigned)(cblocks < 256 ? cblocks : 256)), dim3(256), 0, stream, misc,
                       nclear);
    if (hipGetLastError() != hipSuccess) return CUSZP_ERROR_HIP;
    if (nbEle) {
        const ScanCtx sc = make_scan_ctx(status + kStatusGuard, err, misc + kMiscDiag, kNoAllocTag);
        const Item item = make_item(d_oriData, d_cmpBytes, cmpCapacity, d_result, err, compress_param(errorBound, type),
                                    dim, type, mode, g);
        const hipError_t e = type == CUSZP_TYPE_FLOAT ? launch_compress_f32(dim, mode, plan.waves, item, sc, stream)
                                                      : launch_compress_f64(dim, mode, plan.waves, item, sc, stream);
        if (e != hipSuccess) return CUSZP_ERROR_HIP;
    }
    hipLaunchKernelGGL(finish_result, dim3(1), dim3(64), 0, stream, (const unsigned int*)err, d_result,
                       nbEle ? 0 : 1);
    if (hipGetLastError() != hipSuccess) return CUSZP_ERROR_HIP;
    return CUSZP_SUCCESS;
}
