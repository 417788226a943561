// C ABI of the MI355X fused compressor: argument validation, geometry, the
// per-(device, stream) scan-state pool and the blocking / stream-ordered entry
// points declared in include/cuSZp.h.
//
// Reference behaviour mirrored:
//   generic dispatch .......... src/cuSZp.cu:4-165 (float eb widened for f64, :45/:127)
//   per-dimension entries ..... src/cuSZp_entry_{1,2,3}D_f{32,64}.cu
//   launch geometry / sizes ... src/cuSZp_entry_1D_f32.cu:20-43, src/cuSZp_entry_2D_f32.cu:21-47,
//                               src/cuSZp_entry_3D_f32.cu:21-48
// What differs: scratch comes from a persistent pool instead of three
// cudaMalloc/cudaMemset/cudaFree per call, the size readback goes through
// pinned memory with a stream-local synchronisation, and every argument is
// checked before anything is launched.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <utility>

#include "../../include/cuSZp.h"
#include "common.hpp"

namespace cuszp {
namespace {


struct Slot {
    std::mutex mu;
    unsigned long long* d_status = nullptr;   // tile status words (kStatusGuard guard words first)
    uint64_t cap_tiles = 0;
    unsigned long long* d_misc = nullptr;     // kMiscWords: stream length, sticky bits, diagnostics, counters
    unsigned long long* h_res = nullptr;      // pinned host copies of d_misc words
    uint32_t epoch = 0;
    unsigned info = 0;                        // information bits harvested since the last scan report
    unsigned long long* d_offs = nullptr;     // decompression: payload offset of every decode wave
    uint64_t cap_units = 0;
};

std::mutex g_pool_mu;
std::map<std::pair<int, hipStream_t>, std::unique_ptr<Slot>> g_pool;

Slot& slot_for(int dev, hipStream_t st) {
    std::lock_guard<std::mutex> lk(g_pool_mu);
    auto& p = g_pool[{dev, st}];
    if (!p) p.reset(new Slot());
    return *p;
}

#define CUSZP_HIP(call)                                \
    do {                                               \
        hipError_t e_ = (call);                        \
        if (e_ != hipSuccess) return CUSZP_ERROR_HIP;  \
    } while (0)

// Make room for `ntiles` status words and advance the epoch. Caller holds s.mu.
// A failed allocation leaves the slot consistent: either the old buffer and
// capacity are kept, or there is no buffer and the capacity is 0.
int prepare(Slot& s, uint64_t ntiles, hipStream_t st) {
    if (!s.d_misc) {
        unsigned long long* m = nullptr;
        CUSZP_HIP(hipMalloc((void**)&m, kMiscWords * sizeof(unsigned long long)));
        if (hipHostMalloc((void**)&s.h_res, kMiscWords * sizeof(unsigned long long), hipHostMallocDefault) !=
            hipSuccess) {
            hipFree(m);
            s.h_res = nullptr;
            return CUSZP_ERROR_HIP;
        }
        s.d_misc = m;
        CUSZP_HIP(hipMemsetAsync(s.d_misc, 0, kMiscWords * sizeof(unsigned long long), st));
    }
    if (ntiles > s.cap_tiles) {
        uint64_t cap = s.cap_tiles * 2;
        if (cap < ntiles) cap = ntiles;
        if (cap < 4096) cap = 4096;
        if (s.d_status) {
            // the old array may still be read by queued launches on this stream
            CUSZP_HIP(hipStreamSynchronize(st));
            CUSZP_HIP(hipFree(s.d_status));
            s.d_status = nullptr;
            s.cap_tiles = 0;
        }
        // kStatusGuard words in front of the status words (lookback windows of
        // the first tiles read there)
        unsigned long long* d = nullptr;
        CUSZP_HIP(scan_state_alloc(&d, (kStatusGuard + cap) * sizeof(unsigned long long)));
        s.d_status = d;
        s.cap_tiles = cap;
        CUSZP_HIP(hipMemsetAsync(s.d_status, 0, (kStatusGuard + cap) * sizeof(unsigned long long), st));
        s.epoch = 0;
    }
    if (++s.epoch >= (1u << 24)) {   // 24-bit epoch tag wrapped: clear stale words
        CUSZP_HIP(hipMemsetAsync(s.d_status, 0, (kStatusGuard + s.cap_tiles) * sizeof(unsigned long long), st));
        s.epoch = 1;
    }
    return CUSZP_SUCCESS;
}

// Make room for `nunits` decompression offsets. Caller holds s.mu; a failed
// allocation leaves no table and capacity 0.
int prepare_offsets(Slot& s, uint64_t nunits, hipStream_t st) {
    if (nunits <= s.cap_units) return CUSZP_SUCCESS;
    uint64_t cap = s.cap_units * 2;
    if (cap < nunits) cap = nunits;
    if (cap < 32768) cap = 32768;
    if (s.d_offs) {
        CUSZP_HIP(hipStreamSynchronize(st));   // queued launches on this stream may still read it
        CUSZP_HIP(hipFree(s.d_offs));
        s.d_offs = nullptr;
        s.cap_units = 0;
    }
    unsigned long long* d = nullptr;
    CUSZP_HIP(scan_state_alloc(&d, cap * sizeof(unsigned long long)));
    s.d_offs = d;
    s.cap_units = cap;
    return CUSZP_SUCCESS;
}

ScanCtx scan_ctx(const Slot& s) {
    return make_scan_ctx(s.d_status + kStatusGuard, (unsigned int*)(s.d_misc + kMiscErr), s.d_misc + kMiscDiag,
                         (unsigned long long)s.epoch << kEpochShift);
}

// Read {total, sticky bits} after a blocking launch (or for the async status);
// clear the bits once seen and keep the information bits for the scan report.
int finish_blocking(Slot& s, hipStream_t st, size_t* total) {
    CUSZP_HIP(hipMemcpyAsync(s.h_res, s.d_misc + 1, 2 * sizeof(unsigned long long), hipMemcpyDeviceToHost, st));
    CUSZP_HIP(hipStreamSynchronize(st));
    const unsigned long long bits = s.h_res[1] & 0xffffffffull;
    if (total) *total = (size_t)s.h_res[0];
    if (bits & kErrStall) {
        CUSZP_HIP(hipMemcpy(s.h_res + 4, s.d_misc + kMiscDiag, 3 * sizeof(unsigned long long), hipMemcpyDeviceToHost));
        std::fprintf(stderr, "cuSZp: internal error: tile %llu stopped waiting for tile %llu (status word 0x%016llx)\n",
                     s.h_res[4] - 1, s.h_res[5], s.h_res[6]);
    }
    if (bits) {
        s.info |= (unsigned)(bits & kInfoStolen);
        CUSZP_HIP(hipMemsetAsync(s.d_misc + kMiscErr, 0, (kMiscCounters - kMiscErr) * sizeof(unsigned long long), st));
        CUSZP_HIP(hipStreamSynchronize(st));
    }
    return status_from_bits(bits);
}

}  // namespace

TilePlan plan_tiles(int dim, int type, bool compress, const Geom& g) {
    static const int forced = [] {
        const char* e = std::getenv("CUSZP_WAVES");
        const int w = e ? std::atoi(e) : 0;
        return (w == 4 || w == 8) ? w : 0;
    }();
    const uint64_t groups = (dim == CUSZP_DIM_1D) ? g.H / 2 : g.nblocks;   // 64-element units
    int waves = forced;
    if (!waves) {
        // 8-wave tiles halve the lookbacks per byte (measured faster from 16 Mi
        // values up); arrays of fewer than 128 such tiles (16 MiB) keep 4-wave
        // tiles, which spread over twice as many workgroups
        const uint64_t per8 = tile_groups(type == CUSZP_TYPE_DOUBLE, 8);
        waves = (groups + per8 - 1) / per8 >= 128 ? 8 : 4;
    }
    (void)compress;   // both directions use the same tile shape
    const uint64_t per_tile = tile_groups(type == CUSZP_TYPE_DOUBLE, waves);
    return TilePlan{waves, (groups + per_tile - 1) / per_tile};
}

namespace {

int compress_once(const void* in, unsigned char* out, size_t cap, size_t n, size_t* cmp_size,
                  unsigned long long* d_size, double eb, int dim, cuszp_dims_t dims, int type, int mode,
                  hipStream_t st, bool blocking) {
    if (cmp_size) *cmp_size = 0;
    if (bad_enums(dim, type, mode) || bad_eb(eb, type)) return CUSZP_ERROR_INVALID_ARGUMENT;
    Geom g;
    int rc = make_geom(dim, n, dims, g);
    if (rc) return rc;
    if (n == 0) {
        if (d_size) CUSZP_HIP(hipMemsetAsync(d_size, 0, sizeof(*d_size), st));
        return CUSZP_SUCCESS;
    }
    if (!in || !out) return CUSZP_ERROR_INVALID_ARGUMENT;
    if (max_stream_bytes(dim, g) > kMaxStream) return CUSZP_ERROR_TOO_LARGE;
    if (cap < g.H) return CUSZP_ERROR_OUTPUT_TOO_SMALL;
    const TilePlan plan = plan_tiles(dim, type, true, g);
    g.ntiles = plan.ntiles;
    int dev = 0;
    CUSZP_HIP(hipGetDevice(&dev));
    Slot& s = slot_for(dev, st);
    std::lock_guard<std::mutex> lk(s.mu);
    rc = prepare(s, g.ntiles, st);
    if (rc) return rc;
    const ScanCtx sc = scan_ctx(s);
    Item item = make_item(in, out, cap, d_size ? d_size : s.d_misc + 1, (unsigned int*)(s.d_misc + 2),
                          compress_param(eb, type), dim, type, mode, g);
    hipError_t e;
    if (type == CUSZP_TYPE_FLOAT) e = launch_compress_f32(dim, mode, plan.waves, item, sc, st);
    else e = launch_compress_f64(dim, mode, plan.waves, item, sc, st);
    if (e != hipSuccess) return CUSZP_ERROR_HIP;
    if (!blocking) return CUSZP_SUCCESS;
    return finish_blocking(s, st, cmp_size);
}

int decompress_once(void* out, const unsigned char* cmp, size_t n, size_t cmp_size, double eb, int dim,
                    cuszp_dims_t dims, int type, int mode, hipStream_t st, bool blocking) {
    if (bad_enums(dim, type, mode) || bad_eb(eb, type)) return CUSZP_ERROR_INVALID_ARGUMENT;
    Geom g;
    int rc = make_geom(dim, n, dims, g);
    if (rc) return rc;
    if (n == 0) return CUSZP_SUCCESS;
    if (!out || !cmp) return CUSZP_ERROR_INVALID_ARGUMENT;
    if (max_stream_bytes(dim, g) > kMaxStream) return CUSZP_ERROR_TOO_LARGE;
    if (cmp_size < g.H) return CUSZP_ERROR_CORRUPT_STREAM;
    const TilePlan plan = plan_tiles(dim, type, false, g);
    g.ntiles = plan.ntiles;
    const bool f64 = type == CUSZP_TYPE_DOUBLE;
    const uint64_t nunits = plan.ntiles * (uint64_t)plan.waves;   // decode waves
    int dev = 0;
    CUSZP_HIP(hipGetDevice(&dev));
    Slot& s = slot_for(dev, st);
    std::lock_guard<std::mutex> lk(s.mu);
    // the offset pass scans its segments with the slot's status words
    rc = prepare(s, unit_segments(dim, f64, nunits), st);
    if (!rc) rc = prepare_offsets(s, nunits, st);
    if (rc) return rc;
    const ScanCtx sc = scan_ctx(s);
    Item item = make_item(cmp, out, cmp_size, nullptr, (unsigned int*)(s.d_misc + 2), decompress_param(eb, type),
                          dim, type, mode, g);
    hipError_t e = f64 ? launch_unit_offsets_f64(dim, mode, item, nunits, s.d_offs, sc, st)
                       : launch_unit_offsets_f32(dim, mode, item, nunits, s.d_offs, sc, st);
    if (e != hipSuccess) return CUSZP_ERROR_HIP;
    e = f64 ? launch_decompress_f64(dim, mode, plan.waves, item, sc, s.d_offs, st)
            : launch_decompress_f32(dim, mode, plan.waves, item, sc, s.d_offs, st);
    if (e != hipSuccess) return CUSZP_ERROR_HIP;
    if (!blocking) return CUSZP_SUCCESS;
    return finish_blocking(s, st, nullptr);
}

void report(const char* fn, int rc) {
    if (rc != CUSZP_SUCCESS) std::fprintf(stderr, "cuSZp: %s: %s\n", fn, cuSZp_status_string(rc));
}

const cuszp_dims_t kNoDims = {0, 0, 0};

}  // namespace
}  // namespace cuszp

using namespace cuszp;

extern "C" {

const char* cuSZp_status_string(int status) {
    switch (status) {
        case CUSZP_SUCCESS: return "success";
        case CUSZP_ERROR_INVALID_ARGUMENT: return "invalid argument (enum, pointer or error bound)";
        case CUSZP_ERROR_INVALID_DIMS: return "invalid dims (need x % 4 == 0, non-zero extents, nbEle == x*y*z)";
        case CUSZP_ERROR_TOO_LARGE: return "input too large for the stream format";
        case CUSZP_ERROR_OUTPUT_TOO_SMALL: return "compressed output exceeds the buffer capacity";
        case CUSZP_ERROR_CORRUPT_STREAM: return "compressed stream is shorter than its header describes";
        case CUSZP_ERROR_HIP: return "HIP runtime error";
        case CUSZP_ERROR_INTERNAL: return "internal error: the device-wide scan stalled (output invalid)";
        case CUSZP_WARNING_BOUND_EXCEEDED:
            return "warning: a 2D/3D outlier-mode block holds a residual of magnitude 2^31 (saturated codes); "
                   "it decodes to zeros, beyond the error bound";
        default: return "unknown status";
    }
}

const char* cuSZp_version(void) { return "cuszp_amd 0.1.0 gfx950"; }

size_t cuSZp_max_compressed_size(size_t nbEle, cuszp_dim_t dim, cuszp_dims_t dims, cuszp_type_t type,
                                 cuszp_mode_t mode) {
    if (bad_enums(dim, type, mode)) return 0;
    Geom g;
    if (make_geom(dim, nbEle, dims, g)) return 0;
    if (nbEle == 0) return 0;
    return (size_t)max_stream_bytes(dim, g);
}

int cuSZp_compress_ex(const void* d_oriData, unsigned char* d_cmpBytes, size_t cmpCapacity, size_t nbEle,
                      size_t* cmpSize, double errorBound, cuszp_dim_t dim, cuszp_dims_t dims,
                      cuszp_type_t type, cuszp_mode_t mode, hipStream_t stream) {
    return compress_once(d_oriData, d_cmpBytes, cmpCapacity, nbEle, cmpSize, nullptr, errorBound, dim, dims,
                         type, mode, stream, true);
}

int cuSZp_decompress_ex(void* d_decData, const unsigned char* d_cmpBytes, size_t nbEle, size_t cmpSize,
                        double errorBound, cuszp_dim_t dim, cuszp_dims_t dims, cuszp_type_t type,
                        cuszp_mode_t mode, hipStream_t stream) {
    return decompress_once(d_decData, d_cmpBytes, nbEle, cmpSize, errorBound, dim, dims, type, mode, stream,
                           true);
}

int cuSZp_compress_async(const void* d_oriData, unsigned char* d_cmpBytes, size_t cmpCapacity, size_t nbEle,
                         unsigned long long* d_cmpSize, double errorBound, cuszp_dim_t dim, cuszp_dims_t dims,
                         cuszp_type_t type, cuszp_mode_t mode, hipStream_t stream) {
    return compress_once(d_oriData, d_cmpBytes, cmpCapacity, nbEle, nullptr, d_cmpSize, errorBound, dim, dims,
                         type, mode, stream, false);
}

int cuSZp_decompress_async(void* d_decData, const unsigned char* d_cmpBytes, size_t nbEle, size_t cmpSize,
                           double errorBound, cuszp_dim_t dim, cuszp_dims_t dims, cuszp_type_t type,
                           cuszp_mode_t mode, hipStream_t stream) {
    return decompress_once(d_decData, d_cmpBytes, nbEle, cmpSize, errorBound, dim, dims, type, mode, stream,
                           false);
}

int cuSZp_last_async_status(hipStream_t stream) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return CUSZP_ERROR_HIP;
    Slot& s = slot_for(dev, stream);
    std::lock_guard<std::mutex> lk(s.mu);
    if (!s.d_misc) return CUSZP_SUCCESS;
    return finish_blocking(s, stream, nullptr);
}

int cuSZp_scan_report(hipStream_t stream, cuszp_scan_report_t* report) {
    if (!report) return CUSZP_ERROR_INVALID_ARGUMENT;
    std::memset(report, 0, sizeof(*report));
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return CUSZP_ERROR_HIP;
    Slot& s = slot_for(dev, stream);
    std::lock_guard<std::mutex> lk(s.mu);
    if (!s.d_misc) return CUSZP_SUCCESS;
    CUSZP_HIP(hipMemcpyAsync(s.h_res, s.d_misc, kMiscWords * sizeof(unsigned long long), hipMemcpyDeviceToHost, stream));
    CUSZP_HIP(hipStreamSynchronize(stream));
    report->info = s.info | (unsigned)(s.h_res[kMiscErr] & kInfoStolen);
    report->polls = s.h_res[kMiscCounters + 0];
    report->waits = s.h_res[kMiscCounters + 1];
    report->claims = s.h_res[kMiscCounters + 2];
    report->wait_ticks = s.h_res[kMiscCounters + 3];
    s.info = 0;
    CUSZP_HIP(hipMemsetAsync(s.d_misc + kMiscCounters, 0, 4 * sizeof(unsigned long long), stream));
    CUSZP_HIP(hipStreamSynchronize(stream));
    return CUSZP_SUCCESS;
}

int cuSZp_debug_fail_allocs(unsigned int count) {
    failing_allocs().store(count, std::memory_order_relaxed);
    return CUSZP_SUCCESS;
}

int cuSZp_scan_config(int steal_ticks, unsigned int hold_every, unsigned int hold_ticks, int stats) {
    // holds are bounded (100 ms) so that a stressed launch still ends promptly
    if (hold_ticks > 10000000u) return CUSZP_ERROR_INVALID_ARGUMENT;
    ScanSettings& k = scan_settings();
    k.steal_ticks = steal_ticks < 0 ? (uint32_t)kStealTicks : (uint32_t)steal_ticks;
    k.hold_every = hold_ticks ? hold_every : 0;
    k.hold_ticks = hold_ticks;
    k.stats = stats ? 1u : 0u;
    return CUSZP_SUCCESS;
}

// ---- generic drop-ins: src/cuSZp.cu:4-83, :86-165 ------------------------
void cuSZp_compress(void* d_oriData, unsigned char* d_cmpBytes, size_t nbEle, size_t* cmpSize, float errorBound,
                    cuszp_dim_t dim, cuszp_dims_t dims, cuszp_type_t type, cuszp_mode_t mode, hipStream_t stream) {
    // f64 receives the float error bound widened to double (src/cuSZp.cu:45)
    report("cuSZp_compress", compress_once(d_oriData, d_cmpBytes, (size_t)-1, nbEle, cmpSize, nullptr,
                                           (double)errorBound, dim, dims, type, mode, stream, true));
}

void cuSZp_decompress(void* d_decData, unsigned char* d_cmpBytes, size_t nbEle, size_t cmpSize, float errorBound,
                      cuszp_dim_t dim, cuszp_dims_t dims, cuszp_type_t type, cuszp_mode_t mode,
                      hipStream_t stream) {
    report("cuSZp_decompress", decompress_once(d_decData, d_cmpBytes, nbEle, cmpSize, (double)errorBound, dim,
                                               dims, type, mode, stream, true));
}

// ---- per-dimension drop-ins: src/cuSZp_entry_{1,2,3}D_f{32,64}.cu --------
#define CUSZP_ENTRY_1D(MODE_NAME, MODE_ENUM, T, TNAME, TENUM)                                               \
    void cuSZp_compress_1D_##MODE_NAME##_##TNAME(T* d_oriData, unsigned char* d_cmpBytes, size_t nbEle,     \
                                                  size_t* cmpSize, T errorBound, hipStream_t stream) {      \
        report("cuSZp_compress_1D_" #MODE_NAME "_" #TNAME,                                                  \
               compress_once(d_oriData, d_cmpBytes, (size_t)-1, nbEle, cmpSize, nullptr, (double)errorBound, \
                             CUSZP_DIM_1D, kNoDims, TENUM, MODE_ENUM, stream, true));                       \
    }                                                                                                       \
    void cuSZp_decompress_1D_##MODE_NAME##_##TNAME(T* d_decData, unsigned char* d_cmpBytes, size_t nbEle,   \
                                                    size_t cmpSize, T errorBound, hipStream_t stream) {     \
        report("cuSZp_decompress_1D_" #MODE_NAME "_" #TNAME,                                                \
               decompress_once(d_decData, d_cmpBytes, nbEle, cmpSize, (double)errorBound, CUSZP_DIM_1D,      \
                               kNoDims, TENUM, MODE_ENUM, stream, true));                                   \
    }

#define CUSZP_ENTRY_ND(DIMN, DIM_ENUM, MODE_NAME, MODE_ENUM, T, TNAME, TENUM)                               \
    void cuSZp_compress_##DIMN##_##MODE_NAME##_##TNAME(T* d_oriData, unsigned char* d_cmpBytes,             \
                                                        size_t nbEle, size_t* cmpSize, cuszp_dims_t dims,   \
                                                        T errorBound, hipStream_t stream) {                 \
        report("cuSZp_compress_" #DIMN "_" #MODE_NAME "_" #TNAME,                                           \
               compress_once(d_oriData, d_cmpBytes, (size_t)-1, nbEle, cmpSize, nullptr, (double)errorBound, \
                             DIM_ENUM, dims, TENUM, MODE_ENUM, stream, true));                              \
    }                                                                                                       \
    void cuSZp_decompress_##DIMN##_##MODE_NAME##_##TNAME(T* d_decData, unsigned char* d_cmpBytes,           \
                                                          size_t nbEle, size_t cmpSize, cuszp_dims_t dims,  \
                                                          T errorBound, hipStream_t stream) {               \
        report("cuSZp_decompress_" #DIMN "_" #MODE_NAME "_" #TNAME,                                         \
               decompress_once(d_decData, d_cmpBytes, nbEle, cmpSize, (double)errorBound, DIM_ENUM, dims,    \
                               TENUM, MODE_ENUM, stream, true));                                            \
    }

#define CUSZP_ENTRIES_T(T, TNAME, TENUM)                                                   \
    CUSZP_ENTRY_1D(fixed, CUSZP_MODE_FIXED, T, TNAME, TENUM)                              \
    CUSZP_ENTRY_1D(plain, CUSZP_MODE_PLAIN, T, TNAME, TENUM)                              \
    CUSZP_ENTRY_1D(outlier, CUSZP_MODE_OUTLIER, T, TNAME, TENUM)                          \
    CUSZP_ENTRY_ND(2D, CUSZP_DIM_2D, fixed, CUSZP_MODE_FIXED, T, TNAME, TENUM)            \
    CUSZP_ENTRY_ND(2D, CUSZP_DIM_2D, plain, CUSZP_MODE_PLAIN, T, TNAME, TENUM)            \
    CUSZP_ENTRY_ND(2D, CUSZP_DIM_2D, outlier, CUSZP_MODE_OUTLIER, T, TNAME, TENUM)        \
    CUSZP_ENTRY_ND(3D, CUSZP_DIM_3D, fixed, CUSZP_MODE_FIXED, T, TNAME, TENUM)            \
    CUSZP_ENTRY_ND(3D, CUSZP_DIM_3D, plain, CUSZP_MODE_PLAIN, T, TNAME, TENUM)            \
    CUSZP_ENTRY_ND(3D, CUSZP_DIM_3D, outlier, CUSZP_MODE_OUTLIER, T, TNAME, TENUM)

CUSZP_ENTRIES_T(float, f32, CUSZP_TYPE_FLOAT)
CUSZP_ENTRIES_T(double, f64, CUSZP_TYPE_DOUBLE)

}  // extern "C"
