// Batched C ABI: many arrays in one launch (compress or decompress) with one
// workgroup per tile of the concatenated arrays,
// optionally held as a captured hipGraph, with a launch-shape autotuner.
//
// A plan validates every array once (the same checks as cuSZp_compress_ex /
// cuSZp_decompress_ex), lays the arrays' tiles end to end, uploads one Item
// descriptor per non-empty array and owns the scan state: tile status words,
// an epoch word bumped on the device before each launch (so replays of a
// captured graph never see the previous run's status words as current), a
// launch-wide error word and one error word and one stream length per array.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "common.hpp"

using namespace cuszp;

struct cuszp_batch_plan_s {
    int dev = 0;
    int dir = CUSZP_BATCH_COMPRESS;
    int flags = 0;
    int waves = 4;
    int window = kLookbackWords;   // lookback words per lane (2 or 4)
    std::vector<cuszp_batch_item_t> src;
    std::vector<Item> items;      // non-empty arrays, in caller order
    uint64_t ntiles = 0;
    Item* d_items = nullptr;
    size_t cap_items = 0;
    unsigned long long* d_status = nullptr;
    uint64_t cap_tiles = 0;
    unsigned long long* d_misc = nullptr;   // kMiscWords: [0] epoch, [2] launch bits, [4..7] diagnostics, [8..11] counters
    unsigned int* d_err = nullptr;          // per array
    unsigned long long* d_sizes = nullptr;  // per array (compress)
    hipStream_t cap_stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    unsigned info = 0;                      // information bits harvested since the last scan report
};

namespace {

#define BATCH_HIP(call)                                \
    do {                                               \
        hipError_t e_ = (call);                        \
        if (e_ != hipSuccess) return CUSZP_ERROR_HIP;  \
    } while (0)

int check_item(const cuszp_batch_item_t& it, int dir, Geom& g) {
    const int dim = it.dim, type = it.type, mode = it.mode;
    if (bad_enums(dim, type, mode) || bad_eb(it.errorBound, type)) return CUSZP_ERROR_INVALID_ARGUMENT;
    int rc = make_geom(dim, it.nbEle, it.dims, g);
    if (rc) return rc;
    if (it.nbEle == 0) return CUSZP_SUCCESS;
    if (!it.d_in || !it.d_out) return CUSZP_ERROR_INVALID_ARGUMENT;
    if (max_stream_bytes(dim, g) > kMaxStream) return CUSZP_ERROR_TOO_LARGE;
    if (it.capacity < g.H) return dir == CUSZP_BATCH_COMPRESS ? CUSZP_ERROR_OUTPUT_TOO_SMALL : CUSZP_ERROR_CORRUPT_STREAM;
    return CUSZP_SUCCESS;
}

// Lay out the tiles of every non-empty array for `waves` waves per workgroup
// and upload the descriptors (grows device buffers as needed).
int build_table(cuszp_batch_plan_s& p, hipStream_t st) {
    p.items.clear();
    uint64_t tile0 = 0;
    for (size_t i = 0; i < p.src.size(); ++i) {
        const cuszp_batch_item_t& s = p.src[i];
        if (s.nbEle == 0) continue;
        Geom g;
        make_geom(s.dim, s.nbEle, s.dims, g);
        const uint64_t groups = (s.dim == CUSZP_DIM_1D) ? g.H / 2 : g.nblocks;
        const uint64_t per_tile = tile_groups(s.type == CUSZP_TYPE_DOUBLE, p.waves);
        g.ntiles = (groups + per_tile - 1) / per_tile;
        const bool c = p.dir == CUSZP_BATCH_COMPRESS;
        Item it = make_item(s.d_in, s.d_out, s.capacity, c ? p.d_sizes + i : nullptr, p.d_err + i,
                            c ? compress_param(s.errorBound, s.type) : decompress_param(s.errorBound, s.type),
                            s.dim, s.type, s.mode, g);
        it.tile0 = tile0;
        tile0 += g.ntiles;
        p.items.push_back(it);
    }
    p.ntiles = tile0;
    // Growing a buffer: the old one is released before the new one is
    // allocated, and the capacity drops to 0 in between, so a failed
    // allocation leaves no buffer that a later call would take as large enough
    // (every launch of the plan has finished: the table is built only between
    // synchronised calls).
    if (p.items.size() > p.cap_items) {
        if (p.d_items) BATCH_HIP(hipFree(p.d_items));
        p.d_items = nullptr;
        p.cap_items = 0;
        BATCH_HIP(hipMalloc((void**)&p.d_items, p.items.size() * sizeof(Item)));
        p.cap_items = p.items.size();
    }
    const uint64_t need = p.ntiles ? p.ntiles : 1;   // an empty table still gets a (cleared) status array
    if (need > p.cap_tiles) {
        if (p.d_status) BATCH_HIP(hipFree(p.d_status));
        p.d_status = nullptr;
        p.cap_tiles = 0;
        // kStatusGuard words in front (lookback windows of first tiles read there)
        BATCH_HIP(scan_state_alloc(&p.d_status, (kStatusGuard + need) * sizeof(unsigned long long)));
        p.cap_tiles = need;
    }
    // stale status words must not carry the current epoch: start from zero and
    // restart the epoch whenever the table changes
    BATCH_HIP(hipMemsetAsync(p.d_status, 0, (kStatusGuard + p.cap_tiles) * sizeof(unsigned long long), st));
    BATCH_HIP(hipMemsetAsync(p.d_misc, 0, sizeof(unsigned long long), st));
    if (!p.items.empty())
        BATCH_HIP(hipMemcpyAsync(p.d_items, p.items.data(), p.items.size() * sizeof(Item), hipMemcpyHostToDevice, st));
    BATCH_HIP(hipStreamSynchronize(st));   // the host table may be rebuilt right after
    return CUSZP_SUCCESS;
}

int enqueue(const cuszp_batch_plan_s& p, hipStream_t st) {
    if (p.items.empty()) return CUSZP_SUCCESS;
    // the kernels read the epoch word for the tag
    const ScanCtx sc = make_scan_ctx(p.d_status + kStatusGuard, (unsigned int*)(p.d_misc + kMiscErr),
                                     p.d_misc + kMiscDiag, 0);
    BATCH_HIP(launch_bump_epoch(p.d_misc, st));
    const uint32_t n = (uint32_t)p.items.size();
    if (p.dir == CUSZP_BATCH_COMPRESS)
        BATCH_HIP(launch_compress_batch(p.waves, p.window, p.d_items, n, p.ntiles, sc, p.d_misc, st));
    else
        BATCH_HIP(launch_decompress_batch(p.waves, p.window, p.d_items, n, p.ntiles, sc, p.d_misc, st));
    return CUSZP_SUCCESS;
}

void drop_graph(cuszp_batch_plan_s& p) {
    if (p.exec) hipGraphExecDestroy(p.exec);
    if (p.graph) hipGraphDestroy(p.graph);
    p.exec = nullptr;
    p.graph = nullptr;
}

int capture(cuszp_batch_plan_s& p) {
    drop_graph(p);
    if (!p.cap_stream) BATCH_HIP(hipStreamCreateWithFlags(&p.cap_stream, hipStreamNonBlocking));
    BATCH_HIP(hipStreamBeginCapture(p.cap_stream, hipStreamCaptureModeThreadLocal));
    const int rc = enqueue(p, p.cap_stream);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(p.cap_stream, &g);
    if (rc) {
        if (g) hipGraphDestroy(g);
        return rc;
    }
    BATCH_HIP(e);
    p.graph = g;
    BATCH_HIP(hipGraphInstantiate(&p.exec, p.graph, nullptr, nullptr, 0));
    return CUSZP_SUCCESS;
}

// Launch shapes chosen by the tuner, by direction and size class of the batch
// (log2 of its tile count and of its array count); later plans of the same
// class start from the recorded shape.
std::mutex g_tune_mu;
std::map<std::tuple<int, int, int>, std::pair<int, int>> g_tuned;   // -> (waves per workgroup, lookback window)

std::tuple<int, int, int> size_class(const cuszp_batch_plan_s& p) {
    uint64_t bytes = 0;
    for (const auto& s : p.src) bytes += s.nbEle * (s.type == CUSZP_TYPE_FLOAT ? 4 : 8);
    int lb = 0, lc = 0;
    while ((bytes >> lb) > 1) ++lb;
    while ((p.src.size() >> lc) > 1) ++lc;
    return std::make_tuple(p.dir, lb, lc);
}

}  // namespace

extern "C" {

int cuSZp_batch_plan_create(cuszp_batch_plan_t* out, const cuszp_batch_item_t* items, size_t count, int direction,
                            int flags, hipStream_t stream, size_t* bad_index) {
    if (!out) return CUSZP_ERROR_INVALID_ARGUMENT;
    *out = nullptr;
    if ((direction != CUSZP_BATCH_COMPRESS && direction != CUSZP_BATCH_DECOMPRESS) || (count && !items) ||
        count > 0xffffffffull)
        return CUSZP_ERROR_INVALID_ARGUMENT;
    for (size_t i = 0; i < count; ++i) {
        Geom g;
        const int rc = check_item(items[i], direction, g);
        if (rc) {
            if (bad_index) *bad_index = i;
            return rc;
        }
    }
    cuszp_batch_plan_s* p = new cuszp_batch_plan_s();
    p->dir = direction;
    p->flags = flags;
    p->src.assign(items, items + count);
    int rc = CUSZP_SUCCESS;
    if (hipGetDevice(&p->dev) != hipSuccess) rc = CUSZP_ERROR_HIP;
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        auto f = g_tuned.find(size_class(*p));
        if (f != g_tuned.end()) {
            p->waves = f->second.first;
            p->window = f->second.second;
        }
    }
    if (!rc && hipMalloc((void**)&p->d_misc, kMiscWords * sizeof(unsigned long long)) != hipSuccess) rc = CUSZP_ERROR_HIP;
    const size_t n = count ? count : 1;
    if (!rc && hipMalloc((void**)&p->d_err, n * sizeof(unsigned int)) != hipSuccess) rc = CUSZP_ERROR_HIP;
    if (!rc && hipMalloc((void**)&p->d_sizes, n * sizeof(unsigned long long)) != hipSuccess) rc = CUSZP_ERROR_HIP;
    if (!rc && hipMemsetAsync(p->d_misc, 0, kMiscWords * sizeof(unsigned long long), stream) != hipSuccess)
        rc = CUSZP_ERROR_HIP;
    if (!rc && hipMemsetAsync(p->d_err, 0, n * sizeof(unsigned int), stream) != hipSuccess) rc = CUSZP_ERROR_HIP;
    if (!rc && hipMemsetAsync(p->d_sizes, 0, n * sizeof(unsigned long long), stream) != hipSuccess) rc = CUSZP_ERROR_HIP;
    if (!rc) rc = build_table(*p, stream);
    if (!rc && (flags & CUSZP_BATCH_TUNE)) rc = cuSZp_batch_plan_tune(p, stream, nullptr, nullptr);
    if (!rc && (flags & CUSZP_BATCH_GRAPH) && !p->exec) rc = capture(*p);
    if (rc) {
        cuSZp_batch_plan_destroy(p);
        return rc;
    }
    *out = p;
    return CUSZP_SUCCESS;
}

int cuSZp_batch_plan_execute(cuszp_batch_plan_t p, hipStream_t stream) {
    if (!p) return CUSZP_ERROR_INVALID_ARGUMENT;
    if (p->exec) {
        BATCH_HIP(hipGraphLaunch(p->exec, stream));
        return CUSZP_SUCCESS;
    }
    return enqueue(*p, stream);
}

int cuSZp_batch_plan_results(cuszp_batch_plan_t p, size_t* cmpSizes, int* statuses, hipStream_t stream) {
    if (!p) return CUSZP_ERROR_INVALID_ARGUMENT;
    const size_t n = p->src.size();
    std::vector<unsigned long long> sizes(n ? n : 1);
    std::vector<unsigned int> errs(n ? n : 1);
    unsigned long long misc[kMiscWords];
    BATCH_HIP(hipMemcpyAsync(misc, p->d_misc, sizeof(misc), hipMemcpyDeviceToHost, stream));
    BATCH_HIP(hipMemcpyAsync(errs.data(), p->d_err, errs.size() * sizeof(unsigned int), hipMemcpyDeviceToHost, stream));
    BATCH_HIP(hipMemcpyAsync(sizes.data(), p->d_sizes, sizes.size() * sizeof(unsigned long long),
                             hipMemcpyDeviceToHost, stream));
    BATCH_HIP(hipStreamSynchronize(stream));
    const unsigned long long launch_bits = misc[kMiscErr] & 0xffffffffull;
    p->info |= (unsigned)(launch_bits & kInfoStolen);
    if (launch_bits & kErrStall)
        std::fprintf(stderr, "cuSZp: internal error: batched tile %llu stopped waiting for tile %llu\n", misc[4] - 1,
                     misc[5]);
    int first = CUSZP_SUCCESS;
    for (size_t i = 0; i < n; ++i) {
        const int st = status_from_bits(errs[i] | (launch_bits & kErrStall));
        if (statuses) statuses[i] = st;
        if (cmpSizes) cmpSizes[i] = (p->dir == CUSZP_BATCH_COMPRESS && status_delivers(st)) ? (size_t)sizes[i] : 0;
        if (st && !first) first = st;
    }
    if (launch_bits) {
        BATCH_HIP(hipMemsetAsync(p->d_misc + kMiscErr, 0, (kMiscCounters - kMiscErr) * sizeof(unsigned long long),
                                 stream));
    }
    if (first) BATCH_HIP(hipMemsetAsync(p->d_err, 0, (n ? n : 1) * sizeof(unsigned int), stream));
    BATCH_HIP(hipStreamSynchronize(stream));
    return first;
}

const unsigned long long* cuSZp_batch_plan_device_sizes(cuszp_batch_plan_t p) { return p ? p->d_sizes : nullptr; }

int cuSZp_batch_plan_scan_report(cuszp_batch_plan_t p, hipStream_t stream, cuszp_scan_report_t* report) {
    if (!p || !report) return CUSZP_ERROR_INVALID_ARGUMENT;
    std::memset(report, 0, sizeof(*report));
    unsigned long long misc[kMiscWords];
    BATCH_HIP(hipMemcpyAsync(misc, p->d_misc, sizeof(misc), hipMemcpyDeviceToHost, stream));
    BATCH_HIP(hipStreamSynchronize(stream));
    report->info = p->info | (unsigned)(misc[kMiscErr] & kInfoStolen);
    report->polls = misc[kMiscCounters + 0];
    report->waits = misc[kMiscCounters + 1];
    report->claims = misc[kMiscCounters + 2];
    report->wait_ticks = misc[kMiscCounters + 3];
    p->info = 0;
    BATCH_HIP(hipMemsetAsync(p->d_misc + kMiscCounters, 0, 4 * sizeof(unsigned long long), stream));
    BATCH_HIP(hipStreamSynchronize(stream));
    return CUSZP_SUCCESS;
}

int cuSZp_batch_plan_tune(cuszp_batch_plan_t p, hipStream_t stream, int* waves_out, unsigned* grid_out) {
    if (!p) return CUSZP_ERROR_INVALID_ARGUMENT;
    if (p->items.empty()) return CUSZP_SUCCESS;
    const int kShapes[4][2] = {{4, 4}, {8, 4}, {4, 2}, {8, 2}};   // (waves, lookback window)
    hipEvent_t e0, e1;
    BATCH_HIP(hipEventCreate(&e0));
    BATCH_HIP(hipEventCreate(&e1));
    float best = 1e30f;
    int best_w = p->waves, best_k = p->window;
    int rc = CUSZP_SUCCESS;
    for (const auto& shape : kShapes) {
        p->waves = shape[0];
        p->window = shape[1];
        rc = build_table(*p, stream);
        if (rc) break;
        rc = enqueue(*p, stream);   // warm-up
        if (rc) break;
        hipEventRecord(e0, stream);
        for (int r = 0; r < 3 && !rc; ++r) rc = enqueue(*p, stream);
        hipEventRecord(e1, stream);
        if (rc || hipEventSynchronize(e1) != hipSuccess) {
            rc = rc ? rc : CUSZP_ERROR_HIP;
            break;
        }
        float ms = 0;
        hipEventElapsedTime(&ms, e0, e1);
        if (ms < best) {
            best = ms;
            best_w = shape[0];
            best_k = shape[1];
        }
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    if (rc) return rc;
    p->waves = best_w;
    p->window = best_k;
    rc = build_table(*p, stream);
    if (rc) return rc;
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        g_tuned[size_class(*p)] = {best_w, best_k};
    }
    if (p->exec || (p->flags & CUSZP_BATCH_GRAPH)) rc = capture(*p);
    // tuning runs leave per-array state from the candidate launches: clear it
    const size_t n = p->src.size() ? p->src.size() : 1;
    if (!rc && hipMemsetAsync(p->d_err, 0, n * sizeof(unsigned int), stream) != hipSuccess) rc = CUSZP_ERROR_HIP;
    if (!rc && hipMemsetAsync(p->d_misc + kMiscErr, 0, (kMiscWords - kMiscErr) * sizeof(unsigned long long),
                              stream) != hipSuccess)
        rc = CUSZP_ERROR_HIP;
    if (waves_out) *waves_out = best_w;
    if (grid_out) *grid_out = (unsigned)(p->ntiles < kMaxGrid ? p->ntiles : kMaxGrid);
    return rc;
}

void cuSZp_batch_plan_destroy(cuszp_batch_plan_t p) {
    if (!p) return;
    drop_graph(*p);
    if (p->cap_stream) {
        hipStreamSynchronize(p->cap_stream);
        hipStreamDestroy(p->cap_stream);
    }
    hipDeviceSynchronize();   // no launch of this plan may still use its buffers
    if (p->d_items) hipFree(p->d_items);
    if (p->d_status) hipFree(p->d_status);
    if (p->d_misc) hipFree(p->d_misc);
    if (p->d_err) hipFree(p->d_err);
    if (p->d_sizes) hipFree(p->d_sizes);
    delete p;
}

int cuSZp_compress_batch(const cuszp_batch_item_t* items, size_t count, size_t* cmpSizes, int* statuses,
                         hipStream_t stream) {
    cuszp_batch_plan_t p = nullptr;
    int rc = cuSZp_batch_plan_create(&p, items, count, CUSZP_BATCH_COMPRESS, 0, stream, nullptr);
    if (rc) return rc;
    rc = cuSZp_batch_plan_execute(p, stream);
    if (!rc) rc = cuSZp_batch_plan_results(p, cmpSizes, statuses, stream);
    cuSZp_batch_plan_destroy(p);
    return rc;
}

int cuSZp_decompress_batch(const cuszp_batch_item_t* items, size_t count, int* statuses, hipStream_t stream) {
    cuszp_batch_plan_t p = nullptr;
    int rc = cuSZp_batch_plan_create(&p, items, count, CUSZP_BATCH_DECOMPRESS, 0, stream, nullptr);
    if (rc) return rc;
    rc = cuSZp_batch_plan_execute(p, stream);
    if (!rc) rc = cuSZp_batch_plan_results(p, nullptr, statuses, stream);
    cuSZp_batch_plan_destroy(p);
    return rc;
}

}  // extern "C"
