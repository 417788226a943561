// Serving form: a request server with continuous batching on top of the
// batched kernels (no counterpart in the reference, whose API is one blocking
// call per array).
//
// Requests (compress or decompress of one array) are queued in arrival order.
// Each one needs its output in the server's pool: a device buffer the caller
// hands over at creation, sub-allocated first-fit in 256-byte units. The head
// of the queue is admitted when its output fits in the pool (strict FIFO, so a
// large request is never starved by smaller ones behind it). Admitted requests
// wait for a free lane; a lane is a stream with its own item table and scan
// state, running one batch at a time: the batch's compress requests in one
// compress_batch_kernel launch and its decompress requests in one
// decompress_batch_kernel launch. Several lanes run at once, so requests that
// arrive while a batch is on the GPU start on another lane without waiting
// for it. Each request's result is its single-array result: the batched
// kernels give every array its own lookback segment.
//
// Completed requests keep their output in the pool until the caller releases
// them.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <unordered_map>
#include <vector>

#include "common.hpp"

using namespace cuszp;

namespace {

constexpr uint64_t kPoolAlign = 256;
constexpr int kServeWaves = 4;
constexpr uint64_t kEpochClearEvery = 1ull << 23;   // launches between status clears (24-bit epoch)

struct Req {
    unsigned long long id = 0;
    int dir = CUSZP_BATCH_COMPRESS;
    cuszp_batch_item_t item{};
    Geom geo{};
    uint64_t out_bytes = 0;   // pool bytes of the output
    uint64_t off = 0;         // pool offset once admitted
};

struct Lane {
    hipStream_t st = nullptr;
    hipEvent_t ev = nullptr;
    bool busy = false;
    Item* d_items = nullptr;
    Item* h_items = nullptr;   // pinned staging of the table
    unsigned long long* d_status = nullptr;
    uint64_t cap_tiles = 0;
    uint64_t launches = 0;
    unsigned long long* d_misc = nullptr;   // kMiscWords: [0] epoch, [2] launch bits, [4..7] diagnostics
    unsigned int* d_err = nullptr;
    unsigned long long* d_sizes = nullptr;
    unsigned int* h_err = nullptr;
    unsigned long long* h_sizes = nullptr;
    unsigned long long* h_misc = nullptr;
    std::vector<Req> reqs;     // compress requests first, then decompress
    size_t ncomp = 0;
};

// First-fit allocator over [0, size) with coalescing frees.
class PoolAlloc {
public:
    void reset(uint64_t size) {
        free_.clear();
        if (size) free_[0] = size;
        used_ = 0;
    }
    bool alloc(uint64_t bytes, uint64_t& off) {
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < bytes) continue;
            off = it->first;
            const uint64_t rest = it->second - bytes;
            free_.erase(it);
            if (rest) free_[off + bytes] = rest;
            used_ += bytes;
            return true;
        }
        return false;
    }
    void release(uint64_t off, uint64_t bytes) {
        used_ -= bytes;
        auto it = free_.emplace(off, bytes).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
    uint64_t used() const { return used_; }

private:
    std::map<uint64_t, uint64_t> free_;
    uint64_t used_ = 0;
};

uint64_t round_up(uint64_t v, uint64_t a) { return (v + a - 1) / a * a; }

}  // namespace

struct cuszp_server_s {
    int dev = 0;
    uint8_t* pool = nullptr;
    uint64_t pool_bytes = 0;
    size_t max_batch = 0;
    PoolAlloc alloc;
    std::vector<Lane> lanes;
    std::deque<Req> queue;      // submitted, not admitted
    std::deque<Req> admitted;   // output allocated, waiting for a lane
    std::deque<cuszp_completion_t> done;
    struct Hold {
        uint64_t off, bytes;
        bool complete;
    };
    std::unordered_map<unsigned long long, Hold> held;   // every request not yet released
    unsigned long long next_id = 1;
    std::mutex mu;
};

namespace {

#define SRV_HIP(call)                                  \
    do {                                               \
        hipError_t e_ = (call);                        \
        if (e_ != hipSuccess) return CUSZP_ERROR_HIP;  \
    } while (0)

int lane_init(Lane& L, size_t max_batch, uint64_t tiles) {
    SRV_HIP(hipStreamCreateWithFlags(&L.st, hipStreamNonBlocking));
    SRV_HIP(hipEventCreateWithFlags(&L.ev, hipEventDisableTiming));
    SRV_HIP(hipMalloc((void**)&L.d_items, max_batch * sizeof(Item)));
    SRV_HIP(hipHostMalloc((void**)&L.h_items, max_batch * sizeof(Item), hipHostMallocDefault));
    SRV_HIP(hipMalloc((void**)&L.d_misc, kMiscWords * sizeof(unsigned long long)));
    SRV_HIP(hipMalloc((void**)&L.d_err, max_batch * sizeof(unsigned int)));
    SRV_HIP(hipMalloc((void**)&L.d_sizes, max_batch * sizeof(unsigned long long)));
    SRV_HIP(hipHostMalloc((void**)&L.h_err, max_batch * sizeof(unsigned int), hipHostMallocDefault));
    SRV_HIP(hipHostMalloc((void**)&L.h_sizes, max_batch * sizeof(unsigned long long), hipHostMallocDefault));
    SRV_HIP(hipHostMalloc((void**)&L.h_misc, kMiscWords * sizeof(unsigned long long), hipHostMallocDefault));
    SRV_HIP(hipMalloc((void**)&L.d_status, (kStatusGuard + tiles) * sizeof(unsigned long long)));
    L.cap_tiles = tiles;
    SRV_HIP(hipMemsetAsync(L.d_status, 0, (kStatusGuard + tiles) * sizeof(unsigned long long), L.st));
    SRV_HIP(hipMemsetAsync(L.d_misc, 0, kMiscWords * sizeof(unsigned long long), L.st));
    SRV_HIP(hipStreamSynchronize(L.st));
    return CUSZP_SUCCESS;
}

void lane_free(Lane& L) {
    if (L.st) hipStreamSynchronize(L.st);
    if (L.d_items) hipFree(L.d_items);
    if (L.h_items) hipHostFree(L.h_items);
    if (L.d_status) hipFree(L.d_status);
    if (L.d_misc) hipFree(L.d_misc);
    if (L.d_err) hipFree(L.d_err);
    if (L.d_sizes) hipFree(L.d_sizes);
    if (L.h_err) hipHostFree(L.h_err);
    if (L.h_sizes) hipHostFree(L.h_sizes);
    if (L.h_misc) hipHostFree(L.h_misc);
    if (L.ev) hipEventDestroy(L.ev);
    if (L.st) hipStreamDestroy(L.st);
    L = Lane();
}

uint64_t tiles_of(const Req& r) {
    const uint64_t groups = (r.item.dim == CUSZP_DIM_1D) ? r.geo.H / 2 : r.geo.nblocks;
    const uint64_t per = tile_groups(r.item.type == CUSZP_TYPE_DOUBLE, kServeWaves);
    return (groups + per - 1) / per;
}

// Enqueues lane L's batch (L.reqs, compress first) on its stream. The lane is idle.
int lane_launch(cuszp_server_s& s, Lane& L) {
    const size_t n = L.reqs.size();
    uint64_t tc = 0, td = 0;
    for (size_t k = 0; k < n; ++k) {
        Req& r = L.reqs[k];
        const bool c = r.dir == CUSZP_BATCH_COMPRESS;
        Geom g = r.geo;
        g.ntiles = tiles_of(r);
        const cuszp_batch_item_t& it = r.item;
        Item item = make_item(it.d_in, s.pool + r.off, c ? r.out_bytes : it.capacity, c ? L.d_sizes + k : nullptr,
                              L.d_err + k, c ? compress_param(it.errorBound, it.type)
                                             : decompress_param(it.errorBound, it.type),
                              it.dim, it.type, it.mode, g);
        uint64_t& t = c ? tc : td;
        item.tile0 = t;
        t += g.ntiles;
        L.h_items[k] = item;
    }
    const uint64_t need = tc > td ? tc : td;
    if (need > L.cap_tiles) {   // the lane is idle: its status array can be replaced
        uint64_t cap = L.cap_tiles * 2 > need ? L.cap_tiles * 2 : need;
        if (L.d_status) SRV_HIP(hipFree(L.d_status));
        L.d_status = nullptr;
        L.cap_tiles = 0;   // a failed allocation below leaves no array to launch on
        SRV_HIP(scan_state_alloc(&L.d_status, (kStatusGuard + cap) * sizeof(unsigned long long)));
        L.cap_tiles = cap;
        SRV_HIP(hipMemsetAsync(L.d_status, 0, (kStatusGuard + cap) * sizeof(unsigned long long), L.st));
        L.launches = 0;
    }
    if (L.launches >= kEpochClearEvery) {   // before the 24-bit epoch tag can repeat
        SRV_HIP(hipMemsetAsync(L.d_status, 0, (kStatusGuard + L.cap_tiles) * sizeof(unsigned long long), L.st));
        L.launches = 0;
    }
    SRV_HIP(hipMemcpyAsync(L.d_items, L.h_items, n * sizeof(Item), hipMemcpyHostToDevice, L.st));
    SRV_HIP(hipMemsetAsync(L.d_err, 0, n * sizeof(unsigned int), L.st));
    SRV_HIP(hipMemsetAsync(L.d_misc + kMiscErr, 0, (kMiscCounters - kMiscErr) * sizeof(unsigned long long), L.st));
    // the kernels read the epoch word for the tag
    const ScanCtx sc = make_scan_ctx(L.d_status + kStatusGuard, (unsigned int*)(L.d_misc + kMiscErr),
                                     L.d_misc + kMiscDiag, 0);
    const size_t nc = L.ncomp;
    if (nc) {
        SRV_HIP(launch_bump_epoch(L.d_misc, L.st));
        SRV_HIP(launch_compress_batch(kServeWaves, kLookbackWords, L.d_items, (uint32_t)nc, tc, sc, L.d_misc, L.st));
        ++L.launches;
    }
    if (n > nc) {
        SRV_HIP(launch_bump_epoch(L.d_misc, L.st));
        SRV_HIP(launch_decompress_batch(kServeWaves, kLookbackWords, L.d_items + nc, (uint32_t)(n - nc), td, sc, L.d_misc, L.st));
        ++L.launches;
    }
    SRV_HIP(hipMemcpyAsync(L.h_err, L.d_err, n * sizeof(unsigned int), hipMemcpyDeviceToHost, L.st));
    if (nc) SRV_HIP(hipMemcpyAsync(L.h_sizes, L.d_sizes, nc * sizeof(unsigned long long), hipMemcpyDeviceToHost, L.st));
    SRV_HIP(hipMemcpyAsync(L.h_misc, L.d_misc, kMiscWords * sizeof(unsigned long long), hipMemcpyDeviceToHost, L.st));
    SRV_HIP(hipEventRecord(L.ev, L.st));
    L.busy = true;
    return CUSZP_SUCCESS;
}

void lane_retire(cuszp_server_s& s, Lane& L) {
    const unsigned long long stall = L.h_misc[kMiscErr] & kErrStall;
    for (size_t k = 0; k < L.reqs.size(); ++k) {
        const Req& r = L.reqs[k];
        cuszp_completion_t c;
        std::memset(&c, 0, sizeof(c));
        c.id = r.id;
        c.direction = r.dir;
        c.status = status_from_bits(L.h_err[k] | stall);
        c.offset = r.off;
        c.d_out = s.pool + r.off;
        c.bytes = r.dir == CUSZP_BATCH_COMPRESS ? (size_t)L.h_sizes[k]
                                                : r.item.nbEle * (r.item.type == CUSZP_TYPE_FLOAT ? 4 : 8);
        if (!status_delivers(c.status)) c.bytes = 0;
        s.done.push_back(c);
        s.held[r.id].complete = true;
    }
    L.reqs.clear();
    L.ncomp = 0;
    L.busy = false;
}

// Completes every request of lane L with status `rc` without running it (its
// launch failed); the lane is left idle and empty.
void lane_fail(cuszp_server_s& s, Lane& L, int rc) {
    if (L.st) hipStreamSynchronize(L.st);   // whatever was enqueued before the failure has finished
    for (const Req& r : L.reqs) {
        cuszp_completion_t c;
        std::memset(&c, 0, sizeof(c));
        c.id = r.id;
        c.direction = r.dir;
        c.status = rc;
        c.offset = r.off;
        c.d_out = s.pool + r.off;
        s.done.push_back(c);
        s.held[r.id].complete = true;
    }
    L.reqs.clear();
    L.ncomp = 0;
    L.busy = false;
}

// Retires finished lanes, admits queued requests while their outputs fit and
// starts admitted ones on idle lanes. Caller holds s.mu.
int progress(cuszp_server_s& s) {
    for (Lane& L : s.lanes) {
        if (!L.busy) continue;
        const hipError_t q = hipEventQuery(L.ev);
        if (q == hipSuccess) lane_retire(s, L);
        else if (q != hipErrorNotReady) return CUSZP_ERROR_HIP;
    }
    while (!s.queue.empty()) {
        Req& r = s.queue.front();
        if (!s.alloc.alloc(round_up(r.out_bytes, kPoolAlign), r.off)) break;
        s.held[r.id] = {r.off, round_up(r.out_bytes, kPoolAlign), false};
        s.admitted.push_back(r);
        s.queue.pop_front();
    }
    for (Lane& L : s.lanes) {
        if (s.admitted.empty()) break;
        if (L.busy) continue;
        const size_t n = s.admitted.size() < s.max_batch ? s.admitted.size() : s.max_batch;
        std::vector<Req> dec;
        for (size_t k = 0; k < n; ++k) {
            Req r = s.admitted.front();
            s.admitted.pop_front();
            if (r.dir == CUSZP_BATCH_COMPRESS) L.reqs.push_back(r);
            else dec.push_back(r);
        }
        L.ncomp = L.reqs.size();
        L.reqs.insert(L.reqs.end(), dec.begin(), dec.end());
        const int rc = lane_launch(s, L);
        if (rc) {
            // the batch did not start: its requests complete with the error
            // (their outputs stay held until released) and the lane is idle
            // and empty again
            lane_fail(s, L, rc);
            return rc;
        }
    }
    return CUSZP_SUCCESS;
}

bool any_busy(const cuszp_server_s& s) {
    for (const Lane& L : s.lanes)
        if (L.busy) return true;
    return false;
}

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace

extern "C" {

int cuSZp_server_create(cuszp_server_t* out, void* d_pool, size_t pool_bytes, int lanes, size_t max_batch) {
    if (!out) return CUSZP_ERROR_INVALID_ARGUMENT;
    *out = nullptr;
    if (!d_pool || pool_bytes < kPoolAlign || lanes < 1 || lanes > 4 || max_batch < 1 || max_batch > (1u << 20))
        return CUSZP_ERROR_INVALID_ARGUMENT;
    cuszp_server_s* s = new cuszp_server_s();
    s->pool = (uint8_t*)d_pool;
    s->pool_bytes = pool_bytes / kPoolAlign * kPoolAlign;
    s->max_batch = max_batch;
    s->alloc.reset(s->pool_bytes);
    int rc = hipGetDevice(&s->dev) == hipSuccess ? CUSZP_SUCCESS : CUSZP_ERROR_HIP;
    // status words for a batch whose outputs fill the pool: one tile per 64 KiB
    // of f32 input (32 KiB of f64 output), plus one partial tile per request
    // (a 1D request's header covers whole 32768-element spans: up to 4 tiles when small)
    const uint64_t tiles = s->pool_bytes / 32768 + 4 * max_batch + 1;
    s->lanes.resize(lanes);
    for (int i = 0; i < lanes && !rc; ++i) rc = lane_init(s->lanes[i], max_batch, tiles);
    if (rc) {
        cuSZp_server_destroy(s);
        return rc;
    }
    *out = s;
    return CUSZP_SUCCESS;
}

int cuSZp_server_submit(cuszp_server_t s, int direction, const cuszp_batch_item_t* item, unsigned long long* id) {
    if (!s || !item || !id) return CUSZP_ERROR_INVALID_ARGUMENT;
    if (direction != CUSZP_BATCH_COMPRESS && direction != CUSZP_BATCH_DECOMPRESS) return CUSZP_ERROR_INVALID_ARGUMENT;
    Req r;
    r.dir = direction;
    r.item = *item;
    r.item.d_out = nullptr;
    const int dim = item->dim, type = item->type, mode = item->mode;
    if (bad_enums(dim, type, mode) || bad_eb(item->errorBound, type)) return CUSZP_ERROR_INVALID_ARGUMENT;
    int rc = make_geom(dim, item->nbEle, item->dims, r.geo);
    if (rc) return rc;
    if (item->nbEle && !item->d_in) return CUSZP_ERROR_INVALID_ARGUMENT;
    if (max_stream_bytes(dim, r.geo) > kMaxStream) return CUSZP_ERROR_TOO_LARGE;
    if (direction == CUSZP_BATCH_COMPRESS) {
        r.out_bytes = item->nbEle ? max_stream_bytes(dim, r.geo) : 0;
    } else {
        if (item->nbEle && item->capacity < r.geo.H) return CUSZP_ERROR_CORRUPT_STREAM;
        r.out_bytes = item->nbEle * (type == CUSZP_TYPE_FLOAT ? 4 : 8);
    }
    std::lock_guard<std::mutex> lk(s->mu);
    if (round_up(r.out_bytes, kPoolAlign) > s->pool_bytes) return CUSZP_ERROR_TOO_LARGE;
    r.id = s->next_id++;
    *id = r.id;
    if (item->nbEle == 0) {   // nothing to launch: complete at once
        s->held[r.id] = {0, 0, true};
        cuszp_completion_t c;
        std::memset(&c, 0, sizeof(c));
        c.id = r.id;
        c.direction = direction;
        c.status = CUSZP_SUCCESS;
        c.d_out = s->pool;
        s->done.push_back(c);
        return CUSZP_SUCCESS;
    }
    s->held[r.id] = {0, 0, false};
    s->queue.push_back(r);
    return progress(*s);
}

int cuSZp_server_poll(cuszp_server_t s, cuszp_completion_t* done, size_t max_done, size_t* n_done, int wait) {
    if (!s || !n_done || (max_done && !done)) return CUSZP_ERROR_INVALID_ARGUMENT;
    *n_done = 0;
    std::lock_guard<std::mutex> lk(s->mu);
    int rc = progress(*s);
    if (rc) return rc;
    while (wait && s->done.empty() && any_busy(*s)) {
        for (Lane& L : s->lanes) {   // block on one running lane, then look again
            if (!L.busy) continue;
            SRV_HIP(hipEventSynchronize(L.ev));
            break;
        }
        rc = progress(*s);
        if (rc) return rc;
    }
    while (*n_done < max_done && !s->done.empty()) {
        done[(*n_done)++] = s->done.front();
        s->done.pop_front();
    }
    return CUSZP_SUCCESS;
}

int cuSZp_server_release(cuszp_server_t s, unsigned long long id) {
    if (!s) return CUSZP_ERROR_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->held.find(id);
    if (it == s->held.end() || !it->second.complete) return CUSZP_ERROR_INVALID_ARGUMENT;   // unknown or running
    if (it->second.bytes) s->alloc.release(it->second.off, it->second.bytes);
    s->held.erase(it);
    return progress(*s);
}

int cuSZp_server_stats(cuszp_server_t s, size_t* queued, size_t* in_flight, size_t* pool_used) {
    if (!s) return CUSZP_ERROR_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lk(s->mu);
    size_t f = s->admitted.size();
    for (const Lane& L : s->lanes) f += L.reqs.size();
    if (queued) *queued = s->queue.size();
    if (in_flight) *in_flight = f;
    if (pool_used) *pool_used = (size_t)s->alloc.used();
    return CUSZP_SUCCESS;
}

void cuSZp_server_destroy(cuszp_server_t s) {
    if (!s) return;
    for (Lane& L : s->lanes) lane_free(L);
    delete s;
}

int cuSZp_server_replay(cuszp_server_t s, const cuszp_request_t* reqs, size_t count, int sequential,
                        double* latency_s, int* statuses, double* makespan_s) {
    if (!s || (count && (!reqs || !latency_s || !statuses))) return CUSZP_ERROR_INVALID_ARGUMENT;
    for (size_t i = 1; i < count; ++i)
        if (reqs[i].arrival_s < reqs[i - 1].arrival_s) return CUSZP_ERROR_INVALID_ARGUMENT;
    {
        // the replay owns the server: no other request may be outstanding (the
        // sequential form also uses the pool's first bytes as its output)
        std::lock_guard<std::mutex> lk(s->mu);
        if (!s->held.empty() || !s->done.empty()) return CUSZP_ERROR_INVALID_ARGUMENT;
    }
    const double t0 = now_s();
    if (sequential) {
        // one request at a time through the single-array entries, output at the
        // start of the pool, each waiting for the previous one to finish
        hipStream_t st = s->lanes[0].st;
        for (size_t i = 0; i < count; ++i) {
            const cuszp_request_t& q = reqs[i];
            while (now_s() - t0 < q.arrival_s) {
            }
            const cuszp_batch_item_t& it = q.item;
            int rc;
            if (q.direction == CUSZP_BATCH_COMPRESS) {
                size_t sz = 0;
                rc = cuSZp_compress_ex(it.d_in, s->pool, s->pool_bytes, it.nbEle, &sz, it.errorBound, it.dim, it.dims,
                                       it.type, it.mode, st);
            } else {
                rc = cuSZp_decompress_ex(s->pool, (const unsigned char*)it.d_in, it.nbEle, it.capacity, it.errorBound,
                                         it.dim, it.dims, it.type, it.mode, st);
            }
            statuses[i] = rc;
            latency_s[i] = now_s() - t0 - q.arrival_s;
        }
        if (makespan_s) *makespan_s = now_s() - t0;
        return CUSZP_SUCCESS;
    }
    std::unordered_map<unsigned long long, size_t> index;
    std::vector<cuszp_completion_t> buf(256);
    size_t next = 0, finished = 0;
    while (finished < count) {
        const double t = now_s() - t0;
        while (next < count && reqs[next].arrival_s <= t) {
            unsigned long long id = 0;
            const int rc = cuSZp_server_submit(s, reqs[next].direction, &reqs[next].item, &id);
            if (rc) {
                statuses[next] = rc;
                latency_s[next] = 0;
                ++finished;
            } else {
                index[id] = next;
            }
            ++next;
        }
        size_t nd = 0;
        const int rc = cuSZp_server_poll(s, buf.data(), buf.size(), &nd, 0);
        if (rc) return rc;
        const double tn = now_s() - t0;
        for (size_t k = 0; k < nd; ++k) {
            const size_t i = index[buf[k].id];
            index.erase(buf[k].id);
            statuses[i] = buf[k].status;
            latency_s[i] = tn - reqs[i].arrival_s;
            cuSZp_server_release(s, buf[k].id);
            ++finished;
        }
    }
    if (makespan_s) *makespan_s = now_s() - t0;
    return CUSZP_SUCCESS;
}

}  // extern "C"
