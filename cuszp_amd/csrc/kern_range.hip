// Value range of a device array and the relative error bound derived from it,
// the step before compression when the bound is relative: the reference CLI
// computes it on the host (examples/cuSZp.cpp:176-210): min and max over the
// array (comparisons, so NaNs are skipped unless the first element is NaN,
// which makes both NaN), then errorBound = rel * (max - min) in float
// arithmetic; for f64 arrays the difference is taken in double and rounded to
// float first.
//
// One launch reads the array once: each workgroup reduces a strided share
// with 16-byte loads, writes its (min, max) pair, and the last workgroup to
// finish (an agent-scope ticket) reduces the pairs and resets the ticket.
#include <hip/hip_runtime.h>

#include <cmath>
#include <map>
#include <mutex>

#include "../../include/cuSZp.h"

namespace cuszp_range {

constexpr int kThreads = 256;
constexpr unsigned kMaxBlocks = 2048;

template <class T>
struct Vec;
template <>
struct Vec<float> {
    typedef float type __attribute__((ext_vector_type(4)));
    static constexpr int N = 4;
};
template <>
struct Vec<double> {
    typedef double type __attribute__((ext_vector_type(2)));
    static constexpr int N = 2;
};

template <class T>
__device__ __forceinline__ void take(T v, T& mn, T& mx) {
    mn = v < mn ? v : mn;   // NaN compares false: skipped
    mx = v > mx ? v : mx;
}

template <class T>
__device__ __forceinline__ void block_reduce(T& mn, T& mx, T* smn, T* smx) {
    for (int o = 32; o > 0; o >>= 1) {
        const T a = __shfl_xor(mn, o, 64);
        const T b = __shfl_xor(mx, o, 64);
        mn = a < mn ? a : mn;
        mx = b > mx ? b : mx;
    }
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        smn[w] = mn;
        smx[w] = mx;
    }
    __syncthreads();
    mn = smn[0];
    mx = smx[0];
    for (int i = 1; i < kThreads / 64; ++i) {
        mn = smn[i] < mn ? smn[i] : mn;
        mx = smx[i] > mx ? smx[i] : mx;
    }
}

// x: the array; head: elements before the first 16-byte boundary; nv: whole
// vectors after it. part: 2 * gridDim.x values; res: {min, max}.
template <class T>
__global__ void __launch_bounds__(kThreads) range_kernel(const T* __restrict__ x, uint64_t n, uint64_t head,
                                                         uint64_t nv, T* __restrict__ part, unsigned* ticket,
                                                         T* __restrict__ res) {
    typedef typename Vec<T>::type V;
    constexpr int N = Vec<T>::N;
    __shared__ T smn[kThreads / 64], smx[kThreads / 64];
    __shared__ bool last;
    T mn = INFINITY, mx = -INFINITY;
    const V* __restrict__ xv = (const V*)(x + head);
    const uint64_t stride = (uint64_t)gridDim.x * kThreads;
    uint64_t i = (uint64_t)blockIdx.x * kThreads + threadIdx.x;
    for (; i + 3 * stride < nv; i += 4 * stride) {   // four loads in flight per lane
        V v[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) v[k] = __builtin_nontemporal_load(xv + i + k * stride);
#pragma unroll
        for (int k = 0; k < 4; ++k) {
#pragma unroll
            for (int j = 0; j < N; ++j) take((T)v[k][j], mn, mx);
        }
    }
    for (; i < nv; i += stride) {
        const V v = __builtin_nontemporal_load(xv + i);
#pragma unroll
        for (int j = 0; j < N; ++j) take((T)v[j], mn, mx);
    }
    if (blockIdx.x == 0) {   // elements before the first vector and after the last
        const uint64_t tail0 = head + nv * N;
        if (threadIdx.x < head) take(x[threadIdx.x], mn, mx);
        for (uint64_t t = tail0 + threadIdx.x; t < n; t += kThreads) take(x[t], mn, mx);
    }
    block_reduce(mn, mx, smn, smx);
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = mn;
        part[2 * blockIdx.x + 1] = mx;
        const unsigned t = __hip_atomic_fetch_add(ticket, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        last = t == gridDim.x - 1;
    }
    __syncthreads();
    if (!last) return;
    __atomic_thread_fence(__ATOMIC_ACQUIRE);
    mn = INFINITY;
    mx = -INFINITY;
    for (unsigned b = threadIdx.x; b < gridDim.x; b += kThreads) {
        const T a = __hip_atomic_load(part + 2 * b, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const T c = __hip_atomic_load(part + 2 * b + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        mn = a < mn ? a : mn;
        mx = c > mx ? c : mx;
    }
    __syncthreads();   // smn/smx are reused
    block_reduce(mn, mx, smn, smx);
    if (threadIdx.x == 0) {
        const T x0 = x[0];
        const bool nan0 = x0 != x0;   // the reference's loop starts from element 0
        res[0] = nan0 ? x0 : mn;
        res[1] = nan0 ? x0 : mx;
        *ticket = 0u;   // ready for the next launch on this stream
    }
}

struct State {
    void* part = nullptr;      // 2 * kMaxBlocks doubles
    unsigned* ticket = nullptr;
    double* res = nullptr;     // device {min, max}
    double* h_res = nullptr;   // pinned
    std::mutex mu;
};
std::mutex g_mu;
std::map<std::pair<int, hipStream_t>, State*> g_state;

#define RANGE_HIP(call)                                \
    do {                                               \
        hipError_t e_ = (call);                        \
        if (e_ != hipSuccess) return CUSZP_ERROR_HIP;  \
    } while (0)

int state_for(int dev, hipStream_t st, State*& out) {
    std::lock_guard<std::mutex> lk(g_mu);
    State*& s = g_state[{dev, st}];
    if (!s) {
        State* n = new State();
        if (hipMalloc(&n->part, 2 * kMaxBlocks * sizeof(double)) != hipSuccess ||
            hipMalloc((void**)&n->ticket, 256) != hipSuccess || hipMalloc((void**)&n->res, 2 * sizeof(double)) != hipSuccess ||
            hipHostMalloc((void**)&n->h_res, 2 * sizeof(double), hipHostMallocDefault) != hipSuccess ||
            hipMemsetAsync(n->ticket, 0, 256, st) != hipSuccess) {
            delete n;   // (device buffers of a failed set-up are not reclaimed)
            return CUSZP_ERROR_HIP;
        }
        s = n;
    }
    out = s;
    return CUSZP_SUCCESS;
}

template <class T>
int run(const T* x, size_t n, double* mn, double* mx, hipStream_t st) {
    int dev = 0;
    RANGE_HIP(hipGetDevice(&dev));
    State* s = nullptr;
    int rc = state_for(dev, st, s);
    if (rc) return rc;
    std::lock_guard<std::mutex> lk(s->mu);
    constexpr int N = Vec<T>::N;
    const uint64_t mis = (uint64_t)((uintptr_t)x % 16) / sizeof(T);
    uint64_t head = mis ? (uint64_t)N - mis : 0;
    if (head > n) head = n;
    const uint64_t nv = (n - head) / N;
    uint64_t blocks = (nv + 4ull * kThreads - 1) / (4ull * kThreads);
    if (blocks < 1) blocks = 1;
    if (blocks > kMaxBlocks) blocks = kMaxBlocks;
    hipLaunchKernelGGL(range_kernel<T>, dim3((unsigned)blocks), dim3(kThreads), 0, st, x, (uint64_t)n, head, nv,
                       (T*)s->part, s->ticket, (T*)s->res);
    RANGE_HIP(hipGetLastError());
    RANGE_HIP(hipMemcpyAsync(s->h_res, s->res, 2 * sizeof(T), hipMemcpyDeviceToHost, st));
    RANGE_HIP(hipStreamSynchronize(st));
    const T* r = (const T*)s->h_res;
    *mn = (double)r[0];
    *mx = (double)r[1];
    return CUSZP_SUCCESS;
}

}  // namespace cuszp_range

extern "C" {

int cuSZp_value_range(const void* d_data, size_t nbEle, cuszp_type_t type, double* min_val, double* max_val,
                      hipStream_t stream) {
    if (!d_data || !nbEle || !min_val || !max_val) return CUSZP_ERROR_INVALID_ARGUMENT;
    if (type == CUSZP_TYPE_FLOAT) return cuszp_range::run((const float*)d_data, nbEle, min_val, max_val, stream);
    if (type == CUSZP_TYPE_DOUBLE) return cuszp_range::run((const double*)d_data, nbEle, min_val, max_val, stream);
    return CUSZP_ERROR_INVALID_ARGUMENT;
}

int cuSZp_rel_error_bound(const void* d_data, size_t nbEle, cuszp_type_t type, float rel, float* errorBound,
                          hipStream_t stream) {
    if (!errorBound) return CUSZP_ERROR_INVALID_ARGUMENT;
    double mn = 0, mx = 0;
    const int rc = cuSZp_value_range(d_data, nbEle, type, &mn, &mx, stream);
    if (rc) return rc;
    // examples/cuSZp.cpp:186 (f32: float subtraction) and :207 (f64: double, then float)
    const float range = type == CUSZP_TYPE_FLOAT ? (float)mx - (float)mn : (float)(mx - mn);
    *errorBound = rel * range;
    return CUSZP_SUCCESS;
}

}  // extern "C"
