// Reconstruction quality of a decompressed array against its original, on the
// device: the figures of the reference's computePSNR (src/cuSZp_utility.cu:891-950)
// plus the two maxima it tracks, and its windowed SSIM (:761-889), for the
// batched and serving reports.
//
// Per element the arithmetic is the reference's: for f32 data the error
// |dec - ori|, its square, dec * dec and the pointwise relative error are
// float operations, everything is accumulated in double; f64 data is handled
// in double throughout. Two reads of both arrays: pass 1 sums ori, dec, dec^2
// and err^2 and takes the value range and the error maxima; pass 2 (the means
// known) sums the centred products. Each pass is one launch whose last
// workgroup (agent-scope ticket) reduces the per-workgroup partials; pass 2
// reads pass 1's results from device memory, so the host waits once.
#include <hip/hip_runtime.h>

#include <cmath>
#include <map>
#include <mutex>

#include "../../include/cuSZp.h"

namespace cuszp_quality {

constexpr int kThreads = 256;
constexpr unsigned kMaxBlocks = 1024;
constexpr int kP1 = 8;   // sum ori, sum dec, sum dec^2, sum err^2, max ori, min ori, max err, max rel err
constexpr int kP2 = 3;   // sum (ori-m1)(dec-m2), sum (ori-m1)^2, sum (dec-m2)^2

// the reference's per-element operations for T = float (float results), and
// their double counterparts for T = double
template <class T>
__device__ __forceinline__ T abs_err(T d, T o) { return d > o ? d - o : o - d; }

template <class T>
__device__ __forceinline__ void block_sum(double* v, int k, double* sm) {
    for (int i = 0; i < k; ++i)
        for (int o = 32; o > 0; o >>= 1) v[i] += __shfl_xor(v[i], o, 64);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0)
        for (int i = 0; i < k; ++i) sm[w * 16 + i] = v[i];
    __syncthreads();
    for (int i = 0; i < k; ++i) {
        double s = 0;
        for (int j = 0; j < kThreads / 64; ++j) s += sm[j * 16 + i];
        v[i] = s;
    }
    __syncthreads();
}

__device__ __forceinline__ void block_max(double* v, int k, double* sm) {
    for (int i = 0; i < k; ++i)
        for (int o = 32; o > 0; o >>= 1) v[i] = fmax(v[i], __shfl_xor(v[i], o, 64));
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0)
        for (int i = 0; i < k; ++i) sm[w * 16 + i] = v[i];
    __syncthreads();
    for (int i = 0; i < k; ++i) {
        double s = sm[i];
        for (int j = 1; j < kThreads / 64; ++j) s = fmax(s, sm[j * 16 + i]);
        v[i] = s;
    }
    __syncthreads();
}

// Takes a ticket; true in the last workgroup of the launch (which resets it).
__device__ __forceinline__ bool last_block(unsigned* ticket, bool* flag) {
    __syncthreads();
    if (threadIdx.x == 0) {
        const unsigned t = __hip_atomic_fetch_add(ticket, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        *flag = t == gridDim.x - 1;
        if (*flag) *ticket = 0u;
    }
    __syncthreads();
    if (*flag) __atomic_thread_fence(__ATOMIC_ACQUIRE);
    return *flag;
}

__device__ __forceinline__ double ld(const double* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// part: kP1 doubles per workgroup; res1: kP1 results
template <class T>
__global__ void __launch_bounds__(kThreads) pass1_kernel(const T* __restrict__ ori, const T* __restrict__ dec,
                                                         uint64_t n, double* part, unsigned* ticket, double* res1) {
    __shared__ double sm[(kThreads / 64) * 16];
    __shared__ bool flag;
    // sums: ori, dec, dec^2, err^2; maxima: ori, -ori (min), err, rel err
    double s[4] = {0, 0, 0, 0};
    double m[4] = {-INFINITY, -INFINITY, 0.0, 0.0};
    for (uint64_t i = (uint64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kThreads) {
        const T o = ori[i], d = dec[i];
        const T err = abs_err(d, o);
        s[0] += (double)o;
        s[1] += (double)d;
        s[2] += (double)(T)(d * d);
        s[3] += (double)(T)(err * err);
        m[0] = fmax(m[0], (double)o);
        m[1] = fmax(m[1], -(double)o);
        m[2] = fmax(m[2], (double)err);
        if (o != T(0)) m[3] = fmax(m[3], (double)(T)(err / (o < T(0) ? -o : o)));
    }
    block_sum<T>(s, 4, sm);
    block_max(m, 4, sm);
    if (threadIdx.x == 0)
        for (int k = 0; k < 4; ++k) {
            part[kP1 * blockIdx.x + k] = s[k];
            part[kP1 * blockIdx.x + 4 + k] = m[k];
        }
    if (!last_block(ticket, &flag)) return;
    for (int k = 0; k < 4; ++k) {
        s[k] = 0;
        m[k] = k < 2 ? -INFINITY : 0.0;
    }
    for (unsigned b = threadIdx.x; b < gridDim.x; b += kThreads)
        for (int k = 0; k < 4; ++k) {
            s[k] += ld(part + kP1 * b + k);
            m[k] = fmax(m[k], ld(part + kP1 * b + 4 + k));
        }
    block_sum<T>(s, 4, sm);
    block_max(m, 4, sm);
    if (threadIdx.x == 0) {
        for (int k = 0; k < 4; ++k) res1[k] = s[k];
        res1[4] = m[0];
        res1[5] = -m[1];
        res1[6] = m[2];
        res1[7] = m[3];
    }
}

template <class T>
__global__ void __launch_bounds__(kThreads) pass2_kernel(const T* __restrict__ ori, const T* __restrict__ dec,
                                                         uint64_t n, const double* res1, double* part,
                                                         unsigned* ticket, double* res2) {
    __shared__ double sm[(kThreads / 64) * 16];
    __shared__ bool flag;
    const double m1 = res1[0] / (double)n, m2 = res1[1] / (double)n;
    double s[3] = {0, 0, 0};
    for (uint64_t i = (uint64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kThreads) {
        const double a = (double)ori[i] - m1, b = (double)dec[i] - m2;
        s[0] += a * b;
        s[1] += a * a;
        s[2] += b * b;
    }
    block_sum<T>(s, 3, sm);
    if (threadIdx.x == 0)
        for (int k = 0; k < 3; ++k) part[kP2 * blockIdx.x + k] = s[k];
    if (!last_block(ticket, &flag)) return;
    for (int k = 0; k < 3; ++k) s[k] = 0;
    for (unsigned b = threadIdx.x; b < gridDim.x; b += kThreads)
        for (int k = 0; k < 3; ++k) s[k] += ld(part + kP2 * b + k);
    block_sum<T>(s, 3, sm);
    if (threadIdx.x == 0)
        for (int k = 0; k < 3; ++k) res2[k] = s[k];
}

// SSIM over 7x7x7 windows moved by 2 along each axis (src/cuSZp_utility.cu:761-889):
// per window, the original's min and max, both means, both variances and the
// covariance (double sums of the values), then luminance * contrast * structure
// with c1 = (0.01 R)^2, c2 = (0.03 R)^2, c3 = c2 / 2, R the window's original
// range (1 when the range is 0). One thread per window; the mean over windows
// goes through per-workgroup partials and the last workgroup.
constexpr int kWin = 7, kWinStep = 2;

template <class T>
__device__ double window_ssim(const T* __restrict__ x, const T* __restrict__ y, uint64_t s1, uint64_t s0,
                              uint64_t o0, uint64_t o1, uint64_t o2) {
    const uint64_t b = o0 + s0 * (o1 + s1 * o2);
    T xmin = x[b], xmax = x[b];
    double xs = 0, ys = 0;
    for (int k = 0; k < kWin; ++k)
        for (int j = 0; j < kWin; ++j) {
            const uint64_t r = b + s0 * ((uint64_t)j + s1 * (uint64_t)k);
#pragma unroll
            for (int i = 0; i < kWin; ++i) {
                const T a = x[r + i], c = y[r + i];
                xmin = a < xmin ? a : xmin;
                xmax = a > xmax ? a : xmax;
                xs += (double)a;
                ys += (double)c;
            }
        }
    constexpr double np = kWin * kWin * kWin;
    const double mx = xs / np, my = ys / np;
    double vx = 0, vy = 0, vxy = 0;
    for (int k = 0; k < kWin; ++k)
        for (int j = 0; j < kWin; ++j) {
            const uint64_t r = b + s0 * ((uint64_t)j + s1 * (uint64_t)k);
#pragma unroll
            for (int i = 0; i < kWin; ++i) {
                const double a = (double)x[r + i] - mx, c = (double)y[r + i] - my;
                vx += a * a;
                vy += c * c;
                vxy += a * c;
            }
        }
    vx /= np;
    vy /= np;
    vxy /= np;
    const double sx = sqrt(vx), sy = sqrt(vy);
    const double rr = (double)(T)(xmax - xmin);   // the range in the values' type
    const double c1 = rr == 0 ? 0.01 * 0.01 : 0.01 * 0.01 * rr * rr;
    const double c2 = rr == 0 ? 0.03 * 0.03 : 0.03 * 0.03 * rr * rr;
    const double c3 = c2 / 2;
    const double lum = (2 * mx * my + c1) / (mx * mx + my * my + c1);
    const double con = (2 * sx * sy + c2) / (sx * sx + sy * sy + c2);
    const double str = (vxy + c3) / (sx * sy + c3);
    return lum * con * str;
}

template <class T>
__global__ void __launch_bounds__(kThreads) ssim_kernel(const T* __restrict__ x, const T* __restrict__ y,
                                                        uint64_t s1, uint64_t s0, uint64_t nw0, uint64_t nw1,
                                                        uint64_t nw, double* part, unsigned* ticket, double* res) {
    __shared__ double sm[(kThreads / 64) * 16];
    __shared__ bool flag;
    double v[1] = {0};
    for (uint64_t w = (uint64_t)blockIdx.x * kThreads + threadIdx.x; w < nw; w += (uint64_t)gridDim.x * kThreads) {
        const uint64_t w0 = w % nw0, w1 = (w / nw0) % nw1, w2 = w / (nw0 * nw1);
        v[0] += window_ssim(x, y, s1, s0, w0 * kWinStep, w1 * kWinStep, w2 * kWinStep);
    }
    block_sum<T>(v, 1, sm);
    if (threadIdx.x == 0) part[blockIdx.x] = v[0];
    if (!last_block(ticket, &flag)) return;
    v[0] = 0;
    for (unsigned b = threadIdx.x; b < gridDim.x; b += kThreads) v[0] += ld(part + b);
    block_sum<T>(v, 1, sm);
    if (threadIdx.x == 0) res[0] = v[0] / (double)nw;
}

struct State {
    double* part = nullptr;   // kP1 * kMaxBlocks
    unsigned* ticket = nullptr;
    double* res = nullptr;    // kP1 + kP2
    double* h_res = nullptr;  // pinned
    std::mutex mu;
};
std::mutex g_mu;
std::map<std::pair<int, hipStream_t>, State*> g_state;

#define Q_HIP(call)                                    \
    do {                                               \
        hipError_t e_ = (call);                        \
        if (e_ != hipSuccess) return CUSZP_ERROR_HIP;  \
    } while (0)

int state_for(int dev, hipStream_t st, State*& out) {
    std::lock_guard<std::mutex> lk(g_mu);
    State*& s = g_state[{dev, st}];
    if (!s) {
        State* n = new State();
        if (hipMalloc((void**)&n->part, kP1 * kMaxBlocks * sizeof(double)) != hipSuccess ||
            hipMalloc((void**)&n->ticket, 256) != hipSuccess ||
            hipMalloc((void**)&n->res, (kP1 + kP2) * sizeof(double)) != hipSuccess ||
            hipHostMalloc((void**)&n->h_res, (kP1 + kP2) * sizeof(double), hipHostMallocDefault) != hipSuccess ||
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
int run(const T* ori, const T* dec, size_t n, double* out, hipStream_t st) {
    int dev = 0;
    Q_HIP(hipGetDevice(&dev));
    State* s = nullptr;
    int rc = state_for(dev, st, s);
    if (rc) return rc;
    std::lock_guard<std::mutex> lk(s->mu);
    uint64_t blocks = (n + 8ull * kThreads - 1) / (8ull * kThreads);
    if (blocks < 1) blocks = 1;
    if (blocks > kMaxBlocks) blocks = kMaxBlocks;
    hipLaunchKernelGGL(pass1_kernel<T>, dim3((unsigned)blocks), dim3(kThreads), 0, st, ori, dec, (uint64_t)n, s->part,
                       s->ticket, s->res);
    Q_HIP(hipGetLastError());
    hipLaunchKernelGGL(pass2_kernel<T>, dim3((unsigned)blocks), dim3(kThreads), 0, st, ori, dec, (uint64_t)n,
                       (const double*)s->res, s->part, s->ticket, s->res + kP1);
    Q_HIP(hipGetLastError());
    Q_HIP(hipMemcpyAsync(s->h_res, s->res, (kP1 + kP2) * sizeof(double), hipMemcpyDeviceToHost, st));
    Q_HIP(hipStreamSynchronize(st));
    const double* r = s->h_res;
    const double N = (double)n;
    const double mse = r[3] / N, range = r[4] - r[5];
    out[0] = mse;                                                // src/cuSZp_utility.cu:937-946
    out[1] = range;
    out[2] = 20 * log10(range) - 10 * log10(mse);
    out[3] = sqrt(r[3]);
    out[4] = sqrt(r[3]) / sqrt(r[2]);
    out[5] = sqrt(mse) / range;
    out[6] = r[6];                                               // max |dec - ori|
    out[7] = r[7];                                               // max |dec - ori| / |ori| over ori != 0
    out[8] = (r[kP1] / N) / sqrt(r[kP1 + 1] / N) / sqrt(r[kP1 + 2] / N);   // correlation (acEff)
    return CUSZP_SUCCESS;
}

template <class T>
int run_ssim(const T* x, const T* y, cuszp_dims_t d, double* out, hipStream_t st) {
    if (d.x < (unsigned)kWin || d.y < (unsigned)kWin || d.z < (unsigned)kWin) return CUSZP_ERROR_INVALID_DIMS;
    int dev = 0;
    Q_HIP(hipGetDevice(&dev));
    State* s = nullptr;
    int rc = state_for(dev, st, s);
    if (rc) return rc;
    std::lock_guard<std::mutex> lk(s->mu);
    const uint64_t nw0 = (d.x - kWin) / kWinStep + 1, nw1 = (d.y - kWin) / kWinStep + 1,
                   nw2 = (d.z - kWin) / kWinStep + 1, nw = nw0 * nw1 * nw2;
    uint64_t blocks = (nw + kThreads - 1) / kThreads;
    if (blocks > kMaxBlocks) blocks = kMaxBlocks;
    hipLaunchKernelGGL(ssim_kernel<T>, dim3((unsigned)blocks), dim3(kThreads), 0, st, x, y, (uint64_t)d.y,
                       (uint64_t)d.x, nw0, nw1, nw, s->part, s->ticket, s->res);
    Q_HIP(hipGetLastError());
    Q_HIP(hipMemcpyAsync(s->h_res, s->res, sizeof(double), hipMemcpyDeviceToHost, st));
    Q_HIP(hipStreamSynchronize(st));
    *out = s->h_res[0];
    return CUSZP_SUCCESS;
}

}  // namespace cuszp_quality

extern "C" int cuSZp_ssim(const void* d_oriData, const void* d_decData, cuszp_dims_t dims, cuszp_type_t type,
                          double* ssim, hipStream_t stream) {
    if (!d_oriData || !d_decData || !ssim) return CUSZP_ERROR_INVALID_ARGUMENT;
    if (type == CUSZP_TYPE_FLOAT)
        return cuszp_quality::run_ssim((const float*)d_oriData, (const float*)d_decData, dims, ssim, stream);
    if (type == CUSZP_TYPE_DOUBLE)
        return cuszp_quality::run_ssim((const double*)d_oriData, (const double*)d_decData, dims, ssim, stream);
    return CUSZP_ERROR_INVALID_ARGUMENT;
}

extern "C" int cuSZp_quality(const void* d_oriData, const void* d_decData, size_t nbEle, cuszp_type_t type,
                             double* metrics, hipStream_t stream) {
    if (!d_oriData || !d_decData || !nbEle || !metrics) return CUSZP_ERROR_INVALID_ARGUMENT;
    if (type == CUSZP_TYPE_FLOAT)
        return cuszp_quality::run((const float*)d_oriData, (const float*)d_decData, nbEle, metrics, stream);
    if (type == CUSZP_TYPE_DOUBLE)
        return cuszp_quality::run((const double*)d_oriData, (const double*)d_decData, nbEle, metrics, stream);
    return CUSZP_ERROR_INVALID_ARGUMENT;
}
