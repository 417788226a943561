// cuSZp CPU oracle — TEST INFRASTRUCTURE ONLY.
//
// A plain C++17/OpenMP restatement of the per-block semantics of cuSZp's fused
// compress/decompress kernels. It is the checker for the HIP product path:
// only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
// it. Nothing in cuszp_amd/ links or calls it.
//
// Build: oracle/Makefile (g++ -O2 -ffp-contract=off -fno-fast-math -fopenmp).
//
// Reference semantics followed (paths relative to the cuSZp source tree):
//   quantisation f32 ........ src/cuSZp_kernels_1D_f32.cu:3-20 (PTX mul/setp/add/cvt.rzi/sub)
//   quantisation f64 ........ src/cuSZp_kernels_1D_f64.cu:3-8
//   reciprocal 0.5/eb ....... src/cuSZp_kernels_1D_f32.cu:48, src/cuSZp_kernels_1D_f64.cu:33
//   bit length .............. src/cuSZp_kernels_1D_f32.cu:23-28
//   1D residual + rate ...... src/cuSZp_kernels_1D_f32.cu:561-645 (plain), :69-142 (fixed)
//   1D signed max ........... src/cuSZp_kernels_1D_f32.cu:552,585 (int maxQuant)
//   2D residual + rate ...... src/cuSZp_kernels_2D_f32.cu:695-815 (unsigned maxQuant :709)
//   3D residual + rate ...... src/cuSZp_kernels_3D_f32.cu:661-757
//   outlier selection ....... src/cuSZp_kernels_1D_f32.cu:1165-1183, 2D :1488-1506, 3D :1385-1399
//   payload layout .......... src/cuSZp_kernels_1D_f32.cu:723-798, :1259-1661 (outlier bytes LE)
//   decoder ................. src/cuSZp_kernels_1D_f32.cu:912-1033, :1700-2259;
//                             src/cuSZp_kernels_2D_f32.cu:1249-1320; 3D :2174-3094
//   header size 1D .......... src/cuSZp_entry_1D_f32.cu:43 (ceil(N/32768)*1024)
//   header size 2D/3D ....... src/cuSZp_kernels_2D_f32.cu:46 (roundup4(blockNum))
//   block counts ............ src/cuSZp_entry_2D_f32.cu:21-23, src/cuSZp_entry_3D_f32.cu:21-24
//   decoded value ........... src/cuSZp_kernels_1D_f32.cu:490 (currQuant * eb * 2)
//   relative bound .......... examples/cuSZp.cpp:176-210 (value range loop, float product)
//   quality metrics ......... src/cuSZp_utility.cu:891-950 (computePSNR), :761-889 (SSIM)
//
// Decided semantics where the reference is ambiguous or buggy (each pinned by a
// test in tests/):
//   * quantisation is unfused: RN(x*r), then RN(t+0.5), then saturating trunc
//     (NaN -> 0), then a wrapping int32 subtract;
//   * cmpSize is the true stream length H + sum(block bytes);
//   * 2D/3D header pad bytes [blockNum, roundup4(blockNum)) are 0;
//   * decompression writes +0.0 for every in-range element of a rate-0 block;
//   * each block's payload holds exactly the fields the reference decoder parses
//     from its header byte (only differs from the reference compressor when a
//     2D/3D outlier-mode residual equals -2^31, where the reference is broken).
#include <cmath>
#include <cstdint>
#include <cstring>
#include <cstddef>
#include <vector>
#include <algorithm>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

enum { MODE_FIXED = 0, MODE_PLAIN = 1, MODE_OUTLIER = 2 };
enum { TYPE_F32 = 0, TYPE_F64 = 1 };

// cvt.rzi.s32 semantics: truncate toward zero, saturate, NaN -> 0
// (src/cuSZp_kernels_1D_f32.cu:15).
template <class T> inline int32_t sat_trunc(T y) {
    if (y != y) return 0;
    if (!((double)y < 2147483648.0)) return INT32_MAX;
    if (!((double)y > -2147483649.0)) return INT32_MIN;
    return (int32_t)y;
}

// src/cuSZp_kernels_1D_f32.cu:3-20 / src/cuSZp_kernels_1D_f64.cu:3-8
template <class T> inline int32_t quantize(T x, T r) {
    T t = x * r;                     // RN_T(x*r); no contraction (-ffp-contract=off)
    uint32_t s = (t >= T(-0.5)) ? 0u : 1u;
    T y = t + T(0.5);                // RN_T(t+0.5)
    return (int32_t)((uint32_t)sat_trunc<T>(y) - s);
}

inline int bitlen(uint32_t v) { return v ? 32 - __builtin_clz(v) : 0; }

struct Geo {
    int dim;            // 1, 2, 3
    size_t n;           // element count (1D)
    uint32_t X, Y, Z;   // x fastest
    size_t Bx, By, Bz;  // block grid
    size_t nblocks;     // real blocks
    size_t H;           // header bytes
    int L;              // elements per block
    int S;              // L/8
};

Geo make_geo(int dim, size_t n, uint32_t X, uint32_t Y, uint32_t Z) {
    Geo g{};
    g.dim = dim; g.n = n; g.X = X; g.Y = Y; g.Z = Z;
    if (dim == 1) {
        g.L = 32;
        g.nblocks = (n + 31) / 32;
        g.H = (n + 32768 - 1) / 32768 * 1024;           // src/cuSZp_entry_1D_f32.cu:43
    } else if (dim == 2) {
        g.L = 64;
        g.Bx = (X + 7) / 8; g.By = (Y + 7) / 8; g.Bz = Z;  // src/cuSZp_entry_2D_f32.cu:21-23
        g.nblocks = g.Bz * g.By * g.Bx;
        g.H = (g.nblocks + 3) / 4 * 4;
        g.n = (size_t)X * Y * Z;
    } else {
        g.L = 64;
        g.Bx = (X + 3) / 4; g.By = (Y + 3) / 4; g.Bz = (Z + 3) / 4;  // src/cuSZp_entry_3D_f32.cu:21-24
        g.nblocks = g.Bz * g.By * g.Bx;
        g.H = (g.nblocks + 3) / 4 * 4;
        g.n = (size_t)X * Y * Z;
    }
    g.S = g.L / 8;
    return g;
}

// Global element index of element e of block b, or SIZE_MAX if outside the array.
inline size_t elem_index(const Geo& g, size_t b, int e) {
    if (g.dim == 1) {
        size_t i = b * 32 + (size_t)e;
        return i < g.n ? i : SIZE_MAX;
    }
    size_t bx = b % g.Bx, by = (b / g.Bx) % g.By, bz = b / (g.Bx * g.By);
    size_t x, y, z;
    if (g.dim == 2) {
        z = bz; y = by * 8 + (size_t)(e >> 3); x = bx * 8 + (size_t)(e & 7);
    } else {
        z = bz * 4 + (size_t)(e >> 4); y = by * 4 + (size_t)((e >> 2) & 3); x = bx * 4 + (size_t)(e & 3);
    }
    if (x >= g.X || y >= g.Y || z >= g.Z) return SIZE_MAX;
    return (z * g.Y + y) * (size_t)g.X + x;
}

// Index of the predecessor of element e in the block's Lorenzo chain, -1 if none.
//   1D: src/cuSZp_kernels_1D_f32.cu:568,580 ; 2D: src/cuSZp_kernels_2D_f32.cu:711-737 ;
//   3D: src/cuSZp_kernels_3D_f32.cu:677-734
inline int pred_of(int dim, int e) {
    if (dim == 1) return e ? e - 1 : -1;
    if (dim == 2) {
        int c = e & 7, r = e >> 3;
        if (c) return e - 1;
        return r ? e - 8 : -1;
    }
    int x = e & 3, y = (e >> 2) & 3, z = e >> 4;
    if (x) return e - 1;
    if (y) return e - 4;
    return z ? e - 16 : -1;
}

struct BlockCode {
    uint8_t byte;        // header byte as written
    int sel;             // decoder view: outlier-selected
    int nb;              // decoder view: outlier bytes
    int R;               // decoder view: planes
    uint32_t size;       // decoder view: payload bytes
    uint32_t a[64];      // magnitudes
    uint64_t sgn;        // sign bits, bit e = element e
};

inline void decode_view(int mode, int S, uint8_t byte, int& sel, int& nb, int& R, uint32_t& size) {
    if (mode == MODE_OUTLIER) {   // src/cuSZp_kernels_1D_f32.cu:1705-1711
        sel = byte >> 7;
        nb = ((byte >> 5) & 3) + 1;
        R = byte & 0x1f;
        size = sel ? (uint32_t)(S + S * R + nb) : (R ? (uint32_t)(S + S * R) : 0u);
    } else {
        sel = 0; nb = 0; R = byte;
        size = R ? (uint32_t)(S + S * R) : 0u;
    }
}

template <class T>
void encode_block(const Geo& g, int mode, const T* data, T r, size_t b, BlockCode& bc) {
    const int L = g.L;
    int32_t q[64];
    bool valid[64];
    for (int e = 0; e < L; ++e) {
        size_t i = elem_index(g, b, e);
        valid[e] = (i != SIZE_MAX);
        q[e] = valid[e] ? quantize<T>(data[i], r) : 0;
    }
    bc.sgn = 0;
    for (int e = 0; e < L; ++e) {
        int32_t d = 0;
        if (valid[e]) {
            if (mode == MODE_FIXED) {
                d = q[e];
            } else {
                int p = pred_of(g.dim, e);
                d = (int32_t)((uint32_t)q[e] - (uint32_t)(p >= 0 ? q[p] : 0));
            }
        }
        bc.a[e] = d < 0 ? 0u - (uint32_t)d : (uint32_t)d;
        if (d < 0) bc.sgn |= (uint64_t)1 << e;
    }
    // 1D keeps the running maximum in a signed int (src/cuSZp_kernels_1D_f32.cu:552),
    // so |INT_MIN| (0x80000000 as int) never wins; 2D/3D use unsigned
    // (src/cuSZp_kernels_2D_f32.cu:709).
    auto rank = [&](uint32_t v) -> uint32_t { return (g.dim == 1 && v > 0x7fffffffu) ? 0u : v; };
    uint32_t mx = 0, mx2 = 0;
    for (int e = 0; e < L; ++e) {
        uint32_t v = rank(bc.a[e]);
        mx = std::max(mx, v);
        if (e) mx2 = std::max(mx2, v);
    }
    int byte;
    if (mode != MODE_OUTLIER) {
        byte = bitlen(mx);
    } else {
        const int S = g.S;
        int fr1 = bitlen(mx), fr2 = bitlen(mx2);
        int nb = (bitlen(bc.a[0]) + 7) / 8;
        int ofs1 = fr1 ? S + fr1 * S : 0;
        int ofs2 = fr2 ? S + fr2 * S + nb : S + nb;
        byte = (ofs1 <= ofs2) ? fr1 : (fr2 | 0x80 | ((nb - 1) << 5));
    }
    bc.byte = (uint8_t)(byte & 0xff);
    decode_view(mode, g.S, bc.byte, bc.sel, bc.nb, bc.R, bc.size);
}

// Serialise one block's payload (bc.size bytes) at dst.
void write_payload(const Geo& g, const BlockCode& bc, uint8_t* dst) {
    if (!bc.size) return;
    const int L = g.L, S = g.S;
    uint8_t* p = dst;
    if (bc.sel) {
        uint32_t o = bc.a[0];
        for (int k = 0; k < bc.nb; ++k) *p++ = (uint8_t)((o >> (8 * k)) & 0xff);
    }
    for (int k = 0; k < S; ++k) {
        uint8_t v = 0;
        for (int j = 0; j < 8; ++j) v |= (uint8_t)(((bc.sgn >> (8 * k + j)) & 1) << (7 - j));
        *p++ = v;
    }
    for (int pl = 0; pl < bc.R; ++pl) {
        for (int k = 0; k < S; ++k) {
            uint8_t v = 0;
            for (int j = 0; j < 8; ++j) {
                int e = 8 * k + j;
                uint32_t av = (bc.sel && e == 0) ? 0u : bc.a[e];
                v |= (uint8_t)(((av >> pl) & 1) << (7 - j));
            }
            *p++ = v;
        }
    }
    (void)L;
}

template <class T>
long long compress_t(const Geo& g, int mode, const T* data, T eb, uint8_t* out, size_t cap, int nthreads) {
    const T r = T(0.5f) / eb;  // src/cuSZp_kernels_1D_f32.cu:48 (f32), _1D_f64.cu:33 (f64)
    const size_t nb = g.nblocks;
    std::vector<uint32_t> sizes(nb + 1, 0);
    if (cap < g.H) return -1;
    std::memset(out, 0, g.H);
#pragma omp parallel for schedule(static) num_threads(nthreads)
    for (long long b = 0; b < (long long)nb; ++b) {
        BlockCode bc;
        encode_block<T>(g, mode, data, r, (size_t)b, bc);
        out[b] = bc.byte;
        sizes[b] = bc.size;
    }
    // exclusive scan of block sizes
    std::vector<size_t> offs(nb + 1, 0);
    size_t acc = 0;
    for (size_t b = 0; b < nb; ++b) { offs[b] = acc; acc += sizes[b]; }
    offs[nb] = acc;
    const size_t total = g.H + acc;
    if (total > cap) return -1;
#pragma omp parallel for schedule(static) num_threads(nthreads)
    for (long long b = 0; b < (long long)nb; ++b) {
        if (!sizes[b]) continue;
        BlockCode bc;
        encode_block<T>(g, mode, data, r, (size_t)b, bc);
        write_payload(g, bc, out + g.H + offs[b]);
    }
    return (long long)total;
}

template <class T>
int decompress_t(const Geo& g, int mode, T* out, T eb, const uint8_t* cmp, size_t cmp_size, int nthreads) {
    const size_t nb = g.nblocks;
    if (cmp_size < g.H) return -2;
    std::vector<size_t> offs(nb + 1, 0);
    size_t acc = 0;
    for (size_t b = 0; b < nb; ++b) {
        int sel, nbb, R; uint32_t size;
        decode_view(mode, g.S, cmp[b], sel, nbb, R, size);
        offs[b] = acc; acc += size;
    }
    if (g.H + acc > cmp_size) return -3;
    const int L = g.L, S = g.S;
#pragma omp parallel for schedule(static) num_threads(nthreads)
    for (long long bb = 0; bb < (long long)nb; ++bb) {
        const size_t b = (size_t)bb;
        int sel, nbb, R; uint32_t size;
        decode_view(mode, S, cmp[b], sel, nbb, R, size);
        if (!size) {  // rate-0 block: all residuals zero -> +0.0 (decided: written, not skipped)
            for (int e = 0; e < L; ++e) {
                size_t i = elem_index(g, b, e);
                if (i != SIZE_MAX) out[i] = T(0);
            }
            continue;
        }
        const uint8_t* p = cmp + g.H + offs[b];
        uint32_t a[64] = {0};
        uint32_t o = 0;
        if (sel) {
            for (int k = 0; k < nbb; ++k) o |= (uint32_t)(*p++) << (8 * k);
        }
        uint64_t sgn = 0;
        for (int k = 0; k < S; ++k) {
            uint8_t v = *p++;
            for (int j = 0; j < 8; ++j) sgn |= (uint64_t)((v >> (7 - j)) & 1) << (8 * k + j);
        }
        for (int pl = 0; pl < R; ++pl) {
            for (int k = 0; k < S; ++k) {
                uint8_t v = *p++;
                for (int j = 0; j < 8; ++j) a[8 * k + j] |= (uint32_t)((v >> (7 - j)) & 1) << pl;
            }
        }
        if (sel) a[0] = o;   // element 0's plane bits are ignored (src/cuSZp_kernels_1D_f32.cu:1887)
        int32_t q[64];
        for (int e = 0; e < L; ++e) {
            int32_t d = ((sgn >> e) & 1) ? (int32_t)(0u - a[e]) : (int32_t)a[e];
            if (mode == MODE_FIXED) {
                q[e] = d;
            } else {
                int pe = pred_of(g.dim, e);
                q[e] = (int32_t)((uint32_t)d + (uint32_t)(pe >= 0 ? q[pe] : 0));
            }
        }
        for (int e = 0; e < L; ++e) {
            size_t i = elem_index(g, b, e);
            if (i == SIZE_MAX) continue;
            T m = (T)q[e] * eb;              // RN_T(q*eb)   (src/cuSZp_kernels_1D_f32.cu:490)
            out[i] = m * T(2);               // RN_T(.*2)
        }
    }
    return 0;
}

bool geo_ok(int dim, size_t n, uint32_t X, uint32_t Y, uint32_t Z) {
    if (dim == 1) return true;
    if (dim != 2 && dim != 3) return false;
    if (!X || !Y || !Z || (X % 4)) return false;
    return n == (size_t)X * Y * Z;
}

// Reconstruction quality as src/cuSZp_utility.cu:891-950 computes it, one
// element at a time in the reference's order (f32: float error, squares and
// relative error, double sums; f64: double throughout). out: mse, range, psnr,
// normErr, normErr_norm, nrmse, max |err|, max pointwise relative error,
// correlation (acEff).
template <class T>
void quality_t(const T* ori, const T* dec, size_t n, double* out) {
    double mx = ori[0], mn = ori[0];
    double diffMax = dec[0] > ori[0] ? dec[0] - ori[0] : ori[0] - dec[0];
    double sum1 = 0, sum2 = 0, sum22 = 0;
    for (size_t i = 0; i < n; i++) {
        sum1 += ori[i];
        sum2 += dec[i];
        sum22 += (T)(dec[i] * dec[i]);
    }
    const double mean1 = sum1 / n, mean2 = sum2 / n;
    double sum3 = 0, sum4 = 0, sum = 0, prodSum = 0, maxrel = 0;
    for (size_t i = 0; i < n; i++) {
        if (mx < ori[i]) mx = ori[i];
        if (mn > ori[i]) mn = ori[i];
        const T d = dec[i] - ori[i];
        const T err = d < 0 ? -d : d;
        if (ori[i] != 0) {
            const T a = ori[i] < 0 ? -ori[i] : ori[i];
            const double rel = (T)(err / a);
            if (maxrel < rel) maxrel = rel;
        }
        if (diffMax < err) diffMax = err;
        prodSum += (ori[i] - mean1) * (dec[i] - mean2);
        sum3 += (ori[i] - mean1) * (ori[i] - mean1);
        sum4 += (dec[i] - mean2) * (dec[i] - mean2);
        sum += (T)(err * err);
    }
    const double mse = sum / n, range = mx - mn;
    out[0] = mse;
    out[1] = range;
    out[2] = 20 * std::log10(range) - 10 * std::log10(mse);
    out[3] = std::sqrt(sum);
    out[4] = std::sqrt(sum) / std::sqrt(sum22);
    out[5] = std::sqrt(mse) / range;
    out[6] = diffMax;
    out[7] = maxrel;
    out[8] = (prodSum / n) / std::sqrt(sum3 / n) / std::sqrt(sum4 / n);
}

// Mean SSIM over 7x7x7 windows moved by 2 (src/cuSZp_utility.cu:761-889),
// windows visited z-outer, x-inner as the reference does; per window the
// original's range in the values' type, double sums, then the three factors.
template <class T>
double ssim_t(const T* x, const T* y, size_t X, size_t Y, size_t Z) {
    const int W = 7, step = 2;
    double total = 0;
    long count = 0;
    for (size_t o2 = 0; o2 + W <= Z; o2 += step)
        for (size_t o1 = 0; o1 + W <= Y; o1 += step)
            for (size_t o0 = 0; o0 + W <= X; o0 += step) {
                const size_t base = o0 + X * (o1 + Y * o2);
                T lo = x[base], hi = x[base];
                double sx = 0, sy = 0;
                for (int k = 0; k < W; k++)
                    for (int j = 0; j < W; j++)
                        for (int i = 0; i < W; i++) {
                            const size_t e = base + i + X * (j + Y * (size_t)k);
                            if (lo > x[e]) lo = x[e];
                            if (hi < x[e]) hi = x[e];
                            sx += x[e];
                            sy += y[e];
                        }
                const double n = W * W * W, mx = sx / n, my = sy / n;
                double vx = 0, vy = 0, vxy = 0;
                for (int k = 0; k < W; k++)
                    for (int j = 0; j < W; j++)
                        for (int i = 0; i < W; i++) {
                            const size_t e = base + i + X * (j + Y * (size_t)k);
                            vx += (x[e] - mx) * (x[e] - mx);
                            vy += (y[e] - my) * (y[e] - my);
                            vxy += (x[e] - mx) * (y[e] - my);
                        }
                vx /= n;
                vy /= n;
                vxy /= n;
                const double sgx = std::sqrt(vx), sgy = std::sqrt(vy);
                const T r = hi - lo;
                const double c1 = r == 0 ? 0.01 * 0.01 : 0.01 * 0.01 * r * r;
                const double c2 = r == 0 ? 0.03 * 0.03 : 0.03 * 0.03 * r * r;
                const double c3 = c2 / 2;
                total += (2 * mx * my + c1) / (mx * mx + my * my + c1) * ((2 * sgx * sgy + c2) / (sgx * sgx + sgy * sgy + c2)) *
                         ((vxy + c3) / (sgx * sgy + c3));
                count++;
            }
    return total / count;
}

}  // namespace

extern "C" {

int32_t cuszp_oracle_quantize_f32(float x, float r) { return quantize<float>(x, r); }
int32_t cuszp_oracle_quantize_f64(double x, double r) { return quantize<double>(x, r); }
float cuszp_oracle_recip_f32(float eb) { return 0.5f / eb; }
double cuszp_oracle_recip_f64(double eb) { return 0.5 / eb; }

size_t cuszp_oracle_header_size(int dim, size_t n, uint32_t X, uint32_t Y, uint32_t Z) {
    if (!geo_ok(dim, n, X, Y, Z)) return 0;
    return make_geo(dim, n, X, Y, Z).H;
}

size_t cuszp_oracle_num_blocks(int dim, size_t n, uint32_t X, uint32_t Y, uint32_t Z) {
    if (!geo_ok(dim, n, X, Y, Z)) return 0;
    return make_geo(dim, n, X, Y, Z).nblocks;
}

// Worst case: every block at 32 planes plus the largest outlier prefix.
size_t cuszp_oracle_max_size(int dim, size_t n, uint32_t X, uint32_t Y, uint32_t Z) {
    if (!geo_ok(dim, n, X, Y, Z)) return 0;
    Geo g = make_geo(dim, n, X, Y, Z);
    return g.H + g.nblocks * (size_t)(g.S + 32 * g.S + 4);
}

// Returns the compressed size, or -1 on bad arguments / insufficient capacity.
long long cuszp_oracle_compress(const void* data, int type, int dim, int mode, size_t n,
                                uint32_t X, uint32_t Y, uint32_t Z, double eb,
                                uint8_t* out, size_t cap, int nthreads) {
    if (!geo_ok(dim, n, X, Y, Z) || mode < 0 || mode > 2 || (type != 0 && type != 1)) return -1;
    Geo g = make_geo(dim, n, X, Y, Z);
    if (nthreads <= 0) nthreads = 1;
    if (g.n == 0) return 0;
    if (type == TYPE_F32)
        return compress_t<float>(g, mode, (const float*)data, (float)eb, out, cap, nthreads);
    return compress_t<double>(g, mode, (const double*)data, eb, out, cap, nthreads);
}

// Returns 0 on success, <0 on malformed / truncated streams.
int cuszp_oracle_decompress(void* out, int type, int dim, int mode, size_t n,
                            uint32_t X, uint32_t Y, uint32_t Z, double eb,
                            const uint8_t* cmp, size_t cmp_size, int nthreads) {
    if (!geo_ok(dim, n, X, Y, Z) || mode < 0 || mode > 2 || (type != 0 && type != 1)) return -1;
    Geo g = make_geo(dim, n, X, Y, Z);
    if (nthreads <= 0) nthreads = 1;
    if (g.n == 0) return 0;
    if (type == TYPE_F32)
        return decompress_t<float>(g, mode, (float*)out, (float)eb, cmp, cmp_size, nthreads);
    return decompress_t<double>(g, mode, (double*)out, eb, cmp, cmp_size, nthreads);
}

// Relative error bound as the reference CLI computes it, examples/cuSZp.cpp:176-210:
// one pass from element 0 keeping max, and min only when the element is not
// above max; then rel * (max - min) in float (f64: the difference in double,
// rounded to float). Writes min and max; returns the absolute bound.
float cuszp_oracle_rel_eb(const void* data, int type, size_t n, float rel, double* mn, double* mx) {
    if (!n) return 0.0f;
    if (type == TYPE_F32) {
        const float* x = (const float*)data;
        float hi = x[0], lo = x[0];
        for (size_t i = 0; i < n; i++) {
            if (x[i] > hi) hi = x[i];
            else if (x[i] < lo) lo = x[i];
        }
        *mn = lo;
        *mx = hi;
        return rel * (hi - lo);
    }
    const double* x = (const double*)data;
    double hi = x[0], lo = x[0];
    for (size_t i = 0; i < n; i++) {
        if (x[i] > hi) hi = x[i];
        else if (x[i] < lo) lo = x[i];
    }
    *mn = lo;
    *mx = hi;
    return rel * (float)(hi - lo);
}

void cuszp_oracle_quality(const void* ori, const void* dec, int type, size_t n, double* out) {
    if (!n) return;
    if (type == TYPE_F32) quality_t<float>((const float*)ori, (const float*)dec, n, out);
    else quality_t<double>((const double*)ori, (const double*)dec, n, out);
}

double cuszp_oracle_ssim(const void* ori, const void* dec, int type, size_t X, size_t Y, size_t Z) {
    if (X < 7 || Y < 7 || Z < 7) return -2.0;
    if (type == TYPE_F32) return ssim_t<float>((const float*)ori, (const float*)dec, X, Y, Z);
    return ssim_t<double>((const double*)ori, (const double*)dec, X, Y, Z);
}

int cuszp_oracle_max_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

}  // extern "C"
