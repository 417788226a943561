// Fused error-bounded compress / decompress kernels for gfx950 (MI355X).
//
// One compress kernel and one decompress kernel, templated over
// {1D, 2D, 3D} x {fixed, plain, outlier} x {f32, f64}. They replace the 36
// per-variant kernels of cuSZp (src/cuSZp_kernels_{1D,2D,3D}_f{32,64}.cu) and
// produce/consume the same byte stream.
//
// Work decomposition (wave64-native; see the repository's design notes):
//   * a "group" is 64 consecutive elements of cuSZp block order: two 32-element
//     blocks in 1D, one 8x8 (2D) or 4x4x4 (3D) block otherwise. One wave holds
//     one group, one element per lane;
//   * a workgroup is 4 waves; each wave owns G consecutive groups, so a tile is
//     4*G groups and tiles are claimed in ticket order;
//   * lanes hold elements in "mirror" order: lane L holds element L^7, which is
//     exactly the bit order of cuSZp's sign/plane bytes (element e at byte e/8,
//     bit 7-e%8). A 64-lane carry-out mask of v_add_co_u32 therefore IS one
//     bit-plane of two 1D blocks (or of one 2D/3D block) in stream byte order;
//   * the inter-tile exclusive scan of compressed bytes is a single-pass
//     decoupled lookback: one 64-bit status word per tile packs
//     {epoch, flag, value}, is read 64 predecessors at a time by one wave, and
//     needs only relaxed agent-scope atomics (no fences: the value travels in
//     the same word as its flag). Tickets make forward progress independent of
//     workgroup dispatch order.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cuszp {

constexpr int kWaveSize = 64;
constexpr int kWaves = 4;
constexpr int kThreads = kWaveSize * kWaves;

enum : int { kFixed = 0, kPlain = 1, kOutlier = 2 };

// Tile status word: [63:40] epoch tag, [39:38] flag, [37:0] byte count.
constexpr int kEpochShift = 40;
constexpr uint64_t kFlagAgg = 1ull << 38;
constexpr uint64_t kFlagPrefix = 2ull << 38;
constexpr uint64_t kFlagMask = 3ull << 38;
constexpr uint64_t kValueMask = (1ull << 38) - 1;

// Error bits reported through ScanArgs::err.
constexpr uint32_t kErrOverflow = 1u;    // compressed stream exceeds output capacity
constexpr uint32_t kErrTruncated = 2u;   // stream shorter than its header describes

struct Geom {
    uint64_t n;         // elements
    uint64_t nblocks;   // cuSZp blocks
    uint64_t H;         // header bytes
    uint64_t ntiles;    // workgroups in the launch
    uint64_t XY;        // X*Y
    uint32_t X, Y, Z;
    uint32_t Bx, By;    // blocks along x and y (2D/3D)
};

struct ScanArgs {
    unsigned long long* status;     // [ntiles] tile status words
    unsigned long long* ticket;     // tile ticket counter; 0 at launch, reset to 0 by the last claim
    unsigned long long* total;      // out: total stream bytes (written by the last tile), may be null
    unsigned int* err;              // sticky error bits (kErr*)
    unsigned long long tag;         // epoch << kEpochShift
};

// ---------------------------------------------------------------------------
// wave64 primitives
// ---------------------------------------------------------------------------
template <int CTRL, int ROWS = 0xf, int BANKS = 0xf>
__device__ __forceinline__ uint32_t dpp(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, ROWS, BANKS, false);
}
constexpr int kDppRowShr1 = 0x111, kDppRowShr2 = 0x112, kDppRowShr4 = 0x114, kDppRowShr8 = 0x118;
constexpr int kDppWaveShr1 = 0x138, kDppRowMirror = 0x140, kDppHalfMirror = 0x141;
constexpr int kDppBcast15 = 0x142, kDppBcast31 = 0x143;
constexpr int kDppQuadSwap1 = 0xB1;   // quad_perm [1,0,3,2]
constexpr int kDppQuadSwap2 = 0x4E;   // quad_perm [2,3,0,1]
constexpr int kDppQuadShr1 = 0x90;    // quad_perm [0,0,1,2]
constexpr int kDppQuadShr2 = 0x40;    // quad_perm [0,0,0,1]

__device__ __forceinline__ uint32_t rdlane(uint32_t v, uint32_t l) {
    return (uint32_t)__builtin_amdgcn_readlane((int)v, (int)l);
}
// old[l] <- v (v and l wave-uniform)
__device__ __forceinline__ uint32_t wrlane(uint32_t v, uint32_t l, uint32_t old) {
    asm("v_writelane_b32 %0, %1, m0" : "+v"(old) : "s"(v), "{m0}"(l));
    return old;
}
__device__ __forceinline__ uint64_t ballot(bool p) { return __builtin_amdgcn_ballot_w64(p); }

// b <- 2b; returns the 64-lane mask of the bits shifted out (bit 31 of old b).
__device__ __forceinline__ uint64_t take_top_bit(uint32_t& b) {
    uint32_t nb;
    uint64_t m;
    asm("v_add_co_u32 %0, %1, %2, %2" : "=v"(nb), "=s"(m) : "v"(b));
    b = nb;
    return m;
}
// a <- 2a + (bit `lane` of m)
__device__ __forceinline__ uint32_t push_bit(uint32_t a, uint64_t m) {
    uint32_t r;
    uint64_t c;
    asm("v_addc_co_u32 %0, %1, %2, %2, %3" : "=v"(r), "=s"(c) : "v"(a), "s"(m));
    return r;
}

// OR over each 16-lane row; every lane of a row receives the row's OR.
__device__ __forceinline__ uint32_t row_or(uint32_t v) {
    v |= dpp<kDppQuadSwap1>(v);
    v |= dpp<kDppQuadSwap2>(v);
    v |= dpp<kDppHalfMirror>(v);
    v |= dpp<kDppRowMirror>(v);
    return v;
}

// Inclusive prefix sums (wrapping u32).
__device__ __forceinline__ uint32_t scan_row16(uint32_t v) {
    v += dpp<kDppRowShr1>(v);
    v += dpp<kDppRowShr2>(v);
    v += dpp<kDppRowShr4>(v);
    v += dpp<kDppRowShr8>(v);
    return v;
}
__device__ __forceinline__ uint32_t scan_seg32(uint32_t v) {
    v = scan_row16(v);
    v += dpp<kDppBcast15, 0xa>(v);
    return v;
}
__device__ __forceinline__ uint32_t scan_wave(uint32_t v) {
    v = scan_seg32(v);
    v += dpp<kDppBcast31, 0xc>(v);
    return v;
}

__device__ __forceinline__ uint64_t wave_sum_u64(uint64_t v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, kWaveSize);
    return v;
}

__device__ __forceinline__ int bitlen(uint32_t v) { return v ? 32 - __builtin_clz(v) : 0; }

// ---------------------------------------------------------------------------
// quantisation: t = RN(x*r); s = (t >= -0.5 ? 0 : 1); q = sat_trunc(RN(t+0.5)) - s
// (src/cuSZp_kernels_1D_f32.cu:3-20, src/cuSZp_kernels_1D_f64.cu:3-8).
// The kernels are built with -ffp-contract=off so mul and add round separately;
// v_cvt_i32_f{32,64} truncates, saturates and maps NaN to 0 like cvt.rzi.s32.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int32_t cvt_i32_sat(float y) {
    int32_t r;
    asm("v_cvt_i32_f32 %0, %1" : "=v"(r) : "v"(y));
    return r;
}
__device__ __forceinline__ int32_t cvt_i32_sat(double y) {
    int32_t r;
    asm("v_cvt_i32_f64 %0, %1" : "=v"(r) : "v"(y));
    return r;
}
template <class T>
__device__ __forceinline__ uint32_t quantize(T x, T r) {
    const T t = x * r;
    const uint32_t s = (t >= T(-0.5)) ? 0u : 1u;
    const T y = t + T(0.5);
    return (uint32_t)cvt_i32_sat(y) - s;
}

// ---------------------------------------------------------------------------
// block-byte bookkeeping (decoder view of a header byte; cuSZp's decoders
// src/cuSZp_kernels_1D_f32.cu:1705-1711 for outlier, :837-843 otherwise)
// ---------------------------------------------------------------------------
struct BlockView {
    uint32_t sel, nb, R, size, nwords, head;
};
template <int MODE, int S>
__device__ __forceinline__ BlockView view_of(uint32_t byte) {
    BlockView v;
    if constexpr (MODE == kOutlier) {
        v.sel = byte >> 7;
        v.nb = ((byte >> 5) & 3) + 1;
        v.R = byte & 0x1f;
        v.size = v.sel ? S + S * v.R + v.nb : (v.R ? S + S * v.R : 0);
        v.head = v.sel ? v.nb : 0;
    } else {
        v.sel = 0;
        v.nb = 0;
        v.R = byte;
        v.size = v.R ? S + S * v.R : 0;
        v.head = 0;
    }
    v.nwords = v.size ? (S / 4) * (1 + v.R) : 0;   // sign + planes, in u32 words
    return v;
}

// Header byte for one block given the OR of its ranked magnitudes excluding
// element 0 (or2), element 0's ranked magnitude (r0) and raw magnitude (o).
// src/cuSZp_kernels_1D_f32.cu:1165-1183 (outlier), :641 (fixed/plain).
template <int MODE, int S>
__device__ __forceinline__ uint32_t rate_byte(uint32_t or2, uint32_t r0, uint32_t o) {
    const int fr1 = bitlen(or2 | r0);
    if constexpr (MODE != kOutlier) {
        return (uint32_t)fr1;
    } else {
        const int fr2 = bitlen(or2);
        const int nb = (bitlen(o) + 7) / 8;
        const int ofs1 = fr1 ? S + fr1 * S : 0;
        const int ofs2 = fr2 ? S + fr2 * S + nb : S + nb;
        return (ofs1 <= ofs2) ? (uint32_t)fr1 : (uint32_t)((fr2 | 0x80 | ((nb - 1) << 5)) & 0xff);
    }
}

// Byte-granular little-endian store / load of 1..4 bytes (outlier prefixes).
__device__ __forceinline__ void store_le(uint8_t* p, uint32_t v, uint32_t nb) {
    typedef uint32_t __attribute__((aligned(1))) u32u;
    typedef uint16_t __attribute__((aligned(1))) u16u;
    if (nb == 4) {
        *(u32u*)p = v;
    } else {
        if (nb & 2) { *(u16u*)p = (uint16_t)v; p += 2; v >>= 16; }
        if (nb & 1) { *p = (uint8_t)v; }
    }
}
__device__ __forceinline__ uint32_t load_le(const uint8_t* p, uint32_t nb) {
    uint32_t v = 0;
    for (uint32_t k = 0; k < nb; ++k) v |= (uint32_t)p[k] << (8 * k);
    return v;
}
__device__ __forceinline__ void store_u32u(uint8_t* p, uint32_t v) {
    typedef uint32_t __attribute__((aligned(1))) u32u;
    *(u32u*)p = v;
}
__device__ __forceinline__ uint32_t load_u32u(const uint8_t* p) {
    typedef uint32_t __attribute__((aligned(1))) u32u;
    return *(const u32u*)p;
}

// ---------------------------------------------------------------------------
// decoupled lookback (single pass, wave-parallel, ticketed tiles)
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint64_t ld_status(const unsigned long long* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_status(unsigned long long* p, uint64_t v) {
    __hip_atomic_store(p, (unsigned long long)v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Called by one whole wave. Publishes the tile aggregate, walks back over
// predecessors 64 at a time, publishes the inclusive prefix and returns the
// exclusive prefix (same value in every lane).
__device__ __forceinline__ uint64_t tile_lookback(const ScanArgs& sa, uint64_t tile, uint64_t agg, int lane) {
    const uint64_t tag = sa.tag;
    if (tile == 0) {
        if (lane == 0) st_status(sa.status, tag | kFlagPrefix | agg);
        return 0;
    }
    if (lane == 0) st_status(sa.status + tile, tag | kFlagAgg | agg);
    uint64_t excl = 0;
    int64_t hi = (int64_t)tile;
    for (;;) {
        const int64_t idx = hi - 1 - lane;
        uint64_t w = tag | kFlagPrefix;   // virtual tile -1: inclusive prefix 0
        bool ready = true;
        if (idx >= 0) {
            w = ld_status(sa.status + idx);
            ready = ((w & ~(kFlagMask | kValueMask)) == tag) && (w & kFlagMask);
        }
        // wait until the lanes up to the nearest published prefix are all ready
        for (;;) {
            const uint64_t notready = ballot(!ready);
            const uint64_t pm = ballot(ready && (w & kFlagMask) == kFlagPrefix);
            uint64_t need = ~0ull;
            if (pm) {
                const int k = __builtin_ctzll(pm);
                need = (k == 63) ? ~0ull : ((2ull << k) - 1);
            }
            if (!(notready & need)) break;
            __builtin_amdgcn_s_sleep(1);
            if (!ready) {
                w = ld_status(sa.status + idx);
                ready = ((w & ~(kFlagMask | kValueMask)) == tag) && (w & kFlagMask);
            }
        }
        const uint64_t pm = ballot((w & kFlagMask) == kFlagPrefix && ready);
        if (pm) {
            const int k = __builtin_ctzll(pm);
            excl += wave_sum_u64(lane <= k ? (w & kValueMask) : 0);
            break;
        }
        excl += wave_sum_u64(w & kValueMask);
        hi -= kWaveSize;
    }
    if (lane == 0) st_status(sa.status + tile, tag | kFlagPrefix | (excl + agg));
    return excl;
}

// Workgroup-level: combine the 4 wave totals, run the lookback on wave 0 and
// return {wave's exclusive offset within the stream payload, tile inclusive}.
struct TileOffsets {
    uint64_t wave_base;   // payload offset (excluding H) of this wave's first block
    uint64_t tile_incl;   // payload bytes up to and including this tile
};
__device__ __forceinline__ TileOffsets tile_scan(const ScanArgs& sa, uint64_t tile, uint64_t wtot,
                                                 int wave, int lane, uint64_t* s_wtot, uint64_t* s_excl) {
    if (lane == 0) s_wtot[wave] = wtot;
    __syncthreads();
    uint64_t wexcl = 0, ttot = 0;
#pragma unroll
    for (int w = 0; w < kWaves; ++w) {
        const uint64_t v = s_wtot[w];
        wexcl += (w < wave) ? v : 0;
        ttot += v;
    }
    if (wave == 0) {
        const uint64_t excl = tile_lookback(sa, tile, ttot, lane);
        if (lane == 0) *s_excl = excl;
    }
    __syncthreads();
    const uint64_t excl = *s_excl;
    return TileOffsets{excl + wexcl, excl + ttot};
}

// Tiles are numbered in the order workgroups start, so a tile only ever waits
// on tiles whose workgroups are already resident. The workgroup that draws the
// last ticket returns the counter to 0 for the next launch on this scan state:
// every other draw of this launch precedes its own in the counter's order.
__device__ __forceinline__ uint64_t claim_tile(const ScanArgs& sa, uint64_t ntiles, uint64_t* s_tile) {
    if (threadIdx.x == 0) {
        const uint64_t t = __hip_atomic_fetch_add(sa.ticket, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (t == ntiles - 1) __hip_atomic_store(sa.ticket, 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        *s_tile = t;
    }
    __syncthreads();
    return *s_tile;
}

// ---------------------------------------------------------------------------
// Block walking for 2D/3D: wave-uniform block coordinates advanced by +1 block.
// Block order: b = z*By*Bx + by*Bx + bx (src/cuSZp_kernels_2D_f32.cu:69-73).
// ---------------------------------------------------------------------------
struct Walker {
    uint64_t b;
    uint32_t bx, by, bz;
    __device__ __forceinline__ void init(uint64_t b0, const Geom& g) {
        b = b0;
        const uint64_t t = b0 / g.Bx;
        bx = (uint32_t)(b0 - t * g.Bx);
        by = (uint32_t)(t % g.By);
        bz = (uint32_t)(t / g.By);
    }
    __device__ __forceinline__ void next(const Geom& g) {
        ++b;
        if (++bx == g.Bx) {
            bx = 0;
            if (++by == g.By) { by = 0; ++bz; }
        }
    }
};

// Per-lane geometry of an element of a group (natural element order e = lane).
template <int DIM>
struct LaneGeo;

template <>
struct LaneGeo<1> {
    __device__ __forceinline__ void init(int, const Geom&) {}
};
template <>
struct LaneGeo<2> {
    uint32_t r, c;
    uint64_t off;
    __device__ __forceinline__ void init(int e, const Geom& g) {
        r = (uint32_t)e >> 3;
        c = (uint32_t)e & 7;
        off = (uint64_t)r * g.X + c;
    }
};
template <>
struct LaneGeo<3> {
    uint32_t zz, yy, xx;
    uint64_t off;
    __device__ __forceinline__ void init(int e, const Geom& g) {
        zz = (uint32_t)e >> 4;
        yy = ((uint32_t)e >> 2) & 3;
        xx = (uint32_t)e & 3;
        off = (uint64_t)zz * g.XY + (uint64_t)yy * g.X + xx;
    }
};

// Element index and validity of lane `e` of the current group.
template <int DIM>
__device__ __forceinline__ bool elem_of(const Geom& g, const LaneGeo<DIM>& lg, const Walker& wk,
                                        uint64_t grp, int e, uint64_t& idx) {
    if constexpr (DIM == 1) {
        idx = grp * 64 + (uint64_t)e;
        return idx < g.n;
    } else if constexpr (DIM == 2) {
        const uint32_t y = wk.by * 8 + lg.r, x = wk.bx * 8 + lg.c;
        idx = ((uint64_t)wk.bz * g.Y + (uint64_t)wk.by * 8) * g.X + (uint64_t)wk.bx * 8 + lg.off;
        return (wk.b < g.nblocks) & (y < g.Y) & (x < g.X);
    } else {
        const uint32_t z = wk.bz * 4 + lg.zz, y = wk.by * 4 + lg.yy;
        idx = ((uint64_t)wk.bz * 4 * g.Y + (uint64_t)wk.by * 4) * g.X + (uint64_t)wk.bx * 4 + lg.off;
        return (wk.b < g.nblocks) & (z < g.Z) & (y < g.Y);
    }
}

// Predecessor lane of element e in the block's Lorenzo chain (-1: none).
template <int DIM>
__device__ __forceinline__ int pred_lane(int e) {
    if constexpr (DIM == 1) {
        return (e & 31) ? e - 1 : -1;
    } else if constexpr (DIM == 2) {
        return (e & 7) ? e - 1 : ((e >> 3) ? e - 8 : -1);
    } else {
        return (e & 3) ? e - 1 : (((e >> 2) & 3) ? e - 4 : ((e >> 4) ? e - 16 : -1));
    }
}

// Inverse Lorenzo: rebuild q from residuals d in natural element order.
template <int DIM>
__device__ __forceinline__ uint32_t inverse_lorenzo(uint32_t d, int e) {
    if constexpr (DIM == 1) {
        return scan_seg32(d);
    } else if constexpr (DIM == 2) {
        // q(r,c) = sum_{r'<=r} d(r',0) + sum_{1<=c'<=c} d(r,c')
        const bool c0 = (e & 7) == 0;
        const uint32_t col = scan_wave(c0 ? d : 0u);
        uint32_t v = c0 ? 0u : d;
        uint32_t t = dpp<kDppRowShr1>(v);
        v += ((e & 7) >= 1) ? t : 0u;
        t = dpp<kDppRowShr2>(v);
        v += ((e & 7) >= 2) ? t : 0u;
        v += dpp<kDppRowShr4, 0xf, 0xa>(v);
        return col + v;
    } else {
        // q(z,y,x) = sum_{z'<=z} d(z',0,0) + sum_{1<=y'<=y} d(z,y',0) + sum_{1<=x'<=x} d(z,y,x')
        const int x = e & 3, y = (e >> 2) & 3;
        const uint32_t vz = (x == 0 && y == 0) ? d : 0u;
        const uint32_t vy = (x == 0 && y != 0) ? d : 0u;
        uint32_t vx = x ? d : 0u;
        vx += dpp<kDppQuadShr1>(vx);
        vx += dpp<kDppQuadShr2>(vx);
        return scan_wave(vz) + scan_row16(vy) + vx;
    }
}

// ---------------------------------------------------------------------------
// compress
// ---------------------------------------------------------------------------
template <int DIM, int MODE, class T, int G>
__global__ void __launch_bounds__(kThreads)
compress_kernel(const T* __restrict__ in, uint8_t* __restrict__ out, const Geom geo, const T r,
                const ScanArgs sa, const uint64_t cap) {
    constexpr int S = (DIM == 1) ? 4 : 8;            // bytes per plane
    __shared__ uint64_t s_tile, s_excl;
    __shared__ uint64_t s_wtot[kWaves];
    __shared__ uint32_t s_aux[kWaves][G][2];         // outlier prefix values / top planes

    const int lane = (int)__lane_id();
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const uint64_t tile = claim_tile(sa, geo.ntiles, &s_tile);
    if (tile >= geo.ntiles) return;   // unreachable: exactly ntiles draws per launch
    const uint64_t grp0 = (tile * kWaves + (uint64_t)wave) * G;

    LaneGeo<DIM> lg;
    lg.init(lane, geo);
    Walker wk;
    if constexpr (DIM != 1) wk.init(grp0, geo);
    const int plane = pred_lane<DIM>(lane);
    const int plane_src = plane < 0 ? lane : plane;
    const bool is_e0 = (DIM == 1) ? ((lane & 31) == 7) : (lane == 7);   // element 0, mirror order

    uint32_t W[G];
    uint32_t hv = 0;  // lane g: header byte(s) of group g
    uint64_t wtot = 0;

#pragma unroll
    for (int g = 0; g < G; ++g) {
        const uint64_t grp = grp0 + g;
        uint64_t idx;
        const bool valid = elem_of<DIM>(geo, lg, wk, grp, lane, idx);
        uint32_t q = 0;
        if (valid) q = quantize<T>(in[idx], r);
        uint32_t d;
        if constexpr (MODE == kFixed) {
            d = q;
        } else if constexpr (DIM == 1) {
            uint32_t qp = dpp<kDppWaveShr1>(q);
            qp = (lane & 31) ? qp : 0u;
            d = q - qp;
        } else {
            uint32_t qp = (uint32_t)__builtin_amdgcn_ds_bpermute(plane_src << 2, (int)q);
            qp = (plane >= 0) ? qp : 0u;
            d = q - qp;
        }
        d = valid ? d : 0u;
        // mirror order: lane L now holds element L^7
        const uint32_t dm = dpp<kDppHalfMirror>(d);
        const uint64_t sgn = ballot((int32_t)dm < 0);
        const uint32_t a = ((int32_t)dm < 0) ? 0u - dm : dm;
        // 1D ranks |INT_MIN| as 0 (signed running max, src/cuSZp_kernels_1D_f32.cu:552)
        const uint32_t ra = (DIM == 1) ? (a & 0x7fffffffu) : a;
        const uint32_t orow = row_or(is_e0 ? 0u : ra);
        const uint32_t o0 = rdlane(orow, 0), o1 = rdlane(orow, 16), o2 = rdlane(orow, 32), o3 = rdlane(orow, 48);

        uint32_t bsrc;
        uint32_t w = 0;
        if constexpr (DIM == 1) {
            const uint32_t oA = rdlane(a, 7), oB = rdlane(a, 39);
            const uint32_t byteA = rate_byte<MODE, 4>(o0 | o1, oA & 0x7fffffffu, oA);
            const uint32_t byteB = rate_byte<MODE, 4>(o2 | o3, oB & 0x7fffffffu, oB);
            const BlockView vA = view_of<MODE, 4>(byteA);
            const BlockView vB = view_of<MODE, 4>(byteB);
            hv = wrlane(byteA | (byteB << 8), g, hv);
            wtot += vA.size + vB.size;
            if constexpr (MODE == kOutlier) {
                if (lane == 0) { s_aux[wave][g][0] = oA; s_aux[wave][g][1] = oB; }
            }
            const bool sel = (lane < 32) ? vA.sel : vB.sel;
            bsrc = __builtin_bitreverse32((sel && is_e0) ? 0u : ra);
            w = wrlane((uint32_t)sgn, 0, w);
            w = wrlane((uint32_t)(sgn >> 32), 32, w);
            const uint32_t Rmax = vA.R > vB.R ? vA.R : vB.R;
            for (uint32_t t = 0; t < Rmax; ++t) {
                const uint64_t m = take_top_bit(bsrc);
                w = wrlane((uint32_t)m, 1 + t, w);
                w = wrlane((uint32_t)(m >> 32), 33 + t, w);
            }
        } else {
            const uint32_t o = rdlane(a, 7);
            const uint32_t byte = rate_byte<MODE, 8>(o0 | o1 | o2 | o3, o, o);
            const BlockView v = view_of<MODE, 8>(byte);
            hv = wrlane(byte, g, hv);
            wtot += v.size;
            bsrc = __builtin_bitreverse32((v.sel && is_e0) ? 0u : a);
            w = wrlane((uint32_t)sgn, 0, w);
            w = wrlane((uint32_t)(sgn >> 32), 1, w);
            const uint32_t R31 = v.R < 31 ? v.R : 31;
            for (uint32_t t = 0; t < R31; ++t) {
                const uint64_t m = take_top_bit(bsrc);
                w = wrlane((uint32_t)m, 2 + 2 * t, w);
                w = wrlane((uint32_t)(m >> 32), 3 + 2 * t, w);
            }
            if constexpr (MODE == kOutlier) {
                if (lane == 0) s_aux[wave][g][0] = o;
            } else {
                if (v.R == 32) {   // only when a residual is -2^31: plane 31 lives in LDS
                    const uint64_t m = take_top_bit(bsrc);
                    if (lane == 0) { s_aux[wave][g][0] = (uint32_t)m; s_aux[wave][g][1] = (uint32_t)(m >> 32); }
                }
            }
            wk.next(geo);
        }
        W[g] = w;
    }

    // header bytes of this wave's groups
    {
        if (lane < G) {
            const uint64_t grp = grp0 + lane;
            if constexpr (DIM == 1) {
                typedef uint16_t __attribute__((aligned(1))) u16u;
                *(u16u*)(out + 2 * grp) = (uint16_t)hv;
            } else {
                if (grp < geo.H) out[grp] = (uint8_t)hv;
            }
        }
    }

    const TileOffsets to = tile_scan(sa, tile, wtot, wave, lane, s_wtot, &s_excl);
    if (tile == geo.ntiles - 1 && threadIdx.x == 0 && sa.total) *sa.total = geo.H + to.tile_incl;

    uint64_t base = geo.H + to.wave_base;
    if (base + wtot > cap) {   // never write past the caller's buffer
        if (lane == 0) atomicOr(sa.err, kErrOverflow);
        return;
    }
#pragma unroll
    for (int g = 0; g < G; ++g) {
        if constexpr (DIM == 1) {
            const uint32_t hg = rdlane(hv, g);
            const BlockView vA = view_of<MODE, 4>(hg & 0xff);
            const BlockView vB = view_of<MODE, 4>(hg >> 8);
            const uint64_t offA = base, offB = base + vA.size;
            base += vA.size + vB.size;
            const uint32_t l = (uint32_t)lane & 31;
            const bool hiH = lane >= 32;
            const uint32_t nw = hiH ? vB.nwords : vA.nwords;
            const uint64_t wo = hiH ? offB + vB.head : offA + vA.head;
            if (l < nw) store_u32u(out + wo + 4 * l, W[g]);
            if constexpr (MODE == kOutlier) {
                if (lane == 0 && vA.sel) store_le(out + offA, s_aux[wave][g][0], vA.nb);
                if (lane == 32 && vB.sel) store_le(out + offB, s_aux[wave][g][1], vB.nb);
            }
        } else {
            const BlockView v = view_of<MODE, 8>(rdlane(hv, g));
            const uint64_t off = base;
            base += v.size;
            const uint32_t nw = v.nwords < 64 ? v.nwords : 64;
            if ((uint32_t)lane < nw) store_u32u(out + off + v.head + 4 * lane, W[g]);
            if constexpr (MODE == kOutlier) {
                if (lane == 0 && v.sel) store_le(out + off, s_aux[wave][g][0], v.nb);
            } else {
                if (v.R == 32 && lane < 2) store_u32u(out + off + 256 + 4 * lane, s_aux[wave][g][lane]);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// decompress
// ---------------------------------------------------------------------------
template <int DIM, int MODE, class T, int G>
__global__ void __launch_bounds__(kThreads)
decompress_kernel(T* __restrict__ out, const uint8_t* __restrict__ cmp, const Geom geo, const T eb,
                  const ScanArgs sa, const uint64_t cmp_size) {
    constexpr int S = (DIM == 1) ? 4 : 8;
    __shared__ uint64_t s_tile, s_excl;
    __shared__ uint64_t s_wtot[kWaves];

    const int lane = (int)__lane_id();
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const uint64_t tile = claim_tile(sa, geo.ntiles, &s_tile);
    if (tile >= geo.ntiles) return;   // unreachable: exactly ntiles draws per launch
    const uint64_t grp0 = (tile * kWaves + (uint64_t)wave) * G;

    // header bytes and block sizes of this wave's groups: lane g <-> group g
    uint32_t hv = 0, sz = 0;
    if (lane < G) {
        const uint64_t grp = grp0 + lane;
        if constexpr (DIM == 1) {
            typedef uint16_t __attribute__((aligned(1))) u16u;
            hv = *(const u16u*)(cmp + 2 * grp);
            sz = view_of<MODE, 4>(hv & 0xff).size + view_of<MODE, 4>(hv >> 8).size;
        } else {
            if (grp < geo.nblocks) {
                hv = cmp[grp];
                sz = view_of<MODE, 8>(hv).size;
            }
        }
    }
    const uint32_t incl = scan_wave(sz);
    const uint32_t pre = incl - sz;
    const uint64_t wtot = rdlane(incl, 63);
    const TileOffsets to = tile_scan(sa, tile, wtot, wave, lane, s_wtot, &s_excl);
    if (tile == geo.ntiles - 1 && threadIdx.x == 0 && sa.total) *sa.total = geo.H + to.tile_incl;
    const uint64_t base = geo.H + to.wave_base;
    const bool intact = base + wtot <= cmp_size;
    if (!intact && lane == 0) atomicOr(sa.err, kErrTruncated);

    LaneGeo<DIM> lg;
    lg.init(lane, geo);
    Walker wk;
    if constexpr (DIM != 1) wk.init(grp0, geo);

#pragma unroll 2
    for (int g = 0; g < G; ++g) {
        const uint64_t grp = grp0 + g;
        const uint32_t hg = rdlane(hv, g);
        const uint64_t goff = base + rdlane(pre, g);
        uint32_t a = 0, sgnw0, sgnw1;
        bool zero_blk;
        if constexpr (DIM == 1) {
            const BlockView vA = view_of<MODE, 4>(hg & 0xff);
            const BlockView vB = view_of<MODE, 4>(hg >> 8);
            const uint64_t offA = goff, offB = goff + vA.size;
            const uint32_t l = (uint32_t)lane & 31;
            const bool hiH = lane >= 32;
            const uint32_t nw = intact ? (hiH ? vB.nwords : vA.nwords) : 0;
            const uint64_t wo = hiH ? offB + vB.head : offA + vA.head;
            const uint32_t w = (l < nw) ? load_u32u(cmp + wo + 4 * l) : 0u;
            sgnw0 = rdlane(w, 0);
            sgnw1 = rdlane(w, 32);
            const uint32_t Rmax = vA.R > vB.R ? vA.R : vB.R;
            for (int32_t t = (int32_t)Rmax - 1; t >= 0; --t) {
                const uint32_t lo = ((uint32_t)t < vA.R) ? rdlane(w, 1 + t) : 0u;
                const uint32_t hi = ((uint32_t)t < vB.R) ? rdlane(w, 33 + t) : 0u;
                a = push_bit(a, (uint64_t)lo | ((uint64_t)hi << 32));
            }
            if constexpr (MODE == kOutlier) {
                if (intact && vA.sel && lane == 7) a = load_le(cmp + offA, vA.nb);
                if (intact && vB.sel && lane == 39) a = load_le(cmp + offB, vB.nb);
            }
            zero_blk = !intact || (hiH ? vB.size == 0 : vA.size == 0);
        } else {
            const BlockView v = view_of<MODE, 8>(hg);
            const uint32_t nw = intact ? v.nwords : 0;
            const uint64_t wo = goff + v.head;
            const uint32_t w = ((uint32_t)lane < nw) ? load_u32u(cmp + wo + 4 * lane) : 0u;
            sgnw0 = rdlane(w, 0);
            sgnw1 = rdlane(w, 1);
            if constexpr (MODE != kOutlier) {
                if (v.R == 32 && nw) {
                    const uint32_t w2 = (lane < 2) ? load_u32u(cmp + wo + 256 + 4 * lane) : 0u;
                    a = push_bit(a, (uint64_t)rdlane(w2, 0) | ((uint64_t)rdlane(w2, 1) << 32));
                }
            }
            const uint32_t R31 = v.R < 31 ? v.R : 31;
            for (int32_t t = (int32_t)R31 - 1; t >= 0; --t) {
                a = push_bit(a, (uint64_t)rdlane(w, 2 + 2 * t) | ((uint64_t)rdlane(w, 3 + 2 * t) << 32));
            }
            if constexpr (MODE == kOutlier) {
                if (intact && v.sel && lane == 7) a = load_le(cmp + goff, v.nb);
            }
            zero_blk = !intact || v.size == 0;
        }
        const uint64_t sgn = (uint64_t)sgnw0 | ((uint64_t)sgnw1 << 32);
        const uint32_t dm = ((sgn >> lane) & 1) ? 0u - a : a;
        const uint32_t d = dpp<kDppHalfMirror>(dm);          // back to natural order
        uint32_t q;
        if constexpr (MODE == kFixed) q = d;
        else q = inverse_lorenzo<DIM>(d, lane);
        const T val = zero_blk ? T(0) : (T)(int32_t)q * eb * T(2);   // src/cuSZp_kernels_1D_f32.cu:490
        uint64_t idx;
        const bool valid = elem_of<DIM>(geo, lg, wk, grp, lane, idx);
        if (valid) out[idx] = val;
        if constexpr (DIM != 1) wk.next(geo);
    }
}

}  // namespace cuszp
