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
//   * a workgroup is WAVES waves; each wave owns G consecutive groups, so a
//     tile is WAVES*G groups. The grid is persistent and cooperatively
//     launched (every workgroup resident): workgroup w handles tiles w, w+B,
//     w+2B, ... and prefetches its next tile while it scans and stores the
//     current one. No ticket atomics: one hot device-scope atomic serialises
//     at about 12 ns per operation on MI355X;
//   * lanes hold elements in "mirror" order: lane L holds element L^7, which is
//     exactly the bit order of cuSZp's sign/plane bytes (element e at byte e/8,
//     bit 7-e%8). Each lane packs its sign into bit 31 above the 31-bit
//     magnitude; a 32x32 bit-matrix transpose across each 32-lane half (five
//     ds_swizzle xor stages) then leaves plane p's stream word in lane p and the
//     sign word in lane 31 -- for every rate at a fixed cost. Decompression runs
//     the same transpose (it is an involution);
//   * per-block bookkeeping (rates, header bytes, sizes, offsets) runs once per
//     wave with one lane per block;
//   * the inter-tile exclusive scan of compressed bytes is a single-pass
//     decoupled lookback: one 64-bit status word per tile packs
//     {epoch, flag, value}, is read 64 predecessors at a time by one wave, and
//     needs only relaxed agent-scope atomics (no fences: the value travels in
//     the same word as its flag). With all workgroups resident, the lowest
//     unfinished tile can always complete, so the walk always terminates.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

namespace cuszp {

constexpr int kWaveSize = 64;

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
constexpr uint32_t kErrStall = 4u;       // a lookback gave up waiting (internal error; output invalid)

// Bound on lookback polls (each polls the status words once and sleeps ~64
// cycles): far above any legitimate wait, it guarantees that every wave of a
// launch exits even if a predecessor tile never publishes.
constexpr uint32_t kSpinLimit = 1u << 18;

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
    unsigned long long* total;      // out: total stream bytes (written by the last tile), may be null
    unsigned int* err;              // sticky error bits (kErr*)
    unsigned long long* diag;       // [0] first stalled tile + 1, [1] its missing predecessor, [2] tag seen
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
// old[L] <- v for a compile-time lane L (v wave-uniform)
template <int L>
__device__ __forceinline__ uint32_t wrlane(uint32_t v, uint32_t old) {
    asm("v_writelane_b32 %0, %1, %2" : "+v"(old) : "s"(v), "i"(L));
    return old;
}
__device__ __forceinline__ uint64_t ballot(bool p) { return __builtin_amdgcn_ballot_w64(p); }

// ---------------------------------------------------------------------------
// 32x32 bit transpose within each 32-lane half: out lane j bit i = in lane i bit j.
// Stage J exchanges lane-index bit log2(J) with data bit log2(J): the lane whose
// bit is 0 keeps its data bits with that bit 0 and takes its partner's (lane^J)
// other half shifted into place. Stages 16 and 8 move whole bytes (v_perm),
// stages 4, 2, 1 rotate and merge (v_alignbit + v_bfi). The partner value comes
// from ds_swizzle (xor mode), which runs on the LDS pipe, not the VALU.
// ---------------------------------------------------------------------------
struct XposeConsts {
    uint32_t sh1, mk1, sh2, mk2, sh4, mk4, sel8, sel16;
    __device__ __forceinline__ void init(int lane) {
        const int l = lane & 31;
        const bool lo1 = !(l & 1), lo2 = !(l & 2), lo4 = !(l & 4), lo8 = !(l & 8), lo16 = !(l & 16);
        sh1 = lo1 ? 31 : 1;  mk1 = lo1 ? 0xAAAAAAAAu : 0x55555555u;
        sh2 = lo2 ? 30 : 2;  mk2 = lo2 ? 0xCCCCCCCCu : 0x33333333u;
        sh4 = lo4 ? 28 : 4;  mk4 = lo4 ? 0xF0F0F0F0u : 0x0F0F0F0Fu;
        sel8 = lo8 ? 0x06020400u : 0x03070105u;     // [x0,y0,x2,y2] | [y1,x1,y3,x3]
        sel16 = lo16 ? 0x05040100u : 0x03020706u;   // [x0,x1,y0,y1] | [y2,y3,x2,x3]
    }
};
template <int J>
__device__ __forceinline__ uint32_t swizzle_xor(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_ds_swizzle((int)v, 0x1f | (J << 10));
}
__device__ __forceinline__ uint32_t rotr(uint32_t v, uint32_t s) { return __builtin_amdgcn_alignbit(v, v, s); }
__device__ __forceinline__ uint32_t bfi(uint32_t m, uint32_t a, uint32_t b) { return (m & a) | (~m & b); }
__device__ __forceinline__ uint32_t xpose32(uint32_t x, const XposeConsts& c) {
    x = __builtin_amdgcn_perm(swizzle_xor<16>(x), x, c.sel16);
    x = __builtin_amdgcn_perm(swizzle_xor<8>(x), x, c.sel8);
    x = bfi(c.mk4, rotr(swizzle_xor<4>(x), c.sh4), x);
    x = bfi(c.mk2, rotr(swizzle_xor<2>(x), c.sh2), x);
    x = bfi(c.mk1, rotr(swizzle_xor<1>(x), c.sh1), x);
    return x;
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

__device__ __forceinline__ uint32_t bitlen(uint32_t v) { return v ? 32u - (uint32_t)__builtin_clz(v) : 0u; }

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

// Header byte of a block from its rate over all elements (fr1), over elements
// e >= 1 (fr2) and the bit length of element 0's magnitude (outlier mode).
// src/cuSZp_kernels_1D_f32.cu:1165-1183 (outlier), :641 (fixed/plain).
template <int MODE, int S>
__device__ __forceinline__ uint32_t rate_byte(uint32_t fr1, uint32_t fr2, uint32_t o_bits) {
    if constexpr (MODE != kOutlier) {
        return fr1;
    } else {
        const uint32_t nb = (o_bits + 7) / 8;
        const uint32_t ofs1 = fr1 ? S + fr1 * S : 0;
        const uint32_t ofs2 = fr2 ? S + fr2 * S + nb : S + nb;
        return (ofs1 <= ofs2) ? fr1 : ((fr2 | 0x80u | ((nb - 1) << 5)) & 0xffu);
    }
}

// Byte-granular little-endian store of 1..4 bytes (outlier prefixes).
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
__device__ __forceinline__ void store_u32u(uint8_t* p, uint32_t v) {
    typedef uint32_t __attribute__((aligned(1))) u32u;
    *(u32u*)p = v;
}
__device__ __forceinline__ uint32_t load_u32u(const uint8_t* p) {
    typedef uint32_t __attribute__((aligned(1))) u32u;
    return *(const u32u*)p;
}

// ---------------------------------------------------------------------------
// decoupled lookback (single pass, wave-parallel)
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint64_t ld_status(const unsigned long long* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_status(unsigned long long* p, uint64_t v) {
    __hip_atomic_store(p, (unsigned long long)v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Sum over the wave of per-lane values below 2^40: two 20-bit halves are
// reduced with DPP adds (no LDS round trips) and recombined.
__device__ __forceinline__ uint64_t wave_sum_u40(uint64_t v) {
    const uint32_t lo = scan_wave((uint32_t)v & 0xfffffu);
    const uint32_t hi = scan_wave((uint32_t)(v >> 20));
    return (uint64_t)rdlane(lo, kWaveSize - 1) + ((uint64_t)rdlane(hi, kWaveSize - 1) << 20);
}

constexpr int kLookbackWords = 4;   // predecessors read per lane and round trip: 256 per wave

__device__ __forceinline__ bool status_ready(uint64_t w, uint64_t tag) {
    return ((w & ~(kFlagMask | kValueMask)) == tag) && (w & kFlagMask);
}

// Called by one whole wave. Publishes the tile aggregate, walks back over
// predecessors 256 at a time (position p = lane + 64*k is predecessor
// tile - 1 - p), publishes the inclusive prefix and returns the exclusive
// prefix (same value in every lane). One round trip advances the walk by up to
// 256 tiles, so the prefix front keeps up with tiles issued 256x faster than a
// status round trip.
__device__ __forceinline__ uint64_t tile_lookback(const ScanArgs& sa, uint64_t tile, uint64_t agg, int lane) {
    const uint64_t tag = sa.tag;
    if (tile == 0) {
        if (lane == 0) st_status(sa.status, tag | kFlagPrefix | agg);
        return 0;
    }
    if (lane == 0) st_status(sa.status + tile, tag | kFlagAgg | agg);
    uint64_t excl = 0;
    int64_t hi = (int64_t)tile;
    uint32_t spins = 0;
    bool stalled = false;
    for (;;) {
        uint64_t w[kLookbackWords];
        bool ready[kLookbackWords];
#pragma unroll
        for (int k = 0; k < kLookbackWords; ++k) {
            const int64_t idx = hi - 1 - lane - 64 * k;
            w[k] = tag | kFlagPrefix;   // before tile 0: inclusive prefix 0
            ready[k] = true;
            if (idx >= 0) {
                w[k] = ld_status(sa.status + idx);
                ready[k] = status_ready(w[k], tag);
            }
        }
        // nearest published prefix P (position); wait until positions <= P are ready
        int P;
        for (;;) {
            P = kLookbackWords * 64;
            bool all_ready = true;
#pragma unroll
            for (int k = kLookbackWords - 1; k >= 0; --k) {
                const uint64_t pm = ballot(ready[k] && (w[k] & kFlagMask) == kFlagPrefix);
                if (pm) P = 64 * k + __builtin_ctzll(pm);
            }
#pragma unroll
            for (int k = 0; k < kLookbackWords; ++k) {
                const uint64_t nr = ballot(!ready[k]);
                const int lo = 64 * k;
                if (P >= lo + 63) {
                    all_ready &= (nr == 0);
                } else if (P >= lo) {
                    all_ready &= ((nr & ((2ull << (P - lo)) - 1)) == 0);
                }
            }
            if (all_ready) break;
            if (++spins >= kSpinLimit ||
                ((spins & 1023) == 0 && (__hip_atomic_load(sa.err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & kErrStall))) {
                // watchdog: record the first stuck predecessor and give up (the
                // stream is reported invalid; publishing below unblocks successors)
                stalled = true;
                uint64_t first = ~0ull, seen = 0;
#pragma unroll
                for (int k = 0; k < kLookbackWords; ++k) {
                    const uint64_t nr = ballot(!ready[k]);
                    if (nr && first == ~0ull) {
                        const int ln = __builtin_ctzll(nr);
                        first = (uint64_t)(hi - 1 - ln - 64 * k);
                        seen = __builtin_amdgcn_readlane((unsigned)w[k], ln) | ((uint64_t)__builtin_amdgcn_readlane((unsigned)(w[k] >> 32), ln) << 32);
                    }
                }
                if (lane == 0) {
                    atomicOr(sa.err, kErrStall);
                    if (atomicCAS(sa.diag, 0ull, tile + 1) == 0ull) {
                        sa.diag[1] = first;
                        sa.diag[2] = seen;
                    }
                }
                break;
            }
            __builtin_amdgcn_s_sleep(1);
#pragma unroll
            for (int k = 0; k < kLookbackWords; ++k) {
                const int64_t idx = hi - 1 - lane - 64 * k;
                if (!ready[k]) {
                    w[k] = ld_status(sa.status + idx);
                    ready[k] = status_ready(w[k], tag);
                }
            }
        }
        uint64_t part = 0;
#pragma unroll
        for (int k = 0; k < kLookbackWords; ++k) part += (lane + 64 * k <= P) ? (w[k] & kValueMask) : 0;
        excl += wave_sum_u40(part);
        if (P < kLookbackWords * 64 || stalled) break;
        hi -= kLookbackWords * 64;
    }
    if (lane == 0) st_status(sa.status + tile, tag | kFlagPrefix | (excl + agg));
    return excl;
}

// Workgroup-level: combine the wave totals, run the lookback on wave 0 and
// return {wave's exclusive offset within the stream payload, tile inclusive}.
struct TileOffsets {
    uint64_t wave_base;   // payload offset (excluding H) of this wave's first block
    uint64_t tile_incl;   // payload bytes up to and including this tile
};
template <int WAVES>
__device__ __forceinline__ TileOffsets tile_scan(const ScanArgs& sa, uint64_t tile, uint64_t wtot,
                                                 int wave, int lane, uint64_t* s_wtot, uint64_t* s_excl) {
    if (lane == 0) s_wtot[wave] = wtot;
    __syncthreads();
    uint64_t wexcl = 0, ttot = 0;
#pragma unroll
    for (int w = 0; w < WAVES; ++w) {
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
// per-block metadata handed from the bookkeeping lanes to pass 2 through LDS:
//   [19:0] payload offset from the wave's first block, [25:20] planes (<= 32),
//   [26] block has a payload, [27] outlier-selected, [30:28] prefix bytes
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t pack_meta(uint32_t off, uint32_t R, bool nz, uint32_t sel, uint32_t head) {
    return off | ((R < 32 ? R : 32u) << 20) | ((uint32_t)nz << 26) | (sel << 27) | (head << 28);
}
struct Meta {
    uint32_t off, R, nz, sel, head;
    __device__ __forceinline__ explicit Meta(uint32_t m)
        : off(m & 0xfffff), R((m >> 20) & 63), nz((m >> 26) & 1), sel((m >> 27) & 1), head((m >> 28) & 7) {}
};

template <int I, int N, class F>
__device__ __forceinline__ void static_for(F&& f) {
    if constexpr (I < N) {
        f(std::integral_constant<int, I>{});
        static_for<I + 1, N>(f);
    }
}

// Byte offset of lane (l, h)'s word inside a block payload (after the outlier
// prefix): lane 31 carries the sign word, lane l < 31 plane l; 2D/3D blocks
// interleave the two 32-element halves h word by word (src/cuSZp_kernels_2D_f32.cu:896-903).
template <int DIM>
__device__ __forceinline__ uint32_t word_offset(uint32_t l, uint32_t h, bool sign_in_lane31) {
    if constexpr (DIM == 1) {
        return (l == 31) ? 0u : 4u + 4u * l;
    } else {
        return (sign_in_lane31 && l == 31) ? 4u * h : 8u + 8u * l + 4u * h;
    }
}

// ---------------------------------------------------------------------------
// compress
// ---------------------------------------------------------------------------
template <int DIM, int MODE, class T, int G, int WAVES>
__global__ void __launch_bounds__(WAVES * kWaveSize)
compress_kernel(const T* __restrict__ in, uint8_t* __restrict__ out, const Geom geo, const T r,
                const ScanArgs sa, const uint64_t cap) {
    constexpr int S = (DIM == 1) ? 4 : 8;        // bytes per plane
    constexpr int BPG = (DIM == 1) ? 2 : 1;      // blocks per group
    constexpr int NB = BPG * G;                  // blocks per wave
    static_assert(NB <= kWaveSize, "one bookkeeping lane per block");
    __shared__ uint64_t s_excl;
    __shared__ uint64_t s_wtot[WAVES];
    __shared__ uint32_t s_meta[WAVES][NB];
    __shared__ uint32_t s_sgn[WAVES][G][2];     // sign words of 2D/3D blocks holding |d| = 2^31

    const int lane = (int)__lane_id();
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const uint32_t l = (uint32_t)lane & 31, h = (uint32_t)lane >> 5;

    LaneGeo<DIM> lg;
    lg.init(lane, geo);
    const int pl = pred_lane<DIM>(lane);
    const int pl_src = pl < 0 ? lane : pl;
    XposeConsts xc;
    xc.init(lane);
    // after the transpose, element 0 of a block is bit 7 of its plane words: in
    // both halves for 1D (two blocks), only in the low half for 2D/3D
    const uint32_t not_e0 = (DIM == 1 || h == 0) ? ~0x80u : ~0u;

    // All loads of a tile's G groups per wave are issued back to back; the next
    // tile's loads are issued before the current tile's scan and stores.
    T xv[G];
    uint64_t vmask = 0;   // bit g: this lane's element of group g lies inside the array
    auto load_tile = [&](uint64_t tl) {
        const uint64_t g0 = (tl * WAVES + (uint64_t)wave) * G;
        Walker wk;
        if constexpr (DIM != 1) wk.init(g0, geo);
        vmask = 0;
#pragma unroll
        for (int g = 0; g < G; ++g) {
            uint64_t idx;
            const bool valid = elem_of<DIM>(geo, lg, wk, g0 + g, lane, idx);
            xv[g] = valid ? __builtin_nontemporal_load(in + idx) : T(0);
            vmask |= (uint64_t)valid << g;
            if constexpr (DIM != 1) wk.next(geo);
        }
    };
    const uint64_t stride = gridDim.x;
    uint64_t tile = blockIdx.x;
    if (tile < geo.ntiles) load_tile(tile);
    for (; tile < geo.ntiles; tile += stride) {
    const uint64_t grp0 = (tile * WAVES + (uint64_t)wave) * G;

    uint32_t W[G];
    uint32_t bk1 = 0, bk2 = 0, bko = 0;   // lane j: block j's nonzero-plane mask (all / e >= 1), |d| of element 0
    uint32_t big = 0;                     // bit g: group g's block has a magnitude 2^31 (2D/3D)
    static_for<0, G>([&](auto gc) {
        constexpr int g = decltype(gc)::value;
        const bool valid = (vmask >> g) & 1;
        const uint32_t q = quantize<T>(xv[g], r);   // x = 0 outside the array gives q = 0
        uint32_t d;
        if constexpr (MODE == kFixed) {
            d = q;
        } else if constexpr (DIM == 1) {
            uint32_t qp = dpp<kDppWaveShr1>(q);
            qp = (lane & 31) ? qp : 0u;
            d = q - qp;
        } else {
            uint32_t qp = (uint32_t)__builtin_amdgcn_ds_bpermute(pl_src << 2, (int)q);
            qp = (pl >= 0) ? qp : 0u;
            d = q - qp;
        }
        d = valid ? d : 0u;
        const uint32_t dm = dpp<kDppHalfMirror>(d);   // lane L now holds element L^7
        const uint32_t a = ((int32_t)dm < 0) ? 0u - dm : dm;
        uint64_t m1, m2 = 0;
        bool wide = false;
        if constexpr (DIM != 1) wide = ballot((int32_t)a < 0) != 0;
        if (!wide) {
            // sign bit over the low 31 magnitude bits; 1D ranks |-2^31| as 0
            // (signed running maximum, src/cuSZp_kernels_1D_f32.cu:552), so bit 31 is free
            const uint32_t w = xpose32(bfi(0x80000000u, dm, a), xc);
            W[g] = w;
            m1 = ballot(w != 0) & 0x7fffffff7fffffffull;
            if constexpr (MODE == kOutlier) m2 = ballot((w & not_e0) != 0) & 0x7fffffff7fffffffull;
        } else {
            const uint32_t w = xpose32(a, xc);
            W[g] = w;
            m1 = ballot(w != 0);
            if constexpr (MODE == kOutlier) m2 = ballot((w & not_e0) != 0);
            const uint64_t sg = ballot((int32_t)dm < 0);
            if (lane == 0) {
                s_sgn[wave][g][0] = (uint32_t)sg;
                s_sgn[wave][g][1] = (uint32_t)(sg >> 32);
            }
            big |= 1u << g;
        }
        if constexpr (DIM == 1) {
            bk1 = wrlane<2 * g>((uint32_t)m1, bk1);
            bk1 = wrlane<2 * g + 1>((uint32_t)(m1 >> 32), bk1);
            if constexpr (MODE == kOutlier) {
                bk2 = wrlane<2 * g>((uint32_t)m2, bk2);
                bk2 = wrlane<2 * g + 1>((uint32_t)(m2 >> 32), bk2);
                bko = wrlane<2 * g>(rdlane(a, 7), bko);
                bko = wrlane<2 * g + 1>(rdlane(a, 39), bko);
            }
        } else {
            bk1 = wrlane<g>((uint32_t)m1 | (uint32_t)(m1 >> 32), bk1);
            if constexpr (MODE == kOutlier) {
                bk2 = wrlane<g>((uint32_t)m2 | (uint32_t)(m2 >> 32), bk2);
                bko = wrlane<g>(rdlane(a, 7), bko);
            }
        }
    });

    if (tile + stride < geo.ntiles) load_tile(tile + stride);   // in flight during scan + stores

    // ---- bookkeeping: lane j <-> block j of this wave -----------------------
    const uint64_t blk_idx = (DIM == 1) ? 2 * grp0 + (uint64_t)lane : grp0 + (uint64_t)lane;
    const bool blk = lane < NB;
    uint32_t byte;
    if constexpr (MODE == kOutlier) byte = rate_byte<MODE, S>(bitlen(bk1), bitlen(bk2), bitlen(bko));
    else byte = bitlen(bk1);
    const BlockView v = view_of<MODE, S>(byte);
    const uint32_t sz = blk ? v.size : 0u;
    const uint32_t incl = scan_wave(sz);
    const uint32_t off = incl - sz;
    const uint64_t wtot = rdlane(incl, kWaveSize - 1);
    if (blk) {
        s_meta[wave][lane] = pack_meta(off, v.R, sz != 0, v.sel, v.head);
        // header byte; 2D/3D pad bytes [nblocks, H) are written as 0
        if (DIM == 1 || blk_idx < geo.H) out[blk_idx] = (uint8_t)byte;
    }

    const TileOffsets to = tile_scan<WAVES>(sa, tile, wtot, wave, lane, s_wtot, &s_excl);
    if (tile == geo.ntiles - 1 && threadIdx.x == 0 && sa.total) *sa.total = geo.H + to.tile_incl;
    uint8_t* const wbase = out + geo.H + to.wave_base;
    if (geo.H + to.wave_base + wtot > cap) {   // never write past the caller's buffer
        if (lane == 0) atomicOr(sa.err, kErrOverflow);
        continue;
    }
    if constexpr (MODE == kOutlier) {
        if (blk && v.sel) store_le(wbase + off, bko, v.nb);   // little-endian |d| of element 0
    }

    // ---- pass 2: stream words ----------------------------------------------
    static_for<0, G>([&](auto gc) {
        constexpr int g = decltype(gc)::value;
        const Meta m(s_meta[wave][BPG * g + (DIM == 1 ? h : 0)]);
        const bool wideg = (DIM != 1) && ((big >> g) & 1);
        uint32_t w = W[g];
        if constexpr (MODE == kOutlier) {
            if (m.sel && l != 31) w &= not_e0;   // element 0 travels in the prefix, not the planes
        }
        const bool st = m.nz && (l < m.R || (l == 31 && !wideg));
        if (st) store_u32u(wbase + m.off + m.head + word_offset<DIM>(l, h, !wideg), w);
        if constexpr (DIM != 1) {
            if (wideg && m.nz && lane < 2) store_u32u(wbase + m.off + m.head + 4 * lane, s_sgn[wave][g][lane]);
        }
    });
    }   // tiles
}

// ---------------------------------------------------------------------------
// decompress
// ---------------------------------------------------------------------------
template <int DIM, int MODE, class T, int G, int WAVES>
__global__ void __launch_bounds__(WAVES * kWaveSize)
decompress_kernel(T* __restrict__ out, const uint8_t* __restrict__ cmp, const Geom geo, const T eb,
                  const ScanArgs sa, const uint64_t cmp_size) {
    constexpr int S = (DIM == 1) ? 4 : 8;
    constexpr int BPG = (DIM == 1) ? 2 : 1;
    constexpr int NB = BPG * G;
    __shared__ uint64_t s_excl;
    __shared__ uint64_t s_wtot[WAVES];
    __shared__ uint32_t s_meta[WAVES][NB];
    __shared__ uint32_t s_o[WAVES][NB];

    const int lane = (int)__lane_id();
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const uint32_t l = (uint32_t)lane & 31, h = (uint32_t)lane >> 5;
    LaneGeo<DIM> lg;
    lg.init(lane, geo);
    XposeConsts xc;
    xc.init(lane);

    for (uint64_t tile = blockIdx.x; tile < geo.ntiles; tile += gridDim.x) {
    const uint64_t grp0 = (tile * WAVES + (uint64_t)wave) * G;

    // ---- bookkeeping: lane j <-> block j of this wave -----------------------
    const uint64_t blk_idx = (DIM == 1) ? 2 * grp0 + (uint64_t)lane : grp0 + (uint64_t)lane;
    const bool blk = (lane < NB) && (DIM == 1 || blk_idx < geo.nblocks);
    const uint32_t byte = blk ? (uint32_t)cmp[blk_idx] : 0u;
    const BlockView v = view_of<MODE, S>(byte);
    const uint32_t sz = blk ? v.size : 0u;
    const uint32_t incl = scan_wave(sz);
    const uint32_t off = incl - sz;
    const uint64_t wtot = rdlane(incl, kWaveSize - 1);
    if (blk) s_meta[wave][lane] = pack_meta(off, v.R, sz != 0, v.sel, v.head);

    const TileOffsets to = tile_scan<WAVES>(sa, tile, wtot, wave, lane, s_wtot, &s_excl);
    if (tile == geo.ntiles - 1 && threadIdx.x == 0 && sa.total) *sa.total = geo.H + to.tile_incl;
    const uint8_t* const wbase = cmp + geo.H + to.wave_base;
    const bool intact = geo.H + to.wave_base + wtot <= cmp_size;
    if (!intact && lane == 0) atomicOr(sa.err, kErrTruncated);
    if constexpr (MODE == kOutlier) {
        // a selected block starts with nb <= 4 prefix bytes and then a 4-byte sign
        // word, so one u32 read at the block start stays inside the block
        uint32_t o = 0;
        if (blk && intact && v.sel) {
            o = load_u32u(wbase + off);
            if (v.nb < 4) o &= (1u << (8 * v.nb)) - 1;
        }
        if (lane < NB) s_o[wave][lane] = o;
        __builtin_amdgcn_wave_barrier();
    }

    // ---- prefetch the stream words of all G groups ---------------------------
    uint32_t W[G];
#pragma unroll
    for (int g = 0; g < G; ++g) {
        const Meta m(s_meta[wave][BPG * g + (DIM == 1 ? h : 0)]);
        const bool wideg = (DIM != 1) && (MODE != kOutlier) && m.R >= 32;
        const bool ld = intact && m.nz && (l < m.R || (l == 31 && !wideg));
        W[g] = ld ? load_u32u(wbase + m.off + m.head + word_offset<DIM>(l, h, !wideg)) : 0u;
    }

    // ---- decode ---------------------------------------------------------------
    Walker wk;
    if constexpr (DIM != 1) wk.init(grp0, geo);
#pragma unroll
    for (int g = 0; g < G; ++g) {
        const Meta m(s_meta[wave][BPG * g + (DIM == 1 ? h : 0)]);
        const bool wideg = (DIM != 1) && (MODE != kOutlier) && m.R >= 32;
        const uint32_t x = xpose32(W[g], xc);   // lane l: sign in bit 31 over the magnitude bits
        uint32_t a, neg;
        if (!wideg) {
            neg = x >> 31;
            a = x & 0x7fffffffu;
        } else {   // 32 planes: the sign words sit in their own two words
            const uint32_t sw = (intact && m.nz) ? load_u32u(wbase + m.off + 4 * h) : 0u;
            neg = (sw >> l) & 1;
            a = x;
        }
        if constexpr (MODE == kOutlier) {
            if (m.sel && l == 7 && (DIM == 1 || h == 0)) a = s_o[wave][BPG * g + (DIM == 1 ? h : 0)];   // element 0
        }
        const uint32_t dm = neg ? 0u - a : a;
        const uint32_t d = dpp<kDppHalfMirror>(dm);   // back to natural element order
        uint32_t q;
        if constexpr (MODE == kFixed) q = d;
        else q = inverse_lorenzo<DIM>(d, lane);
        const bool zero_blk = !intact || !m.nz;
        const T val = zero_blk ? T(0) : (T)(int32_t)q * eb * T(2);   // src/cuSZp_kernels_1D_f32.cu:490
        uint64_t idx;
        const bool valid = elem_of<DIM>(geo, lg, wk, grp0 + g, lane, idx);
        if (valid) out[idx] = val;
        if constexpr (DIM != 1) wk.next(geo);
    }
    }   // tiles
}

}  // namespace cuszp
