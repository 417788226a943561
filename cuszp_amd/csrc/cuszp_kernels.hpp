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
//     tile is WAVES*G groups. The grid is persistent: workgroup w handles
//     tiles w, w+B, w+2B, ... and prefetches its next tile while it scans and
//     stores the current one. No ticket atomics: one hot device-scope atomic
//     serialises at about 12 ns per operation on MI355X;
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
//     {epoch, flag, value}, is read 256 predecessors at a time by one wave, and
//     needs only relaxed agent-scope atomics (no fences: the value travels in
//     the same word as its flag). Forward progress does not rely on every
//     workgroup being resident: a wave that has waited too long for a
//     predecessor claims it and recomputes that tile's byte count itself;
//   * the batched kernels run the same per-variant loops over a table of
//     arrays; each array's tiles form their own scan segment.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

namespace cuszp {

constexpr int kWaveSize = 64;

enum : int { kFixed = 0, kPlain = 1, kOutlier = 2 };

// Tile status word: [63:40] epoch tag, [39:38] flag, [37:0] byte count.
// Flags: 1 = aggregate of the tile, 2 = inclusive prefix, 3 = claimed by a
// waiting wave that is computing the aggregate; 0 = nothing yet.
constexpr int kEpochShift = 40;
constexpr uint64_t kEpochMask = (1ull << 24) - 1;
constexpr uint64_t kFlagAgg = 1ull << 38;
constexpr uint64_t kFlagPrefix = 2ull << 38;
constexpr uint64_t kFlagClaim = 3ull << 38;
constexpr uint64_t kFlagMask = 3ull << 38;
constexpr uint64_t kValueMask = (1ull << 38) - 1;
constexpr uint64_t kTagMask = ~(kFlagMask | kValueMask);

// Bits of the sticky error words.
constexpr uint32_t kErrOverflow = 1u;    // compressed stream exceeds output capacity
constexpr uint32_t kErrTruncated = 2u;   // stream shorter than its header describes
constexpr uint32_t kErrStall = 4u;       // a lookback gave up waiting (internal error; output invalid)
constexpr uint32_t kInfoStolen = 8u;     // a waiting wave computed a predecessor's byte count (not an error)
constexpr uint32_t kErrMask = 7u;

// Bound on lookback polls (each polls the status words once): far above any
// legitimate wait, it guarantees that every wave of a launch exits even if a
// predecessor tile never publishes.
constexpr uint32_t kSpinLimit = 1u << 18;
// A wait longer than this (s_memrealtime ticks: 100 us at the 100 MHz
// reference clock) makes the waiting wave compute pending predecessors itself.
constexpr uint64_t kStealTicks = 10000;

struct Geom {
    uint64_t n;         // elements
    uint64_t nblocks;   // cuSZp blocks
    uint64_t H;         // header bytes
    uint64_t ntiles;    // tiles of the array
    uint64_t XY;        // X*Y
    uint32_t X, Y, Z;
    uint32_t Bx, By;    // blocks along x and y (2D/3D)
    uint32_t aligned;   // 2D/3D: every block lies entirely inside the array (extents multiples of the block)
};

// Launch-wide scan state.
struct ScanCtx {
    unsigned long long* status;     // tile status words, one per tile of the launch
    unsigned int* err;              // sticky error bits of the launch (kErrStall, kInfoStolen)
    unsigned long long* diag;       // [0] first stalled tile + 1, [1] its missing predecessor, [2] word seen
    unsigned long long tag;         // epoch << kEpochShift
};

// One array of a launch. The single-array kernels take it as an argument; the
// batched kernels read a table of them.
struct Item {
    const void* src;                // compress: values; decompress: stream
    void* dst;                      // compress: stream; decompress: values
    uint64_t limit;                 // compress: stream capacity; decompress: stream length
    unsigned long long* total;      // out: stream length, written by the last tile (may be null)
    unsigned int* err;              // sticky error bits of this array (kErrOverflow, kErrTruncated)
    double param;                   // compress: 0.5/eb, decompress: eb (exact in the array's type)
    uint64_t tile0;                 // first tile of the array in the launch (its scan segment start)
    uint32_t variant;               // type * 9 + (dim - 1) * 3 + mode
    uint32_t pad_;
    Geom geo;                       // geo.ntiles: tiles of this array
};

// ---------------------------------------------------------------------------
// wave64 primitives
// ---------------------------------------------------------------------------
// Lanes whose source is outside the row/wave read 0. With every row and bank
// enabled that is bound_ctrl (no old value to materialise); masked rows/banks
// keep the old value 0.
template <int CTRL, int ROWS = 0xf, int BANKS = 0xf>
__device__ __forceinline__ uint32_t dpp(uint32_t v) {
    if constexpr (ROWS == 0xf && BANKS == 0xf)
        return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, CTRL, ROWS, BANKS, true);
    else
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
// A value the compiler cannot prove wave-uniform but that is (kept in SGPRs,
// so that uniform-base + 32-bit-offset addressing applies).
__device__ __forceinline__ uint64_t uniform64(uint64_t v) {
    return (uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)v) |
           ((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(v >> 32)) << 32);
}

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
// (m & a) | (~m & b) in one v_bfi_b32 (the compiler otherwise emits and + and_or)
__device__ __forceinline__ uint32_t bfi(uint32_t m, uint32_t a, uint32_t b) {
    uint32_t r;
    asm("v_bfi_b32 %0, %1, %2, %3" : "=v"(r) : "v"(m), "v"(a), "v"(b));
    return r;
}
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
// decoupled lookback (single pass, wave-parallel, segmented, stealing)
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint64_t ld_status(const unsigned long long* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_status(unsigned long long* p, uint64_t v) {
    __hip_atomic_store(p, (unsigned long long)v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint64_t cas_status(unsigned long long* p, uint64_t expect, uint64_t desired) {
    unsigned long long e = expect;
    __hip_atomic_compare_exchange_strong(p, &e, (unsigned long long)desired, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                         __HIP_MEMORY_SCOPE_AGENT);
    return e;   // value seen (== expect on success)
}
__device__ __forceinline__ uint64_t rdlane64(uint64_t v, uint32_t l) {
    return (uint64_t)rdlane((uint32_t)v, l) | ((uint64_t)rdlane((uint32_t)(v >> 32), l) << 32);
}

// Sum over the wave of per-lane values below 2^40: two 20-bit halves are
// reduced with DPP adds (no LDS round trips) and recombined.
__device__ __forceinline__ uint64_t wave_sum_u40(uint64_t v) {
    const uint32_t lo = scan_wave((uint32_t)v & 0xfffffu);
    const uint32_t hi = scan_wave((uint32_t)(v >> 20));
    return (uint64_t)rdlane(lo, kWaveSize - 1) + ((uint64_t)rdlane(hi, kWaveSize - 1) << 20);
}

__device__ __forceinline__ bool status_ready(uint64_t w, uint64_t tag) {
    const uint64_t f = w & kFlagMask;
    return ((w & kTagMask) == tag) && f != 0 && f != kFlagClaim;
}

// Lookback state shared by the waves of a workgroup.
template <int WAVES>
struct LookbackShared {
    uint32_t pos[2][WAVES][2];   // per poll round (double-buffered): nearest prefix, nearest pending position
    uint32_t ctl[2];             // wave 0's decisions per poll round: bit 0 steal, bit 1 give up
    uint64_t part[WAVES];        // per-wave sums of a resolved window
};

// Exclusive prefix of `tile` within its segment (tiles first .. tile-1), run by
// every wave of the workgroup (uniform control flow across the workgroup).
// `published`: the caller already published the tile's aggregate (or, for the
// segment's first tile, its inclusive prefix) with publish_tile().
//
// The workgroup publishes the tile aggregate, then reads predecessors a window
// of WAVES * 64 * K at a time: wave v, lane j, word k reads position
// p = 64*(K*v + k) + j, i.e. tile - 1 - p. A window is resolved once every
// position up to the nearest published inclusive prefix P is ready; the sum of
// the values at positions <= P is the exclusive prefix (or, with no prefix in
// the window, its sum carries into the next window). The inclusive prefix is
// then published. In a persistent grid of B workgroups the nearest prefix
// trails a tile by up to about B positions, so the window is sized to cover B
// in one round trip of status loads.
//
// Forward progress does not depend on which workgroups are resident: when a
// window stays unresolved for kStealTicks, each wave claims its farthest
// pending position (CAS from the stale word to CLAIM), computes that tile's
// aggregate with agg_of(tile) and publishes it. A claim is held only by a
// running wave that computes without waiting, so every wait ends; the owner's
// later store of the same aggregate (or of its prefix) is harmless.
template <int WAVES, int K, class AggF>
__device__ __forceinline__ uint64_t wg_lookback(const ScanCtx& sc, uint64_t tile, uint64_t first, uint64_t agg,
                                                int wave, int lane, LookbackShared<WAVES>& ls, AggF&& agg_of,
                                                bool published = false, const uint64_t (*pre)[K] = nullptr) {
    constexpr int SPAN = 64 * K;           // positions per wave and window
    constexpr int WIN = WAVES * SPAN;      // positions per window
    const uint64_t tag = sc.tag;
    const bool leader = (wave == 0) && (lane == 0);
    if (tile == first) {
        if (leader && !published) st_status(sc.status + tile, tag | kFlagPrefix | agg);
        return 0;
    }
    if (leader && !published) st_status(sc.status + tile, tag | kFlagAgg | agg);
    const int pbase = wave * SPAN;
    uint64_t excl = 0;
    int64_t hi = (int64_t)tile;
    uint32_t spins = 0, round = 0;
    const uint64_t t_wait = __builtin_amdgcn_s_memrealtime();
    bool steal = false, stalled = false;
    bool use_pre = pre != nullptr;
    for (;;) {
        uint64_t w[K];
        bool ready[K];
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const int64_t idx = hi - 1 - (pbase + 64 * k + lane);
            w[k] = tag | kFlagPrefix;   // before the segment: inclusive prefix 0
            ready[k] = true;
            if (idx >= (int64_t)first) {
                w[k] = use_pre ? (*pre)[k] : ld_status(sc.status + idx);
                ready[k] = status_ready(w[k], tag);
            }
        }
        use_pre = false;   // only the first window comes pre-loaded
        int P;
        for (;;) {
            int Pv = WIN, Nv = WIN;
#pragma unroll
            for (int k = K - 1; k >= 0; --k) {
                const uint64_t pm = ballot(ready[k] && (w[k] & kFlagMask) == kFlagPrefix);
                if (pm) Pv = pbase + 64 * k + __builtin_ctzll(pm);
                const uint64_t nr = ballot(!ready[k]);
                if (nr) Nv = pbase + 64 * k + __builtin_ctzll(nr);
            }
            const uint32_t par = round & 1;
            if (lane == 0) {
                ls.pos[par][wave][0] = (uint32_t)Pv;
                ls.pos[par][wave][1] = (uint32_t)Nv;
                if (wave == 0) {
                    uint32_t c = 0;
                    if (__builtin_amdgcn_s_memrealtime() - t_wait > kStealTicks) c |= 1;
                    if (spins + 1 >= kSpinLimit ||
                        (((spins + 1) & 1023) == 0 &&
                         (__hip_atomic_load(sc.err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & kErrStall)))
                        c |= 2;
                    ls.ctl[par] = c;
                }
            }
            __syncthreads();
            P = WIN;
            int N = WIN;
#pragma unroll
            for (int v = 0; v < WAVES; ++v) {
                P = min(P, (int)ls.pos[par][v][0]);
                N = min(N, (int)ls.pos[par][v][1]);
            }
            const uint32_t ctl = ls.ctl[par];
            ++round;
            ++spins;
            if (N > P || N == WIN) break;   // every position up to P is ready
            // nearest pending position of this wave at or below P
            int sk = -1, sl = 0;
#pragma unroll
            for (int k = 0; k < K; ++k) {
                uint64_t nr = ballot(!ready[k]);
                const int lo = pbase + 64 * k;
                if (P < lo) nr = 0;
                else if (P < lo + 63) nr &= (2ull << (P - lo)) - 1;
                if (nr) { sk = k; sl = 63 - __builtin_clzll(nr); }
            }
            uint64_t pend = 0;
#pragma unroll
            for (int k = 0; k < K; ++k) if (k == sk) pend = rdlane64(w[k], sl);
            const uint64_t pidx = (uint64_t)(hi - 1 - (pbase + 64 * sk + sl));
            if (ctl & 2) {
                // last-resort watchdog: record the stuck predecessor and give up
                // (the output is reported invalid; publishing below unblocks successors)
                stalled = true;
                if (lane == 0 && sk >= 0) {
                    atomicOr(sc.err, kErrStall);
                    if (atomicCAS(sc.diag, 0ull, tile + 1) == 0ull) {
                        sc.diag[1] = pidx;
                        sc.diag[2] = pend;
                    }
                }
                break;
            }
            steal |= (ctl & 1) != 0;
            bool took = false;
            if (steal && sk >= 0 && !(((pend & kTagMask) == tag) && ((pend & kFlagMask) == kFlagClaim))) {
                uint64_t seen = 0;
                if (lane == 0) seen = cas_status(sc.status + pidx, pend, tag | kFlagClaim);
                seen = rdlane64(seen, 0);
                if (seen == pend) {
                    const uint64_t a = agg_of(pidx);
                    if (lane == 0) {
                        cas_status(sc.status + pidx, tag | kFlagClaim, tag | kFlagAgg | a);
                        atomicOr(sc.err, kInfoStolen);
                    }
                    took = true;
                }
            }
            if (!took) __builtin_amdgcn_s_sleep(1);
#pragma unroll
            for (int k = 0; k < K; ++k) {
                const int64_t idx = hi - 1 - (pbase + 64 * k + lane);
                if (!ready[k]) {
                    w[k] = ld_status(sc.status + idx);
                    ready[k] = status_ready(w[k], tag);
                }
            }
        }
        uint64_t part = 0;
#pragma unroll
        for (int k = 0; k < K; ++k) part += (pbase + 64 * k + lane <= P) ? (w[k] & kValueMask) : 0;
        part = wave_sum_u40(part);
        if (lane == 0) ls.part[wave] = part;
        __syncthreads();
#pragma unroll
        for (int v = 0; v < WAVES; ++v) excl += ls.part[v];
        if (P < WIN || stalled) break;
        hi -= WIN;
    }
    if (leader) st_status(sc.status + tile, tag | kFlagPrefix | (excl + agg));
    return excl;
}

// Early publication of a tile's aggregate (the segment's first tile publishes
// its inclusive prefix): done as soon as its byte count is known.
__device__ __forceinline__ void publish_tile(const ScanCtx& sc, uint64_t tile, uint64_t first, uint64_t agg) {
    st_status(sc.status + tile, sc.tag | (tile == first ? kFlagPrefix : kFlagAgg) | agg);
}

// Loads of the first lookback window of `tile` (the words wg_lookback reads
// at positions p = 64*(K*wave + k) + lane), issued ahead of time: in a
// pipelined loop they are issued before the next tile's data loads, so waiting
// for them does not wait for that data. Positions before the segment load the
// segment's first word instead (wg_lookback replaces them).
template <int WAVES, int K>
__device__ __forceinline__ void lookback_issue(const ScanCtx& sc, uint64_t tile, uint64_t first, int wave, int lane,
                                               uint64_t (&w)[K]) {
#pragma unroll
    for (int k = 0; k < K; ++k) {
        int64_t idx = (int64_t)tile - 1 - (wave * 64 * K + 64 * k + lane);
        if (idx < (int64_t)first) idx = (int64_t)first;
        w[k] = ld_status(sc.status + idx);
    }
}

// ---------------------------------------------------------------------------
// Stores of the hot loops, written as inline asm: the compiler's s_waitcnt
// insertion does not see them, so its waits for the loop's own loads never
// wait for these stores to complete (gfx9 counts stores in vmcnt and retires
// memory operations in order). The loops never read back what they store.
// Base addresses are wave-uniform (SGPRs), offsets 32-bit per lane.
//
// Each starts with s_nop 4: a VMEM instruction that reads an SGPR written by a
// VALU instruction (here the v_readfirstlane that makes the base uniform) needs
// 5 wait states on gfx9, and the compiler's hazard recognizer cannot see that
// an inline asm block is a VMEM instruction.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void st_async_b32(const void* base, uint32_t off, uint32_t v) {
    asm volatile("s_nop 4\n\tglobal_store_dword %0, %1, %2" : : "v"(off), "v"(v), "s"(uniform64((uint64_t)base)) : "memory");
}
__device__ __forceinline__ void st_async_b16(const void* base, uint32_t off, uint32_t v) {
    asm volatile("s_nop 4\n\tglobal_store_short %0, %1, %2" : : "v"(off), "v"(v), "s"(uniform64((uint64_t)base)) : "memory");
}
__device__ __forceinline__ void st_async_b8(const void* base, uint32_t off, uint32_t v) {
    asm volatile("s_nop 4\n\tglobal_store_byte %0, %1, %2" : : "v"(off), "v"(v), "s"(uniform64((uint64_t)base)) : "memory");
}
// little-endian 1..4 bytes
__device__ __forceinline__ void st_async_le(const void* base, uint32_t off, uint32_t v, uint32_t nb) {
    if (nb == 4) {
        st_async_b32(base, off, v);
    } else {
        if (nb & 2) { st_async_b16(base, off, v); off += 2; v >>= 16; }
        if (nb & 1) st_async_b8(base, off, v);
    }
}
// status word publication (relaxed, agent scope: sc1 like __hip_atomic_store)
__device__ __forceinline__ void st_status_async(const unsigned long long* status, uint32_t tile, uint64_t v) {
    asm volatile("s_nop 4\n\tglobal_store_dwordx2 %0, %1, %2 sc1" : : "v"(tile * 8u), "v"(v), "s"(uniform64((uint64_t)status)) : "memory");
}

// Workgroup-level: combine the wave totals and run the lookback; returns
// {wave's exclusive offset within the stream payload, tile inclusive}.
struct TileOffsets {
    uint64_t wave_base;   // payload offset (excluding H) of this wave's first block
    uint64_t tile_incl;   // payload bytes up to and including this tile
};
template <int WAVES, int K, class AggF>
__device__ __forceinline__ TileOffsets tile_scan(const ScanCtx& sc, uint64_t tile, uint64_t first, uint64_t wtot,
                                                 int wave, int lane, uint64_t* s_wtot, LookbackShared<WAVES>& ls,
                                                 AggF&& agg_of) {
    if (lane == 0) s_wtot[wave] = wtot;
    __syncthreads();
    uint64_t wexcl = 0, ttot = 0;
#pragma unroll
    for (int w = 0; w < WAVES; ++w) {
        const uint64_t v = s_wtot[w];
        wexcl += (w < wave) ? v : 0;
        ttot += v;
    }
    const uint64_t excl = wg_lookback<WAVES, K>(sc, tile, first, ttot, wave, lane, ls, agg_of);
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

// Element index of the first element of the walker's block (2D/3D).
template <int DIM>
__device__ __forceinline__ uint64_t block_origin(const Geom& g, const Walker& wk) {
    if constexpr (DIM == 2) return ((uint64_t)wk.bz * g.Y + (uint64_t)wk.by * 8) * g.X + (uint64_t)wk.bx * 8;
    else return ((uint64_t)wk.bz * 4 * g.Y + (uint64_t)wk.by * 4) * g.X + (uint64_t)wk.bx * 4;
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
// per-block metadata handed from the bookkeeping lanes to the stream-word
// lanes through LDS (one 8-byte entry per block):
//   a: [19:0] byte offset of the block's sign word from the wave's first
//      block (past any outlier prefix), [30] 32 planes (2D/3D: the sign words
//      move in front of the planes), [31] outlier-selected
//   b: bit l set when lane l of a 32-lane half carries one of the block's
//      stream words (planes 0..R-1 in lanes 0..R-1, the sign word in lane 31)
// ---------------------------------------------------------------------------
struct Meta {
    uint32_t a, b;
    __device__ __forceinline__ uint32_t off() const { return a & 0xfffffu; }
    __device__ __forceinline__ bool wide() const { return (a >> 30) & 1; }
    __device__ __forceinline__ bool sel() const { return (int32_t)a < 0; }
    __device__ __forceinline__ bool nz() const { return b != 0; }
    __device__ __forceinline__ bool has(uint32_t l) const { return (b >> l) & 1; }
};
__device__ __forceinline__ Meta make_meta(uint32_t off, const BlockView& v, bool nz) {
    Meta m;
    const bool wide = v.R >= 32;
    m.a = (off + v.head) | ((uint32_t)wide << 30) | (v.sel << 31);
    m.b = !nz ? 0u : wide ? ~0u : (((1u << v.R) - 1u) | 0x80000000u);
    return m;
}

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

// Item sources: where a persistent loop finds the array a tile belongs to.
// OneItem serves the single-array kernels (the descriptor is a kernel
// argument); ItemTable serves the batched kernels (descriptors in device
// memory, ordered by tile0).
struct OneItem {
    Item item;
    __device__ __forceinline__ uint32_t locate(uint64_t, uint32_t) const { return 0; }
    __device__ __forceinline__ Item get(uint32_t) const { return item; }
    __device__ __forceinline__ uint32_t variant(uint32_t) const { return item.variant; }
};
struct ItemTable {
    const Item* items;
    uint32_t count;
    // last item whose tile0 <= tile, searching from `from` (tiles only move forward)
    __device__ __forceinline__ uint32_t locate(uint64_t tile, uint32_t from) const {
        uint32_t lo = from, hi = count;   // answer in [lo, hi)
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (items[mid].tile0 <= tile) lo = mid;
            else hi = mid;
        }
        return lo;
    }
    __device__ __forceinline__ Item get(uint32_t i) const { return items[i]; }
    __device__ __forceinline__ uint32_t variant(uint32_t i) const { return items[i].variant; }
};

// ---------------------------------------------------------------------------
// Aggregates recomputed by a waiting wave (see tile_lookback).
// ---------------------------------------------------------------------------
// OR of v over each 32-lane half (every lane of the half gets the result).
__device__ __forceinline__ uint32_t or_half(uint32_t v) {
    v |= swizzle_xor<16>(v);
    v |= swizzle_xor<8>(v);
    v |= swizzle_xor<4>(v);
    v |= swizzle_xor<2>(v);
    v |= swizzle_xor<1>(v);
    return v;
}

// Payload bytes of local tile `lt` of a compress item, by one wave. Same block
// sizes as pass 1 of compress_run: a block's rate is the bit length of the OR of
// its magnitudes (the highest non-zero plane), 1D ranks |-2^31| as 0.
template <int DIM, int MODE, class T, int G, int WAVES>
__device__ uint64_t compress_tile_bytes(const T* __restrict__ in, const Geom& geo, T r, uint64_t lt, int lane) {
    constexpr int S = (DIM == 1) ? 4 : 8;
    LaneGeo<DIM> lg;
    lg.init(lane, geo);
    const int pl = pred_lane<DIM>(lane);
    const int pl_src = pl < 0 ? lane : pl;
    const bool e0 = (DIM == 1) ? ((lane & 31) == 0) : (lane == 0);
    const uint64_t g0 = lt * WAVES * G;
    Walker wk;
    if constexpr (DIM != 1) wk.init(g0, geo);
    uint64_t sum = 0;
    for (int i = 0; i < WAVES * G; ++i) {
        uint64_t idx;
        const bool valid = elem_of<DIM>(geo, lg, wk, g0 + i, lane, idx);
        const T x = valid ? in[idx] : T(0);
        const uint32_t q = quantize<T>(x, r);
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
        const uint32_t a = ((int32_t)d < 0) ? 0u - d : d;
        const uint32_t a1 = (DIM == 1) ? (a & 0x7fffffffu) : a;
        const uint32_t o1 = or_half(a1);
        const uint32_t o2 = (MODE == kOutlier) ? or_half(e0 ? 0u : a1) : 0u;
        if constexpr (DIM == 1) {
#pragma unroll
            for (int hh = 0; hh < 2; ++hh) {
                const uint32_t byte = rate_byte<MODE, S>(bitlen(rdlane(o1, 32 * hh)), bitlen(rdlane(o2, 32 * hh)),
                                                         bitlen(rdlane(a, 32 * hh)));
                sum += view_of<MODE, S>(byte).size;
            }
        } else {
            const uint32_t byte = rate_byte<MODE, S>(bitlen(rdlane(o1, 0) | rdlane(o1, 32)),
                                                     bitlen(rdlane(o2, 0) | rdlane(o2, 32)), bitlen(rdlane(a, 0)));
            sum += view_of<MODE, S>(byte).size;
            wk.next(geo);
        }
    }
    return sum;
}

// Payload bytes of local tile `lt` of a decompress item, from its header bytes.
template <int DIM, int MODE, int G, int WAVES>
__device__ uint64_t header_tile_bytes(const uint8_t* __restrict__ cmp, const Geom& geo, uint64_t lt, int lane) {
    constexpr int S = (DIM == 1) ? 4 : 8;
    constexpr int BPT = WAVES * G * ((DIM == 1) ? 2 : 1);   // blocks per tile
    const uint64_t limit = (DIM == 1) ? geo.H : geo.nblocks;
    uint64_t sum = 0;
#pragma unroll
    for (int c = 0; c < (BPT + kWaveSize - 1) / kWaveSize; ++c) {
        const uint64_t b = lt * BPT + (uint64_t)(c * kWaveSize + lane);
        const bool in_tile = c * kWaveSize + lane < BPT;
        const uint32_t byte = (in_tile && b < limit) ? (uint32_t)cmp[b] : 0u;
        sum += wave_sum_u40(view_of<MODE, S>(byte).size);
    }
    return sum;
}

// ---------------------------------------------------------------------------
// compress: persistent loop over the tiles of consecutive items of one variant
// ---------------------------------------------------------------------------
// Shared memory of a compress workgroup, sized for the largest variant. Three
// tiles are in flight per workgroup (see compress_run); everything a tile needs
// between its pass 1 and its stores lives here, indexed by tile slot.
template <int WAVES, int GMAX>
struct SmemC {
    LookbackShared<WAVES> lb;
    uint64_t wtot[3][WAVES];                    // per-wave byte counts
    Meta meta[3][WAVES][2 * GMAX];              // per-block metadata
    uint32_t aux[3][WAVES][2 * GMAX];           // sign words of 32-plane 2D/3D groups
    uint32_t words[3][WAVES][GMAX][kWaveSize];  // stream words
    uint32_t lane[3][WAVES][2 * GMAX][2];       // per block: {off | prefix bytes << 20 | header byte << 24, |d0|}
    uint32_t wave[3][WAVES][4];                 // {wave exclusive offset in the tile, wave bytes, 32-plane groups, -}
};

// Processes tiles tile, tile + gridDim.x, ... while they belong to items of this
// variant; returns with `tile` at the first tile it did not process and `it` at
// that tile's item (or tile >= ntiles).
//
// Two-deep software pipeline built around gfx9's in-order vmcnt (stores and
// loads share one counter, so a wait for any load also waits for every older
// store): iteration i
//   1. resolves the lookback of t_{i-2} (its first window was loaded at the
//      start of iteration i-1) and writes t_{i-2}'s header bytes and stream
//      words from LDS (stores are inline asm the compiler's waits do not see);
//   2. loads the first lookback window of t_{i-1} (whose byte count and those
//      of its round were published at the end of iteration i-1);
//   3. runs pass 1 of t_i on data loaded during iteration i-1 and, group by
//      group, loads t_{i+1} into the same registers;
//   4. publishes t_i's byte count and parks its words and metadata in LDS.
// Every wait of iteration i is for loads issued about one pass 1 earlier, so
// no memory latency is exposed while the lookback has a whole iteration to
// see its predecessors. The last tiles of an item (per workgroup) drain the
// pipeline in extra iterations without a pass 1.
template <int DIM, int MODE, class T, int G, int WAVES, int GMAX, class Src>
__device__ __forceinline__ void compress_run(const Src& src, uint32_t& it, uint64_t& tile, const uint64_t ntiles,
                                             const ScanCtx& sc, SmemC<WAVES, GMAX>& sm) {
    constexpr int S = (DIM == 1) ? 4 : 8;        // bytes per plane
    constexpr int BPG = (DIM == 1) ? 2 : 1;      // blocks per group
    constexpr int NB = BPG * G;                  // blocks per wave
    constexpr int KLB = (16 / WAVES > 1) ? 16 / WAVES : 1;   // lookback words per lane
    static_assert(NB <= kWaveSize, "one bookkeeping lane per block");
    static_assert(G <= GMAX, "shared memory sized for the largest variant");

    const int lane = (int)__lane_id();
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const uint32_t l = (uint32_t)lane & 31, h = (uint32_t)lane >> 5;
    const int pl = pred_lane<DIM>(lane);
    const int pl_src = pl < 0 ? lane : pl;
    XposeConsts xc;
    xc.init(lane);
    const uint32_t wo = word_offset<DIM>(l, h, true);        // this lane's word in a block's payload
    const uint32_t wo_wide = word_offset<DIM>(l, h, false);  // same, 32-plane 2D/3D layout
    // after the transpose, element 0 of a block is bit 7 of its plane words: in
    // both halves for 1D (two blocks), only in the low half for 2D/3D
    const uint32_t not_e0 = (DIM == 1 || h == 0) ? ~0x80u : ~0u;
    const uint32_t stride = gridDim.x;

    Item cur = src.get(it);
    LaneGeo<DIM> lg;
    lg.init(lane, cur.geo);

    T xv[G];
    uint64_t vmask = 0;   // bit g: this lane's element of group g lies inside the array
    auto load_tile = [&](uint64_t lt) {
        const Geom& geo = cur.geo;
        const T* __restrict__ in = (const T*)cur.src;
        const uint64_t g0 = (lt * WAVES + (uint64_t)wave) * G;
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

    // resolve the lookback of tile t (slot s, ttot bytes) and write its stream (every wave)
    uint64_t lbw[KLB];
    auto finish = [&](uint32_t s, uint64_t t, uint64_t ttot) {
        const Geom& geo = cur.geo;
        const T r = (T)cur.param;
        const uint64_t excl = wg_lookback<WAVES, KLB>(sc, t, cur.tile0, ttot, wave, lane, sm.lb,
            [&](uint64_t p) { return compress_tile_bytes<DIM, MODE, T, G, WAVES>((const T*)cur.src, geo, r, p - cur.tile0, lane); },
            true, &lbw);
        if (t - cur.tile0 == geo.ntiles - 1 && threadIdx.x == 0 && cur.total) *cur.total = geo.H + excl + ttot;
        const uint32_t wexcl = sm.wave[s][wave][0], wbytes = sm.wave[s][wave][1], big = sm.wave[s][wave][2];
        const uint32_t ls0 = lane < NB ? sm.lane[s][wave][lane][0] : 0u;   // bookkeeping lanes only
        // header bytes (lane j: block j of the wave); 2D/3D pad bytes [nblocks, H) are 0
        const uint64_t grp0 = ((uint64_t)(t - cur.tile0) * WAVES + (uint64_t)wave) * G;
        const uint64_t b0 = (DIM == 1) ? 2 * grp0 : grp0;
        if (lane < NB && (DIM == 1 || b0 + (uint64_t)lane < geo.H))
            st_async_b8((uint8_t*)cur.dst + b0, (uint32_t)lane, ls0 >> 24);
        const uint64_t wave_base = excl + wexcl;
        uint8_t* const wbase = (uint8_t*)cur.dst + uniform64(geo.H + wave_base);
        const bool fits = geo.H + wave_base + wbytes <= cur.limit;   // never write past the caller's buffer
        if (!fits && lane == 0) atomicOr(cur.err, kErrOverflow);
        if (fits) {
            if constexpr (MODE == kOutlier) {
                const uint32_t pref = (ls0 >> 20) & 7;
                if (lane < NB && pref) st_async_le(wbase, ls0 & 0xfffffu, sm.lane[s][wave][lane][1], pref);
            }
            static_for<0, G>([&](auto gc) {
                constexpr int g = decltype(gc)::value;
                const Meta m = sm.meta[s][wave][BPG * g + (DIM == 1 ? h : 0)];
                const bool wideg = (DIM != 1) && ((big >> g) & 1);
                uint32_t w = sm.words[s][wave][g][lane];
                if constexpr (MODE == kOutlier) {
                    if (m.sel() && l != 31) w &= not_e0;   // element 0 travels in the prefix, not the planes
                }
                if (m.has(l)) st_async_b32(wbase, m.off() + (wideg ? wo_wide : wo), w);
                if constexpr (DIM != 1) {
                    if (wideg && m.nz() && lane < 2)
                        st_async_b32(wbase, m.off() + 4 * lane, sm.aux[s][wave][2 * g + lane]);
                }
            });
        }
    };

    for (;;) {   // items of this variant
    load_tile(tile - cur.tile0);
    // pipeline: npend tiles (0..2) wait for their stores; slot is where the
    // tile entering pass 1 parks; the pending ones are tile - stride (slot - 1)
    // and tile - 2 * stride (slot - 2), with byte counts ttot1 and ttot2
    uint32_t npend = 0, slot = 0;
    uint64_t ttot1 = 0, ttot2 = 0;
    for (;;) {
    const Geom& geo = cur.geo;
    const uint32_t s2 = (slot + 1) % 3;   // t_{i-2}
    const uint32_t s1 = (slot + 2) % 3;   // t_{i-1}
    // 1. oldest pending tile: lookback + stores
    if (npend == 2) finish(s2, tile - 2 * (uint64_t)stride, ttot2);
    // 2. first lookback window of the younger pending tile
    if (npend >= 1) lookback_issue<WAVES, KLB>(sc, tile - stride, cur.tile0, wave, lane, lbw);

    // 3. pass 1 of t_i
    const uint64_t lt = tile - cur.tile0;
    const T r = (T)cur.param;
    const uint64_t grp0 = (lt * WAVES + (uint64_t)wave) * G;
    const uint64_t next = tile + stride;
    const bool same = next < cur.tile0 + geo.ntiles;   // next tile in this item
    const T* __restrict__ in = (const T*)cur.src;
    const uint64_t gn0 = ((next - cur.tile0) * WAVES + (uint64_t)wave) * G;
    // The next tile's groups load during pass 1 when all of this wave's groups
    // of it lie inside the array (then no per-element checks are needed);
    // otherwise the tile loads after pass 1 with checks. The loads are issued
    // on every iteration (a fixed number per iteration lets the compiler wait
    // for exactly the loads it needs); without a next tile to roll in, every
    // lane re-reads element 0.
    bool roll = same;
    uint64_t nbase = 0;
    if constexpr (DIM == 1) {
        roll = roll && (gn0 + G) * 64 <= geo.n;
        nbase = gn0 * 64;
    } else {
        roll = roll && geo.aligned && gn0 + G <= geo.nblocks;
    }
    Walker wkn;
    if constexpr (DIM != 1) {
        if (roll) wkn.init(gn0, geo);
    }
    const uint32_t lofs = roll ? (uint32_t)lane : 0u;

    uint32_t W[G];
    uint32_t bk1 = 0, bk2 = 0, bko = 0;   // lane j: block j's nonzero-plane mask (all / e >= 1), |d| of element 0
    uint32_t big = 0;                     // bit g: group g's block has a magnitude 2^31 (2D/3D)
    // Tiles entirely inside the array skip the per-element validity select
    // (fixed mode never needs it: x = 0 gives q = 0). 1D instantiates pass 1
    // twice for that; 2D/3D branch per group (two copies cost more registers).
    const bool full = ballot(vmask != (G == 64 ? ~0ull : (1ull << G) - 1)) == 0;
    auto pass1 = [&](auto fullc) {
    constexpr bool FULL = decltype(fullc)::value;
    static_for<0, G>([&](auto gc) {
        constexpr int g = decltype(gc)::value;
        const uint32_t q = quantize<T>(xv[g], r);   // x = 0 outside the array gives q = 0
        if constexpr (DIM == 1) {
            xv[g] = __builtin_nontemporal_load(in + (roll ? nbase + 64 * g : 0) + lofs);
        } else {
            xv[g] = __builtin_nontemporal_load(in + (roll ? block_origin<DIM>(geo, wkn) + lg.off : 0));
            if (roll) wkn.next(geo);
        }
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
        if constexpr (MODE != kFixed && !FULL) {
            if (DIM == 1 || !full) d = ((vmask >> g) & 1) ? d : 0u;
        }
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
                sm.aux[slot][wave][2 * g] = (uint32_t)sg;
                sm.aux[slot][wave][2 * g + 1] = (uint32_t)(sg >> 32);
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
    };
    if (DIM == 1 && full) pass1(std::true_type{});
    else pass1(std::false_type{});
    if (roll) vmask = (G == 64) ? ~0ull : (1ull << G) - 1;
    else if (same) load_tile(next - cur.tile0);   // a partial next tile: loads with checks

    // ---- bookkeeping: lane j <-> block j of this wave -----------------------
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
        sm.meta[slot][wave][lane] = make_meta(off, v, sz != 0);
        sm.lane[slot][wave][lane][0] = off | (((blk && v.sel) ? v.nb : 0u) << 20) | (byte << 24);
        sm.lane[slot][wave][lane][1] = bko;
    }
#pragma unroll
    for (int g = 0; g < G; ++g) sm.words[slot][wave][g][lane] = W[g];
    // combine the wave totals and publish the tile's byte count right away
    if (lane == 0) sm.wtot[slot][wave] = wtot;
    __syncthreads();
    uint64_t wexcl = 0, ttot = 0;
#pragma unroll
    for (int w = 0; w < WAVES; ++w) {
        const uint64_t t = sm.wtot[slot][w];
        wexcl += (w < wave) ? t : 0;
        ttot += t;
    }
    if (wave == 0 && lane == 0)
        st_status_async(sc.status, (uint32_t)tile, sc.tag | (tile == cur.tile0 ? kFlagPrefix : kFlagAgg) | ttot);
    if (lane == 0) {   // read back by this wave only
        sm.wave[slot][wave][0] = (uint32_t)wexcl;
        sm.wave[slot][wave][1] = (uint32_t)wtot;
        sm.wave[slot][wave][2] = big;
    }
    npend = npend == 2 ? 2 : npend + 1;
    slot = (slot + 1) % 3;
    ttot2 = ttot1;
    ttot1 = ttot;

    tile = next;
    if (!same) break;
    }   // tiles of this item
    // drain: the last one or two tiles of this item for this workgroup
    {
        const uint64_t last = tile - stride;
        if (npend == 2) finish((slot + 1) % 3, last - stride, ttot2);
        lookback_issue<WAVES, KLB>(sc, last, cur.tile0, wave, lane, lbw);
        finish((slot + 2) % 3, last, ttot1);
    }
    if (tile >= ntiles) return;
    it = src.locate(tile, it);
    if (src.variant(it) != cur.variant) return;
    cur = src.get(it);
    lg.init(lane, cur.geo);
    }   // items
}


// ---------------------------------------------------------------------------
// decompress: persistent loop, same contract as compress_run
// ---------------------------------------------------------------------------
// Shared memory of a decompress workgroup, sized for the largest variant.
template <int WAVES, int GMAX>
struct SmemD {
    LookbackShared<WAVES> lb;
    uint64_t wtot[2][WAVES];
    Meta meta[2][WAVES][2 * GMAX];    // per-block metadata of the current and the pending tile
    uint32_t aux[WAVES][2 * GMAX];    // outlier |d| of element 0 per block of the tile being decoded
};

// Same one-deep pipeline as compress_run: iteration i turns tile t_i's header
// bytes into block sizes and publishes its byte count at once (its header
// bytes were loaded during iteration i-1), then resolves the lookback of
// t_{i-1} and decodes it.
template <int DIM, int MODE, class T, int G, int WAVES, int GMAX, class Src>
__device__ __forceinline__ void decompress_run(const Src& src, uint32_t& it, uint64_t& tile, const uint64_t ntiles,
                                               const ScanCtx& sc, SmemD<WAVES, GMAX>& sm) {
    constexpr int S = (DIM == 1) ? 4 : 8;
    constexpr int BPG = (DIM == 1) ? 2 : 1;
    constexpr int NB = BPG * G;
    constexpr int KLB = (16 / WAVES > 1) ? 16 / WAVES : 1;   // lookback words per lane
    static_assert(G <= GMAX, "shared memory sized for the largest variant");

    const int lane = (int)__lane_id();
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const uint32_t l = (uint32_t)lane & 31, h = (uint32_t)lane >> 5;
    XposeConsts xc;
    xc.init(lane);
    const uint32_t wo = word_offset<DIM>(l, h, true);        // this lane's word in a block's payload
    const uint32_t wo_wide = word_offset<DIM>(l, h, false);  // same, 32-plane 2D/3D layout
    const uint64_t stride = gridDim.x;
    Item cur = src.get(it);
    LaneGeo<DIM> lg;
    lg.init(lane, cur.geo);
    // header byte of this lane's block in a tile (lane j <-> block j of the wave)
    auto header_byte = [&](uint64_t lt) -> uint32_t {
        const uint64_t g0 = (lt * WAVES + (uint64_t)wave) * G;
        const uint64_t b = (DIM == 1) ? 2 * g0 + (uint64_t)lane : g0 + (uint64_t)lane;
        const bool ok = (lane < NB) && (DIM == 1 || b < cur.geo.nblocks);
        return ok ? (uint32_t)((const uint8_t*)cur.src)[b] : 0u;
    };
    uint32_t byte = header_byte(tile - cur.tile0);

    // the pending tile: block sizes known and byte count published
    bool pend = false;
    uint32_t pbuf = 0;
    uint64_t p_tile = 0, p_wexcl = 0, p_ttot = 0, p_wtot = 0;
    uint32_t p_off = 0, p_pref = 0;   // bookkeeping lane: block offset, outlier prefix bytes (0: none)

    // lookback, stream loads, decode and stores of the pending tile (every wave)
    auto flush = [&]() {
        const Geom& geo = cur.geo;
        const uint8_t* __restrict__ cmp = (const uint8_t*)cur.src;
        T* __restrict__ out = (T*)cur.dst;
        const T eb = (T)cur.param;
        const uint64_t grp0 = ((p_tile - cur.tile0) * WAVES + (uint64_t)wave) * G;
        const uint64_t excl = wg_lookback<WAVES, KLB>(sc, p_tile, cur.tile0, p_ttot, wave, lane, sm.lb,
            [&](uint64_t p) { return header_tile_bytes<DIM, MODE, G, WAVES>(cmp, geo, p - cur.tile0, lane); },
            true);
        const uint64_t wave_base = excl + p_wexcl;
        const uint8_t* const wbase = cmp + uniform64(geo.H + wave_base);
        const bool intact = geo.H + wave_base + p_wtot <= cur.limit;
        if (!intact && lane == 0) atomicOr(cur.err, kErrTruncated);
        if constexpr (MODE == kOutlier) {
            // a selected block starts with nb <= 4 prefix bytes and then a 4-byte sign
            // word, so one u32 read at the block start stays inside the block
            uint32_t o = 0;
            if (p_pref && intact) {
                o = load_u32u(wbase + p_off);
                if (p_pref < 4) o &= (1u << (8 * p_pref)) - 1;
            }
            if (lane < NB) sm.aux[wave][lane] = o;
            __builtin_amdgcn_wave_barrier();
        }
        // ---- the stream words of all G groups -----------------------------------
        uint32_t W[G];
#pragma unroll
        for (int g = 0; g < G; ++g) {
            const Meta m = sm.meta[pbuf][wave][BPG * g + (DIM == 1 ? h : 0)];
            const bool wideg = (DIM != 1) && (MODE != kOutlier) && m.wide();
            W[g] = (intact && m.has(l)) ? load_u32u(wbase + (m.off() + (wideg ? wo_wide : wo))) : 0u;
        }
        // ---- decode ---------------------------------------------------------------
        Walker wk;
        if constexpr (DIM != 1) wk.init(grp0, geo);
#pragma unroll
        for (int g = 0; g < G; ++g) {
            const Meta m = sm.meta[pbuf][wave][BPG * g + (DIM == 1 ? h : 0)];
            const bool wideg = (DIM != 1) && (MODE != kOutlier) && m.wide();
            const uint32_t x = xpose32(W[g], xc);   // lane l: sign in bit 31 over the magnitude bits
            uint32_t a, neg;
            if (!wideg) {
                neg = x >> 31;
                a = x & 0x7fffffffu;
            } else {   // 32 planes: the sign words sit in their own two words
                const uint32_t sw = (intact && m.nz()) ? load_u32u(wbase + (m.off() + 4 * h)) : 0u;
                neg = (sw >> l) & 1;
                a = x;
            }
            if constexpr (MODE == kOutlier) {
                if (m.sel() && l == 7 && (DIM == 1 || h == 0)) a = sm.aux[wave][BPG * g + (DIM == 1 ? h : 0)];   // element 0
            }
            const uint32_t dm = neg ? 0u - a : a;
            const uint32_t d = dpp<kDppHalfMirror>(dm);   // back to natural element order
            uint32_t q;
            if constexpr (MODE == kFixed) q = d;
            else q = inverse_lorenzo<DIM>(d, lane);
            const bool zero_blk = !intact || !m.nz();
            const T val = zero_blk ? T(0) : (T)(int32_t)q * eb * T(2);   // src/cuSZp_kernels_1D_f32.cu:490
            uint64_t idx;
            const bool valid = elem_of<DIM>(geo, lg, wk, grp0 + g, lane, idx);
            if (valid) out[idx] = val;
            if constexpr (DIM != 1) wk.next(geo);
        }
        pend = false;
    };

    for (uint32_t cbuf = 0;; cbuf ^= 1) {
    const Geom& geo = cur.geo;
    const uint64_t lt = tile - cur.tile0;
    const uint64_t grp0 = (lt * WAVES + (uint64_t)wave) * G;

    // ---- bookkeeping: lane j <-> block j of this wave -----------------------
    const uint64_t blk_idx = (DIM == 1) ? 2 * grp0 + (uint64_t)lane : grp0 + (uint64_t)lane;
    const bool blk = (lane < NB) && (DIM == 1 || blk_idx < geo.nblocks);
    const BlockView v = view_of<MODE, S>(byte);
    const uint32_t sz = blk ? v.size : 0u;
    const uint32_t incl = scan_wave(sz);
    const uint32_t off = incl - sz;
    const uint64_t wtot = rdlane(incl, kWaveSize - 1);
    // every entry is rewritten: blocks past the end get an empty lane mask, so no
    // load ever uses an offset left over from an earlier tile or array
    if (lane < NB) sm.meta[cbuf][wave][lane] = make_meta(off, v, blk && sz != 0);
    if (lane == 0) sm.wtot[cbuf][wave] = wtot;
    __syncthreads();
    uint64_t wexcl = 0, ttot = 0;
#pragma unroll
    for (int w = 0; w < WAVES; ++w) {
        const uint64_t t = sm.wtot[cbuf][w];
        wexcl += (w < wave) ? t : 0;
        ttot += t;
    }
    if (wave == 0 && lane == 0) publish_tile(sc, tile, cur.tile0, ttot);

    // the next tile's header bytes load while the pending tile decodes
    const uint64_t next = tile + stride;
    const bool same = next < cur.tile0 + geo.ntiles;
    const uint32_t c_pref = (blk && v.sel) ? v.nb : 0u;
    if (same) byte = header_byte(next - cur.tile0);

    if (pend) flush();   // the previous tile's lookback and decode

    pend = true;
    pbuf = cbuf;
    p_tile = tile;
    p_wexcl = wexcl;
    p_ttot = ttot;
    p_wtot = wtot;
    p_off = off;
    p_pref = c_pref;

    tile = next;
    if (same) continue;
    flush();   // last tile of this item for this workgroup
    if (tile >= ntiles) return;
    it = src.locate(tile, it);
    if (src.variant(it) != cur.variant) return;
    cur = src.get(it);
    lg.init(lane, cur.geo);
    byte = header_byte(tile - cur.tile0);
    }   // tiles
}


// ---------------------------------------------------------------------------
// single-array kernels: one item, every tile of the same variant
// ---------------------------------------------------------------------------
template <int DIM, int MODE, class T, int G, int WAVES>
__global__ void __launch_bounds__(WAVES * kWaveSize)
compress_kernel(const Item item, const ScanCtx sc) {
    __shared__ SmemC<WAVES, G> sm;
    const OneItem src{item};
    uint32_t it = 0;
    uint64_t tile = blockIdx.x;
    if (tile < item.geo.ntiles) compress_run<DIM, MODE, T, G, WAVES>(src, it, tile, item.geo.ntiles, sc, sm);
}

template <int DIM, int MODE, class T, int G, int WAVES>
__global__ void __launch_bounds__(WAVES * kWaveSize)
decompress_kernel(const Item item, const ScanCtx sc) {
    __shared__ SmemD<WAVES, G> sm;
    const OneItem src{item};
    uint32_t it = 0;
    uint64_t tile = blockIdx.x;
    if (tile < item.geo.ntiles) decompress_run<DIM, MODE, T, G, WAVES>(src, it, tile, item.geo.ntiles, sc, sm);
}

}  // namespace cuszp

namespace cuszp {

// ---------------------------------------------------------------------------
// batched kernels: the arrays of a table, of any variant, in one persistent
// launch. Workgroup w takes tiles w, w + B, ... of the concatenated tile space;
// a tile's array is found by binary search on tile0, and each array's tiles
// form their own lookback segment, so every array's stream is exactly its
// single-array stream. f64 arrays use half as many groups per wave (tiles of
// the same byte size as f32 tiles).
// ---------------------------------------------------------------------------
constexpr int kBatchGF32 = 16;
constexpr int kBatchGF64 = 8;

#define CUSZP_BATCH_CASES(RUN)                                                        \
    case 0: RUN(1, kFixed, float, kBatchGF32); break;                                 \
    case 1: RUN(1, kPlain, float, kBatchGF32); break;                                 \
    case 2: RUN(1, kOutlier, float, kBatchGF32); break;                               \
    case 3: RUN(2, kFixed, float, kBatchGF32); break;                                 \
    case 4: RUN(2, kPlain, float, kBatchGF32); break;                                 \
    case 5: RUN(2, kOutlier, float, kBatchGF32); break;                               \
    case 6: RUN(3, kFixed, float, kBatchGF32); break;                                 \
    case 7: RUN(3, kPlain, float, kBatchGF32); break;                                 \
    case 8: RUN(3, kOutlier, float, kBatchGF32); break;                               \
    case 9: RUN(1, kFixed, double, kBatchGF64); break;                                \
    case 10: RUN(1, kPlain, double, kBatchGF64); break;                               \
    case 11: RUN(1, kOutlier, double, kBatchGF64); break;                             \
    case 12: RUN(2, kFixed, double, kBatchGF64); break;                               \
    case 13: RUN(2, kPlain, double, kBatchGF64); break;                               \
    case 14: RUN(2, kOutlier, double, kBatchGF64); break;                             \
    case 15: RUN(3, kFixed, double, kBatchGF64); break;                               \
    case 16: RUN(3, kPlain, double, kBatchGF64); break;                               \
    case 17: RUN(3, kOutlier, double, kBatchGF64); break;                             \
    default: return;

template <int WAVES>
__global__ void __launch_bounds__(WAVES * kWaveSize)
compress_batch_kernel(const Item* __restrict__ items, uint32_t nitems, uint64_t ntiles, ScanCtx sc,
                      const unsigned long long* __restrict__ epoch) {
    __shared__ SmemC<WAVES, kBatchGF32> sm;
    sc.tag = (epoch[0] & kEpochMask) << kEpochShift;
    const ItemTable src{items, nitems};
    uint64_t tile = blockIdx.x;
    if (tile >= ntiles) return;
    uint32_t it = src.locate(tile, 0);
#define CUSZP_RUN_C(D, M, T, G) compress_run<D, M, T, G, WAVES>(src, it, tile, ntiles, sc, sm)
    while (tile < ntiles) {
        switch (src.variant(it)) { CUSZP_BATCH_CASES(CUSZP_RUN_C) }
    }
#undef CUSZP_RUN_C
}

template <int WAVES>
__global__ void __launch_bounds__(WAVES * kWaveSize)
decompress_batch_kernel(const Item* __restrict__ items, uint32_t nitems, uint64_t ntiles, ScanCtx sc,
                        const unsigned long long* __restrict__ epoch) {
    __shared__ SmemD<WAVES, kBatchGF32> sm;
    sc.tag = (epoch[0] & kEpochMask) << kEpochShift;
    const ItemTable src{items, nitems};
    uint64_t tile = blockIdx.x;
    if (tile >= ntiles) return;
    uint32_t it = src.locate(tile, 0);
#define CUSZP_RUN_D(D, M, T, G) decompress_run<D, M, T, G, WAVES>(src, it, tile, ntiles, sc, sm)
    while (tile < ntiles) {
        switch (src.variant(it)) { CUSZP_BATCH_CASES(CUSZP_RUN_D) }
    }
#undef CUSZP_RUN_D
}

}  // namespace cuszp
