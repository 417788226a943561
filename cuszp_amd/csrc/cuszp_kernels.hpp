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
//   * a workgroup is WAVES waves and handles one tile: each wave owns kChunks
//     chunks of G consecutive groups, so a tile is WAVES*kChunks*G groups.
//     The grid has one workgroup per tile (tile = launch base + blockIdx.x).
//     No ticket atomics: one hot device-scope atomic serialises at about
//     12 ns per operation on MI355X;
//   * lane e holds element e of the group. Each lane packs its sign into bit 31
//     above the 31-bit magnitude; a 32x32 bit-matrix transpose across each
//     32-lane half (five ds_swizzle xor stages) then leaves plane p's stream
//     word in lane p and the sign word in lane 31 -- for every rate at a fixed
//     cost. The transpose also puts the elements in the bit order of cuSZp's
//     sign/plane bytes (element e at byte e/8, bit 7-e%8, i.e. bit e^7 of the
//     word) at no extra cost; decompression runs the inverse transpose;
//   * per-block bookkeeping (rates, header bytes, sizes, offsets) runs once per
//     wave with one lane per block;
//   * the inter-tile exclusive scan of compressed bytes is a single-pass
//     decoupled lookback: one 64-bit status word per tile packs
//     {epoch, flag, value}, is read 128 or 256 predecessors at a time by one
//     wave, and
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
constexpr uint32_t kWarnLossy = 16u;     // a 2D/3D outlier block has rate byte 32, which decoders read as rate 0
constexpr uint32_t kErrMask = 7u;

// Item flags. kItemStreamNt: compress writes the stream with nontemporal
// stores. Set for arrays of at least kStreamNtBytes: their streams outgrow the
// memory-side cache at typical ratios, and dirty stream lines left there
// would be written back during the next kernel. A smaller stream is stored
// normally and stays cached for its next reader.
constexpr uint32_t kItemStreamNt = 1u;
constexpr uint64_t kStreamNtBytes = 1ull << 30;

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
    unsigned long long* diag;       // [0] first stalled tile + 1, [1] its missing predecessor, [2] word seen;
                                    // with `stats`: [4] polls that found a window unresolved, [5] lookbacks
                                    // that polled, [6] predecessors claimed, [7] clock ticks spent waiting
    unsigned long long tag;         // epoch << kEpochShift
    uint32_t steal_ticks;           // wait (s_memrealtime ticks) before a pending predecessor is claimed
    uint32_t hold_every;            // scan stress: tile t with t % hold_every == hold_every - 1 waits
    uint32_t hold_ticks;            //   hold_ticks before it publishes anything (0: off)
    uint32_t stats;                 // nonzero: accumulate the lookback counters into diag[4..7]
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
    uint32_t flags;                 // kItemStreamNt
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
constexpr int kDppQuadBcast0 = 0x00;  // quad_perm [0,0,0,0]
constexpr int kDppRowNewBcast0 = 0x150;   // lane 0 of each row to the whole row

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
// 32x32 bit transposes within each 32-lane half, with the element order of
// cuSZp's sign/plane bytes folded in. In the stream, element e of a 32-element
// half sits at bit e^7 of its plane word (byte e/8, bit 7-e%8). Lanes hold
// elements in natural order, so:
//   compress   (ENC): out lane i bit j = in lane j^7 bit i   (lane e's word -> plane words)
//   decompress (DEC): out lane i bit j = in lane j bit i^7   (plane words -> lane e's word)
// Stage J exchanges lane-index bit log2(J) with data bit log2(J). Stages 16 and
// 8 move whole bytes (v_perm); stages 4, 2, 1 rotate the lane's own word and
// merge in its partner's (v_alignbit + v_bfi). Exchanging the bit with its
// complement in those three stages (the order twist e -> e^7) costs nothing:
// it only changes which half of each pair stays in place, i.e. the per-lane
// shift and mask constants. The partner value (lane^J) comes from ds_swizzle
// (xor mode), which runs on the LDS pipe, not the VALU.
// ---------------------------------------------------------------------------
struct XposeConsts {
    uint32_t sh1, mk1, sh2, mk2, sh4, mk4, sel8, sel16;
    // dec: the decompress transpose (plane words -> elements)
    __device__ __forceinline__ void init(int lane, bool dec) {
        const int l = lane & 31;
        const bool lo1 = !(l & 1) != dec, lo2 = !(l & 2) != dec, lo4 = !(l & 4) != dec;
        const bool lo8 = !(l & 8), lo16 = !(l & 16);
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
    x = bfi(c.mk4, rotr(x, c.sh4), swizzle_xor<4>(x));
    x = bfi(c.mk2, rotr(x, c.sh2), swizzle_xor<2>(x));
    x = bfi(c.mk1, rotr(x, c.sh1), swizzle_xor<1>(x));
    return x;
}
// The same transpose on N independent values, stage by stage, so the N
// swizzles of a stage are in flight together (one LDS round trip per stage
// instead of one per value); the lanes' own rotations are computed while the
// swizzles are in flight.
template <int N>
__device__ __forceinline__ void xpose32_n(uint32_t* x, const XposeConsts& c) {
    uint32_t t[N];
#pragma unroll
    for (int i = 0; i < N; ++i) t[i] = swizzle_xor<16>(x[i]);
#pragma unroll
    for (int i = 0; i < N; ++i) x[i] = __builtin_amdgcn_perm(t[i], x[i], c.sel16);
#pragma unroll
    for (int i = 0; i < N; ++i) t[i] = swizzle_xor<8>(x[i]);
#pragma unroll
    for (int i = 0; i < N; ++i) x[i] = __builtin_amdgcn_perm(t[i], x[i], c.sel8);
#pragma unroll
    for (int i = 0; i < N; ++i) t[i] = swizzle_xor<4>(x[i]);
#pragma unroll
    for (int i = 0; i < N; ++i) x[i] = bfi(c.mk4, rotr(x[i], c.sh4), t[i]);
#pragma unroll
    for (int i = 0; i < N; ++i) t[i] = swizzle_xor<2>(x[i]);
#pragma unroll
    for (int i = 0; i < N; ++i) x[i] = bfi(c.mk2, rotr(x[i], c.sh2), t[i]);
#pragma unroll
    for (int i = 0; i < N; ++i) t[i] = swizzle_xor<1>(x[i]);
#pragma unroll
    for (int i = 0; i < N; ++i) x[i] = bfi(c.mk1, rotr(x[i], c.sh1), t[i]);
}
// Bit j of each 32-bit half of m moved to bit j^7 (each byte bit-reversed):
// a per-element mask in natural order as a sign word in stream order.
__device__ __forceinline__ uint64_t stream_order64(uint64_t m) {
    const uint32_t lo = __builtin_bswap32(__builtin_bitreverse32((uint32_t)m));
    const uint32_t hi = __builtin_bswap32(__builtin_bitreverse32((uint32_t)(m >> 32)));
    return (uint64_t)lo | ((uint64_t)hi << 32);
}

// p[lane - 1] - q[lane] in one DPP subtract (wave_shr:1; lane 0 reads p = 0).
// Measured on gfx950 with this assembler: v_sub_u32_dpp d, a, b gives
// a[lane - 1] - b[lane]; v_subrev_u32_dpp d, a, b gives b[lane - 1] - a[lane]
// (the shift is on its second operand), so neither forms q[lane] - q[lane - 1]
// in one instruction. The two wait states a DPP read of a VGPR needs after the
// VALU write of it are in the block, since the compiler cannot see into it.
__device__ __forceinline__ uint32_t sub_prev_wave_shr1(uint32_t p, uint32_t q) {
    uint32_t d;
    asm("s_nop 1\n\tv_sub_u32_dpp %0, %1, %2 wave_shr:1 row_mask:0xf bank_mask:0xf bound_ctrl:1"
        : "=v"(d)
        : "v"(p), "v"(q));
    return d;
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
    uint32_t sel, nb, R, size, head;
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

// The outlier header byte and its decoder view in one step, for fr1, fr2 <= 31
// (no block of 32 planes). There view_of(rate_byte(fr1, fr2, o_bits)) is the
// encoder's own choice: a selected block has an element-0 residual, so nb is
// 1..4 and the byte's fields [4:0] = fr2, [6:5] = nb - 1 decode back exactly;
// an unselected byte fr1 <= 31 decodes to R = fr1 and size ofs1.
template <int S>
__device__ __forceinline__ BlockView outlier_code(uint32_t fr1, uint32_t fr2, uint32_t o_bits, uint32_t& byte) {
    const uint32_t nb = (o_bits + 7) >> 3;
    const uint32_t ofs1 = fr1 ? S + fr1 * S : 0;
    const uint32_t ofs2 = S + fr2 * S + nb;
    const bool sel = ofs1 > ofs2;
    byte = sel ? (fr2 | 0x80u | ((nb - 1) << 5)) : fr1;
    BlockView v;
    v.sel = sel;
    v.nb = nb;
    v.R = sel ? fr2 : fr1;
    v.size = sel ? ofs2 : ofs1;
    v.head = sel ? nb : 0;
    return v;
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
// LDS byte pointers and unaligned LDS words (gfx950 accesses LDS unaligned).
typedef __attribute__((address_space(3))) uint8_t lds_u8;
typedef __attribute__((address_space(3))) uint32_t __attribute__((aligned(1))) lds_u32u;
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

// Exclusive prefix of `tile` within its segment (tiles first .. tile-1), run by
// one whole wave (no workgroup barriers inside).
//
// The wave publishes the tile's aggregate (the segment's first tile publishes
// its inclusive prefix and returns 0), then reads predecessors a window of
// 64 * K at a time: lane j, word k reads position p = 64k + j, i.e. tile
// tile - 1 - p (the K loads of a window are issued back to back: one round
// trip). A window is resolved once every position up to the nearest published
// inclusive prefix P is ready; the values at positions <= P sum to the
// exclusive prefix (with no prefix in the window, the window's sum carries
// into the next one). The wave then publishes the tile's inclusive prefix.
// Positions before the segment read words of an earlier segment or of the
// guard in front of the status array (kStatusGuard words) and count as an
// inclusive prefix of 0.
//
// Tiles are numbered in launch order, so a predecessor normally belongs to a
// workgroup that started earlier. Forward progress does not depend on that:
// when a window stays unresolved for sc.steal_ticks (kStealTicks unless a test
// lowers it), the wave claims the farthest
// pending predecessor of the window (CAS from the word it saw to CLAIM),
// computes that tile's aggregate with agg_of(tile) and publishes it. A claim is
// held only by a running wave that computes without waiting, so every wait
// ends; the owner's later store of the same aggregate (or of its inclusive
// prefix) is harmless.
//
// pre (with has_pre): the window's first read, issued earlier by the caller
// with lookback_window(); its words may be stale (pending ones are read again).
template <int K>
struct LookbackWindow {
    uint64_t w[K];
};
template <int K>
__device__ __forceinline__ LookbackWindow<K> lookback_window(const ScanCtx& sc, uint64_t tile, int lane) {
    LookbackWindow<K> pw;
    const unsigned long long* const wp = sc.status + ((int64_t)tile - 1 - lane);
#pragma unroll
    for (int k = 0; k < K; ++k) pw.w[k] = ld_status(wp - 64 * k);
    return pw;
}
template <int K, class AggF>
__device__ __forceinline__ uint64_t wave_lookback(const ScanCtx& sc, uint64_t tile, uint64_t first, uint64_t agg,
                                                  int lane, AggF&& agg_of, const LookbackWindow<K> pre = {},
                                                  bool has_pre = false) {
    constexpr int WIN = 64 * K;
    const uint64_t tag = sc.tag;
    if (sc.hold_every && tile % sc.hold_every == sc.hold_every - 1) {
        // scan stress (tests): this tile publishes late, so its successors
        // claim it and compute its aggregate themselves; the wait is bounded
        const uint64_t t0 = __builtin_amdgcn_s_memrealtime();
        while (__builtin_amdgcn_s_memrealtime() - t0 < sc.hold_ticks) __builtin_amdgcn_s_sleep(8);
    }
    if (tile == first) {
        if (lane == 0) st_status(sc.status + tile, tag | kFlagPrefix | agg);
        return 0;
    }
    if (lane == 0) st_status(sc.status + tile, tag | kFlagAgg | agg);
    uint64_t excl = 0;
    int64_t hi = (int64_t)tile;
    uint32_t spins = 0;
    const uint64_t t_wait = __builtin_amdgcn_s_memrealtime();
    bool stalled = false;
    int64_t claimed = -1;   // a predecessor this wave claimed and must compute
    for (;;) {
        if (claimed >= 0) {   // computed with no window words live; the window is then re-read
            const uint64_t a = agg_of((uint64_t)claimed);
            if (lane == 0) {
                cas_status(sc.status + claimed, tag | kFlagClaim, tag | kFlagAgg | a);
                atomicOr(sc.err, kInfoStolen);
                if (sc.stats) atomicAdd(sc.diag + 6, 1ull);
            }
            claimed = -1;
        }
        const unsigned long long* const wp = sc.status + (hi - 1 - lane);   // position 64k + lane: wp - 64k
        const int64_t rel = hi - 1 - lane - (int64_t)first;                 // before the segment iff rel < 64k
        uint64_t w[K];
        if (has_pre) {   // the first window, read before (only once: a claim re-reads)
#pragma unroll
            for (int k = 0; k < K; ++k) w[k] = pre.w[k];
            has_pre = false;
        } else {
#pragma unroll
            for (int k = 0; k < K; ++k) w[k] = ld_status(wp - 64 * k);
        }
#pragma unroll
        for (int k = 0; k < K; ++k)
            if (rel < 64 * k) w[k] = tag | kFlagPrefix;
        int P;
        for (;;) {
            P = WIN;
            int N = WIN;
            uint64_t nr[K];
#pragma unroll
            for (int k = K - 1; k >= 0; --k) {
                const bool rd = status_ready(w[k], tag);
                const uint64_t pm = ballot(rd && (w[k] & kFlagMask) == kFlagPrefix);
                if (pm) P = 64 * k + __builtin_ctzll(pm);
                nr[k] = ballot(!rd);
                if (nr[k]) N = 64 * k + __builtin_ctzll(nr[k]);
            }
            if (N > P || N == WIN) break;   // every position up to P is ready
            ++spins;
            // farthest pending position at or below P: (sk, sl)
            int sk = 0, sl = 0;
#pragma unroll
            for (int k = 0; k < K; ++k) {
                const int lo = 64 * k;
                const uint64_t m = P < lo ? 0ull : (P < lo + 63 ? nr[k] & ((2ull << (P - lo)) - 1) : nr[k]);
                if (m) { sk = k; sl = 63 - __builtin_clzll(m); }
            }
            uint64_t pend = 0;
#pragma unroll
            for (int k = 0; k < K; ++k) if (k == sk) pend = rdlane64(w[k], sl);
            const int64_t pidx = hi - 1 - (64 * sk + sl);
            if (spins >= kSpinLimit ||
                ((spins & 1023) == 0 &&
                 (__hip_atomic_load(sc.err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & kErrStall))) {
                // last-resort watchdog: record the stuck predecessor and give up
                // (the output is reported invalid; publishing below unblocks successors)
                stalled = true;
                if (lane == 0) {
                    atomicOr(sc.err, kErrStall);
                    if (atomicCAS(sc.diag, 0ull, tile + 1) == 0ull) {
                        sc.diag[1] = (uint64_t)pidx;
                        sc.diag[2] = pend;
                    }
                }
                break;
            }
            if (__builtin_amdgcn_s_memrealtime() - t_wait > sc.steal_ticks &&
                !(((pend & kTagMask) == tag) && ((pend & kFlagMask) == kFlagClaim))) {
                uint64_t seen = 0;
                if (lane == 0) seen = cas_status(sc.status + pidx, pend, tag | kFlagClaim);
                if (rdlane64(seen, 0) == pend) {
                    claimed = pidx;
                    break;
                }
            }
            __builtin_amdgcn_s_sleep(1);
            // re-read the pending words (all loads first; a ready word stays ready)
#pragma unroll
            for (int k = 0; k < K; ++k)
                if ((nr[k] >> lane) & 1) w[k] = ld_status(wp - 64 * k);
        }
        if (claimed >= 0) continue;
        uint64_t part = 0;
#pragma unroll
        for (int k = 0; k < K; ++k) part += (64 * k + lane <= P) ? (w[k] & kValueMask) : 0;
        excl += wave_sum_u40(part);
        if (P < WIN || stalled) break;
        hi -= WIN;
    }
    if (lane == 0) {
        st_status(sc.status + tile, tag | kFlagPrefix | (excl + agg));
        if (sc.stats && spins) {
            atomicAdd(sc.diag + 4, (unsigned long long)spins);
            atomicAdd(sc.diag + 5, 1ull);
            atomicAdd(sc.diag + 7, (unsigned long long)(__builtin_amdgcn_s_memrealtime() - t_wait));
        }
    }
    return excl;
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
    // advance by n <= Bx blocks (crosses at most one row end)
    __device__ __forceinline__ void step(const Geom& g, uint32_t n) {
        b += n;
        bx += n;
        if (bx >= g.Bx) {
            bx -= g.Bx;
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
        // (lane e = 16z + 4y + x: a DPP row per z, a quad per y)
        const int x = e & 3, y = (e >> 2) & 3;
        uint32_t vx = x ? d : 0u;
        vx += dpp<kDppQuadShr1>(vx);
        vx += dpp<kDppQuadShr2>(vx);
        uint32_t vy = (x == 0 && y != 0) ? d : 0u;   // prefix over the quads' first lanes, then across the quad
        vy += dpp<kDppRowShr4>(vy);
        vy += dpp<kDppRowShr8>(vy);
        vy = dpp<kDppQuadBcast0>(vy);
        // the row's first lane (x = y = 0) across the row, then the prefix over the rows
        uint32_t vz = dpp<kDppRowNewBcast0>(d);
        vz += dpp<kDppBcast15, 0xa>(vz);
        vz += dpp<kDppBcast31, 0xc>(vz);
        return vz + vy + vx;
    }
}

// The same sums for N independent groups, step by step, so each DPP step of
// one group fills the other groups' hazard slots instead of s_nop.
template <int DIM, int N>
__device__ __forceinline__ void inverse_lorenzo_n(uint32_t* d, int e) {
    if constexpr (DIM == 1) {
#pragma unroll
        for (int i = 0; i < N; ++i) d[i] += dpp<kDppRowShr1>(d[i]);
#pragma unroll
        for (int i = 0; i < N; ++i) d[i] += dpp<kDppRowShr2>(d[i]);
#pragma unroll
        for (int i = 0; i < N; ++i) d[i] += dpp<kDppRowShr4>(d[i]);
#pragma unroll
        for (int i = 0; i < N; ++i) d[i] += dpp<kDppRowShr8>(d[i]);
#pragma unroll
        for (int i = 0; i < N; ++i) d[i] += dpp<kDppBcast15, 0xa>(d[i]);
    } else {
#pragma unroll
        for (int i = 0; i < N; ++i) d[i] = inverse_lorenzo<DIM>(d[i], e);
    }
}

// ---------------------------------------------------------------------------
// per-block metadata handed from the bookkeeping lanes to the stream-word
// lanes through LDS (one 16-byte entry per block):
//   a: byte offset of the block's sign word from the wave's first block (past
//      any outlier prefix)
//   b: 0 for a block without payload, else R + 1 (R planes; 33 for a 2D/3D
//      block of 32 planes). Lane l of a 32-lane half carries one of the
//      block's words iff key(l) < b, with key(31) = 0 (the sign word, or plane
//      31 of a 32-plane block) and key(l) = l + 1 (plane l): one compare.
//   c: the plane words' mask of an outlier-selected block (element 0 travels
//      in the prefix: bit 7, element 0's stream bit, cleared), else all ones
// ---------------------------------------------------------------------------
struct Meta {
    uint32_t a, b, c, pad;
    __device__ __forceinline__ uint32_t off() const { return a; }
    __device__ __forceinline__ bool sel() const { return c != ~0u; }
    __device__ __forceinline__ bool nz() const { return b != 0; }
    __device__ __forceinline__ bool has(uint32_t key) const { return key < b; }
};
template <int MODE>
__device__ __forceinline__ Meta make_meta(uint32_t off, const BlockView& v, bool nz) {
    Meta m;
    m.a = off + v.head;
    m.b = nz ? v.R + 1u : 0u;
    m.c = (MODE == kOutlier && v.sel) ? ~0x80u : ~0u;
    m.pad = 0;
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

// The arrays of a batched launch: descriptors in device memory, ordered by
// tile0; a tile belongs to the last array whose tile0 is not above it.
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
// Aggregates recomputed by a waiting wave (see wave_lookback).
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
// sizes as pass 1 of compress_tile: a block's rate is the bit length of the OR of
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

// Lookback window of a tile: 64 * kLookbackWords predecessors per round trip.
constexpr int kLookbackWords = 4;
// Words in front of every status array (window reads before tile 0 land there).
constexpr uint64_t kStatusGuard = 64 * kLookbackWords;

// ---------------------------------------------------------------------------
// Offset table of a single-array decompression: before any block is decoded,
// a pass over the header bytes alone writes the stream offset of every
// decode wave's payload (a "unit": the UB consecutive blocks one wave
// decodes). The decode waves then start from their own offset, with no
// lookback and no workgroup barrier; the header is 1/32 (1D) or 1/64 (2D/3D)
// of the values' bytes, so the pass reads little.
//
// The pass is itself one single-pass decoupled-lookback scan, over segments of
// kPrefixThreads x kPrefixBytes header bytes: thread t of segment s sums the
// block sizes of header bytes [(s * kPrefixThreads + t) * kPrefixBytes, +kPrefixBytes),
// which hold kPrefixBytes / UB whole units.
// ---------------------------------------------------------------------------
constexpr int kPrefixThreads = 256;
constexpr int kPrefixBytes = 128;
constexpr uint64_t kPrefixSegBytes = (uint64_t)kPrefixThreads * kPrefixBytes;

// Payload bytes of the header bytes [b0, b0 + kPrefixBytes) per unit of UB
// blocks; bytes at or past `limit` (2D/3D: the block count, so header pad
// bytes count nothing) are blocks without payload.
template <int MODE, int S, int UB>
__device__ __forceinline__ void unit_sums(const uint8_t* __restrict__ cmp, uint64_t b0, uint64_t limit,
                                          uint32_t (&sums)[kPrefixBytes / UB]) {
    static_assert(kPrefixBytes % UB == 0 && UB % 16 == 0, "whole units of 16-byte words");
    typedef uint32_t v4u __attribute__((ext_vector_type(4)));
    typedef uint32_t v4uu __attribute__((ext_vector_type(4), aligned(1)));
    constexpr int V = kPrefixBytes / 16;
    v4u w[V];
    if (b0 + kPrefixBytes <= limit) {
#pragma unroll
        for (int i = 0; i < V; ++i) w[i] = __builtin_nontemporal_load((const v4uu*)(cmp + b0) + i);
    } else {
        uint32_t* wb = (uint32_t*)w;
#pragma unroll
        for (int i = 0; i < 4 * V; ++i) {
            uint32_t v = 0;
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const uint64_t b = b0 + 4 * i + k;
                v |= (b < limit ? (uint32_t)cmp[b] : 0u) << (8 * k);
            }
            wb[i] = v;
        }
    }
    const uint32_t* wb = (const uint32_t*)w;
#pragma unroll
    for (int u = 0; u < kPrefixBytes / UB; ++u) {
        uint32_t s = 0;
#pragma unroll
        for (int i = 0; i < UB / 4; ++i) {
            const uint32_t x = wb[u * (UB / 4) + i];
#pragma unroll
            for (int k = 0; k < 4; ++k) s += view_of<MODE, S>((x >> (8 * k)) & 0xffu).size;
        }
        sums[u] = s;
    }
}

// Payload bytes of segment `seg`, by one wave (a waiting wave's recomputation
// of a predecessor it claimed; see wave_lookback).
template <int MODE, int S>
__device__ uint64_t segment_bytes(const uint8_t* __restrict__ cmp, uint64_t limit, uint64_t seg, int lane) {
    uint64_t sum = 0;
    for (uint32_t i = (uint32_t)lane; i < kPrefixSegBytes; i += kWaveSize) {
        const uint64_t b = seg * kPrefixSegBytes + i;
        sum += view_of<MODE, S>(b < limit ? (uint32_t)cmp[b] : 0u).size;
    }
    return wave_sum_u40(sum);
}

// offs[u] = H + payload bytes of units 0 .. u-1, for u < nunits; segment =
// base + blockIdx.x of nseg. With `total`, the last segment writes the stream
// length the header describes, H + all payload bytes.
template <int MODE, int S, int UB>
__global__ void __launch_bounds__(kPrefixThreads)
unit_offsets_kernel(const uint8_t* __restrict__ cmp, const uint64_t limit, const uint64_t H, const uint64_t nunits,
                    unsigned long long* __restrict__ offs, unsigned long long* __restrict__ total,
                    const ScanCtx sc, const uint64_t nseg, const uint64_t base) {
    constexpr int UPT = kPrefixBytes / UB;   // units per thread
    constexpr int NW = kPrefixThreads / kWaveSize;
    __shared__ uint32_t wsum[NW];
    __shared__ uint64_t excl;
    const uint64_t seg = base + blockIdx.x;
    const int t = (int)threadIdx.x, lane = (int)__lane_id();
    const int wave = __builtin_amdgcn_readfirstlane(t >> 6);
    const uint64_t th = seg * kPrefixThreads + (uint64_t)t;
    uint32_t sums[UPT];
    unit_sums<MODE, S, UB>(cmp, th * kPrefixBytes, limit, sums);
    uint32_t tot = 0;
#pragma unroll
    for (int k = 0; k < UPT; ++k) tot += sums[k];
    const uint32_t incl = scan_wave(tot);   // < 2^32: a segment holds at most 2^15 blocks of <= 2 KiB
    if (lane == kWaveSize - 1) wsum[wave] = incl;
    // wave 0's first lookback read overlaps the other waves' sums
    LookbackWindow<kLookbackWords> pre = {};
    if (wave == 0 && seg != 0) pre = lookback_window<kLookbackWords>(sc, seg, lane);
    __syncthreads();
    uint32_t wex = 0, stot = 0;
#pragma unroll
    for (int w = 0; w < NW; ++w) {
        const uint32_t v = wsum[w];
        wex += (w < wave) ? v : 0u;
        stot += v;
    }
    if (wave == 0) {
        const uint64_t e = wave_lookback<kLookbackWords>(sc, seg, 0, stot, lane, [&](uint64_t p) {
            return segment_bytes<MODE, S>(cmp, limit, p, lane);
        }, pre, true);
        if (lane == 0) {
            excl = e;
            if (seg == nseg - 1 && total) *total = H + e + stot;
        }
    }
    __syncthreads();
    uint64_t o = H + excl + wex + (incl - tot);
#pragma unroll
    for (int k = 0; k < UPT; ++k) {
        const uint64_t u = th * UPT + (uint64_t)k;
        if (u < nunits) offs[u] = o;
        o += sums[k];
    }
}

// ---------------------------------------------------------------------------
// one tile per workgroup pass
// ---------------------------------------------------------------------------

// Tile shape. A wave works through kChunks chunks of ShapeC<T>::G groups (one
// element per lane per group); a tile is WAVES * kChunks * G groups (16384 f32
// or 8192 f64 elements at 4 waves), for compress and decompress alike.
template <class T> struct ShapeC;
template <> struct ShapeC<float> { static constexpr int G = 8; };
template <> struct ShapeC<double> { static constexpr int G = 4; };
constexpr int kChunks = 8;
__host__ __device__ constexpr uint64_t tile_groups(bool f64, int waves) {
    return (uint64_t)waves * kChunks * (f64 ? 4 : 8);
}

// Compressed bytes a compress wave stages in LDS; a wave whose payload is
// larger (compression ratio below about 2.7 on its share of the tile)
// recomputes its groups once the tile's offset is known and stores them
// directly.
constexpr uint32_t kStageBytes = 6144;
// Waves per SIMD the compress kernels are built for (the register budget):
// six 4-wave workgroups of about 26 KB of LDS fit one CU's 160 KB.
constexpr int kCompressWavesPerSimd = 6;
// Waves per SIMD the decompress kernels are built for: 8 needs at most 64
// VGPRs and few enough SGPRs for eight waves' scalar files.
constexpr int kDecompressWavesPerSimd = 8;
// Raw buffer resources (gfx9): word 3 of the descriptor, and an offset that is
// always past a resource's range (the range of one wave's payload is < 2^31).
constexpr int kBufferWord3 = 0x00020000;
constexpr uint32_t kAbsent = 0x80000000u;
// Groups whose bit transposes run side by side (a divisor of the groups per chunk).
constexpr int kXposeBatch = 4;
// 2D/3D decompress: a chunk's G blocks along x are rows of 128 (3D) or 256 (2D)
// bytes; they pass through LDS, each row padded by 16 bytes (fewer bank
// conflicts), so the global stores write whole rows 16 bytes per lane.
constexpr int kOutPad = 16;
constexpr int kOutStage = 16 * (128 + kOutPad);   // the larger of 16 x 128 B and 8 x 256 B, padded
// Largest chunk payload (from its 4-byte-aligned start) read with dense loads
// into the LDS buffer of a decompress wave; kOutStage bytes hold it.
constexpr int kDenseBytes = 2048;
static_assert(kDenseBytes <= kOutStage, "dense chunk buffer");
// The LDS buffer of a decompress wave: a 1D wave's whole payload, read in one
// burst of up to four 16-byte-per-lane loads where it fits; otherwise one
// chunk's payload at a time (kDenseBytes), or 2D/3D decoded rows (kOutStage).
// 4032 bytes keep four 8-wave workgroups within a CU's 160 KiB of LDS.
constexpr int kWaveBuf = 4032;
static_assert(kWaveBuf >= kOutStage && kWaveBuf % 16 == 0, "wave buffer");

// Shared memory of a compress workgroup.
template <int WAVES>
struct SmemC {
    uint64_t wtot[WAVES];                      // per-wave byte counts
    uint64_t excl;                             // the tile's exclusive prefix (wave 0's lookback)
    uint32_t over[WAVES];                      // the wave's payload did not fit its staging area
    Meta meta[WAVES][16];                      // per-block metadata of the chunk in work
    uint32_t aux[WAVES][16];                   // sign words of the chunk's 32-plane 2D/3D groups
    uint8_t hdr[WAVES][128];                   // header bytes of the wave's blocks, stored after its chunks
    uint4 stage[WAVES][kStageBytes / 16 + 1];  // the wave's payload, as it will appear in the stream (+ a spare word)
};

// Shared memory of a decompress workgroup.
// Per-block metadata of a decompress wave, in LDS (one 8-byte entry per block
// of the wave, written once by the wave's bookkeeping):
//   a:  byte offset of the block's sign word within the wave's payload (past any
//       outlier prefix)
//   nb: [31:24] 0 for a block without payload (or of a truncated stream), else
//       R + 1 (33 for 32 planes); [8] outlier-selected; [9] 32 planes (2D/3D
//       only: the sign words move in front of the planes)
// Lane l of a 32-lane half carries one of the block's words iff key(l) < R + 1,
// with key(31) = 0 (the sign word, or plane 31 of a 32-plane block) and
// key(l) = l + 1 (plane l). A lane holds its key as (key << 24) | 0xffffff, so
// the test is one unsigned compare with nb whatever its low flag bits.
struct DMeta {
    uint32_t a, nb;
    __device__ __forceinline__ uint32_t off() const { return a; }
    __device__ __forceinline__ bool wide() const { return (nb >> 9) & 1; }
    __device__ __forceinline__ bool sel() const { return (nb >> 8) & 1; }
    __device__ __forceinline__ bool nz() const { return nb >= (1u << 24); }
    __device__ __forceinline__ bool has(uint32_t keyx) const { return keyx < nb; }
    // Buffer offset of the lane's word, wo bytes into the block (a block
    // without 32 planes), or an offset past every buffer range (reads 0).
    __device__ __forceinline__ uint32_t word_at(uint32_t keyx, uint32_t wo) const {
        return has(keyx) ? a + wo : kAbsent;
    }
};
template <int DIM>
__device__ __forceinline__ DMeta make_dmeta(uint32_t off, const BlockView& v, bool nz) {
    DMeta m;
    const bool wide = DIM != 1 && v.R >= 32;
    m.a = off + v.head;
    m.nb = (nz ? (v.R + 1u) << 24 : 0u) | (v.sel << 8) | ((uint32_t)wide << 9);
    return m;
}
__device__ __forceinline__ uint32_t lane_keyx(uint32_t l) { return (((l == 31) ? 0u : l + 1u) << 24) | 0xffffffu; }

// Blocks of metadata per decompress wave (its kChunks * blocks-per-chunk).
constexpr int kMetaPerWave = 128;
template <int WAVES>
struct SmemD {
    uint64_t wtot[WAVES];             // per-wave byte counts
    uint64_t excl;                    // the tile's exclusive prefix (wave 0's lookback)
    DMeta meta[WAVES][kMetaPerWave];  // per-block metadata of the wave's blocks
    // the wave's stream words, loaded densely (then read per lane): its whole
    // payload (1D, where it fits) or one chunk's; 2D/3D: after that the
    // chunk's decoded rows, padded
    uint4 ostage[WAVES][kWaveBuf / 16];
    uint32_t zero[WAVES][2];          // two zero words per wave (read by lanes without a word)
};

// Load of one element per lane. 1D loads cover whole cache lines and stream
// (nontemporal); a 2D/3D group reads 16 or 32 bytes of each of its rows' lines,
// whose other bytes belong to the chunk's next groups, so those loads keep
// the lines cached.
template <int DIM, class T>
__device__ __forceinline__ T load_elem(const T* p) {
    if constexpr (DIM == 1) return __builtin_nontemporal_load(p);
    else return *p;
}

// The wave-uniform part of the loads of a wave's G groups.
template <int DIM, class T, int G>
struct GroupLoads {
    // Loads x[g] for the wave's groups grp0 .. grp0+G-1. FULL: every element
    // lies inside the array (no per-element checks); otherwise out-of-range
    // lanes read 0 and vmask records the valid (group, lane) pairs per group.
    template <bool FULL>
    __device__ __forceinline__ static void run(const T* __restrict__ in, const Geom& geo, const LaneGeo<DIM>& lg,
                                               uint64_t grp0, int lane, T (&x)[G], uint64_t& vmask) {
        if constexpr (FULL) {
            vmask = (G == 64) ? ~0ull : (1ull << G) - 1;
            if constexpr (DIM == 1) {
                const T* p = in + grp0 * 64 + (uint64_t)lane;
#pragma unroll
                for (int g = 0; g < G; ++g) x[g] = load_elem<DIM>(p + 64 * g);
            } else {
                Walker wk;
                wk.init(grp0, geo);
                if (wk.bx + G <= geo.Bx) {
                    // the G blocks are consecutive along x: origins 8 (2D) or 4 (3D) elements apart
                    const T* p = in + block_origin<DIM>(geo, wk) + lg.off;
#pragma unroll
                    for (int g = 0; g < G; ++g) x[g] = load_elem<DIM>(p + (DIM == 2 ? 8 : 4) * g);
                } else {
#pragma unroll
                    for (int g = 0; g < G; ++g) {
                        x[g] = load_elem<DIM>(in + block_origin<DIM>(geo, wk) + lg.off);
                        wk.next(geo);
                    }
                }
            }
        } else {
            vmask = 0;
            Walker wk;
            if constexpr (DIM != 1) wk.init(grp0, geo);
#pragma unroll
            for (int g = 0; g < G; ++g) {
                uint64_t idx;
                const bool valid = elem_of<DIM>(geo, lg, wk, grp0 + g, lane, idx);
                x[g] = valid ? load_elem<DIM>(in + idx) : T(0);
                vmask |= (uint64_t)valid << g;
                if constexpr (DIM != 1) wk.next(geo);
            }
        }
    }
};

// Whether every element of groups grp0 .. grp0+G-1 lies inside the array.
template <int DIM, int G>
__device__ __forceinline__ bool groups_full(const Geom& geo, uint64_t grp0) {
    if constexpr (DIM == 1) return (grp0 + G) * 64 <= geo.n;
    else return geo.aligned && grp0 + G <= geo.nblocks;
}

// Compresses tile `tile` of array `cur` (every wave of the workgroup calls it).
// Each wave runs its kChunks chunks of G groups in turn:
//   1. loads (one element per lane per group; a chunk's G loads are issued
//      before the previous chunk is processed);
//   2. pass 1, per group: quantise, Lorenzo residual, sign|magnitude, bit
//      transpose; nonzero-plane masks go to the block's bookkeeping lane;
//   3. bookkeeping, lane j <-> block j of the chunk: header byte (to LDS, stored
//      after the chunks), size, offset; metadata through LDS;
//   4. the chunk's stream words are written into the wave's LDS staging area
//      exactly as they will appear in the stream.
// The wave totals then combine in LDS, wave 0 runs the lookback (publishing
// the tile's byte count first, then its inclusive prefix), and every wave
// copies its staged payload to the stream with 16-byte stores. A wave whose
// payload did not fit its staging area runs its chunks again and stores the
// words directly.
template <int DIM, int MODE, class T, int WAVES, int LB = kLookbackWords>
__device__ __forceinline__ void compress_tile(const Item& cur, const uint64_t tile, const ScanCtx& sc,
                                              SmemC<WAVES>& sm) {
    constexpr int G = ShapeC<T>::G;              // groups per chunk
    constexpr int S = (DIM == 1) ? 4 : 8;        // bytes per plane
    constexpr int BPG = (DIM == 1) ? 2 : 1;      // blocks per group
    constexpr int NB = BPG * G;                  // blocks per chunk
    static_assert(NB <= 16 && 2 * G <= 16, "shared metadata sized for 16 blocks per chunk");

    const int lane = (int)__lane_id();
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const uint32_t l = (uint32_t)lane & 31, h = (uint32_t)lane >> 5;
    const Geom& geo = cur.geo;
    const T r = (T)cur.param;
    const uint64_t lt = tile - cur.tile0;
    uint8_t* __restrict__ const dst = (uint8_t*)cur.dst;
    const T* __restrict__ const in = (const T*)cur.src;
    const uint64_t wgrp0 = (lt * WAVES + (uint64_t)wave) * (kChunks * G);   // this wave's first group

    LaneGeo<DIM> lg;
    lg.init(lane, geo);
    const int pl = pred_lane<DIM>(lane);
    const int pl_src = pl < 0 ? lane : pl;
    XposeConsts xc;
    xc.init(lane, false);
    // after the transpose, element 0 of a block is bit 7 of its plane words: in
    // both halves for 1D (two blocks), only in the low half for 2D/3D
    const uint32_t not_e0 = (DIM == 1 || h == 0) ? ~0x80u : ~0u;
    // bit 7 of this lane's word is not element 0 of a plane (the sign word, or
    // the high half of a 2D/3D block): kept in selected blocks too
    const uint32_t keep_e0 = (l == 31 || (DIM != 1 && h != 0)) ? 0x80u : 0u;
    const uint32_t key = (l == 31) ? 0u : l + 1u;             // see Meta::b
    const uint32_t wo = word_offset<DIM>(l, h, true);        // this lane's word in a block's payload
    const uint32_t wo_wide = word_offset<DIM>(l, h, false);  // same, 32-plane 2D/3D layout
    // LDS addresses of this lane's word in the staging area (offset 0) and of
    // the spare word past it, where lanes without a word write
    lds_u8* const sbase = (lds_u8*)sm.stage[wave];
    const uint32_t s0 = (uint32_t)(size_t)sbase;
    const uint32_t so = s0 + wo, so_wide = s0 + wo_wide, spare = s0 + kStageBytes;

    uint64_t excl = 0, wexcl = 0, ttot = 0;
    uint32_t wave_tot = 0;
    bool over = false;
    bool lossy = false;   // a 2D/3D outlier block got rate byte 32 (see kWarnLossy)

    // Processes chunk c from xv (vmask: valid elements per group); roll(g) runs
    // right after group g's value is consumed (the fast path loads the next
    // chunk's group g into xv[g] there). DIRECT: store the words at gbase
    // instead of staging them (and write no header bytes).
    auto process = [&](auto directc, uint8_t* const gbase, int c, T (&xv)[G], const uint64_t vmask, auto&& roll) {
        constexpr bool DIRECT = decltype(directc)::value;
        const uint64_t grp0 = wgrp0 + (uint64_t)c * G;

        // ---- pass 1 ---------------------------------------------------------
        // WIDE = false: the branch-free form, which assumes no 2D/3D block
        // holds a magnitude of 2^31 and reports the groups where one does;
        // the chunk is then rerun from memory with WIDE = true.
        uint32_t W[G];
        uint32_t bk1, bk2, bko;   // lane j: block j's nonzero-plane mask (all / e >= 1), |d| of element 0
        uint32_t big;             // bit g: group g's block has a magnitude 2^31 (2D/3D)
        auto residual = [&](const uint32_t q) -> uint32_t {
            if constexpr (MODE == kFixed) {
                return q;
            } else if constexpr (DIM == 1) {
                // q - q[lane - 1], negated from one DPP subtract (the negation is
                // the subtraction |d| needs anyway); element 0 of each block has no
                // predecessor: lane 31's code is not passed on, lane 0 reads 0
                const uint32_t p = ((lane & 31) == 31) ? 0u : q;
                return 0u - sub_prev_wave_shr1(p, q);
            } else {
                uint32_t qp = (uint32_t)__builtin_amdgcn_ds_bpermute(pl_src << 2, (int)q);
                qp = (pl >= 0) ? qp : 0u;
                return q - qp;
            }
        };
        auto pass1 = [&](auto widec, auto&& xval, const uint64_t vm, auto&& roll_fn) -> uint64_t {
            constexpr bool WIDE = decltype(widec)::value;
            bk1 = bk2 = bko = big = 0;
            uint64_t anywide = 0;
            if constexpr (!WIDE) {
                // the words of all groups first, then their transposes side by
                // side (one LDS round trip per stage for kXposeBatch groups),
                // then the plane masks
                uint32_t aor = 0;   // OR of the magnitudes: bit 31 set iff some block needs 32 planes
                auto code = [&](auto gc, uint32_t d) {   // d: group g's residual
                    constexpr int g = decltype(gc)::value;
                    d = ((vm >> g) & 1) ? d : 0u;   // codes and residuals of padding are 0
                    const uint32_t a = ((int32_t)d < 0) ? 0u - d : d;
                    if constexpr (DIM != 1) aor |= a;
                    // sign bit over the low 31 magnitude bits; 1D ranks |-2^31| as 0
                    // (signed running maximum, src/cuSZp_kernels_1D_f32.cu:552), so bit 31 is free
                    W[g] = bfi(0x80000000u, d, a);
                    if constexpr (MODE == kOutlier) {   // |d| of each block's element 0
                        if constexpr (DIM == 1) {
                            bko = wrlane<2 * g>(rdlane(a, 0), bko);
                            bko = wrlane<2 * g + 1>(rdlane(a, 32), bko);
                        } else {
                            bko = wrlane<g>(rdlane(a, 0), bko);
                        }
                    }
                };
                if constexpr (DIM == 1 || MODE == kFixed) {
                    static_for<0, G>([&](auto gc) {
                        constexpr int g = decltype(gc)::value;
                        const uint32_t q = quantize<T>(xval(g), r);   // x = 0 outside the array (q = -1 if r = inf: masked below)
                        roll_fn(g);
                        code(gc, residual(q));
                    });
                } else {
                    // 2D/3D Lorenzo: every group's codes first (W holds them), then
                    // the predecessors' codes of all groups in flight together (one
                    // LDS permute each), then the residuals
                    static_for<0, G>([&](auto gc) {
                        constexpr int g = decltype(gc)::value;
                        W[g] = quantize<T>(xval(g), r);
                        roll_fn(g);
                    });
                    uint32_t QP[G];
                    static_for<0, G>([&](auto gc) {
                        constexpr int g = decltype(gc)::value;
                        QP[g] = (uint32_t)__builtin_amdgcn_ds_bpermute(pl_src << 2, (int)W[g]);
                    });
                    static_for<0, G>([&](auto gc) {
                        constexpr int g = decltype(gc)::value;
                        code(gc, W[g] - ((pl >= 0) ? QP[g] : 0u));
                    });
                }
                if constexpr (DIM != 1) anywide = ballot((int32_t)aor < 0);
                constexpr int XB = kXposeBatch < G ? kXposeBatch : G;
#pragma unroll
                for (int g = 0; g < G; g += XB) xpose32_n<XB>(W + g, xc);
                static_for<0, G>([&](auto gc) {
                    constexpr int g = decltype(gc)::value;
                    const uint64_t m1 = ballot(W[g] != 0) & 0x7fffffff7fffffffull;
                    uint64_t m2 = 0;
                    if constexpr (MODE == kOutlier) {
                        // planes nonzero without element 0 (bit 7 of the words of
                        // planes 0..30; 2D/3D: low half only): W not in {0, 0x80}
                        constexpr uint64_t e0lanes = (DIM == 1) ? 0x7fffffff7fffffffull : 0x7fffffffull;
                        m2 = m1 & ~(ballot(W[g] == 0x80u) & e0lanes);
                    }
                    if constexpr (DIM == 1) {
                        bk1 = wrlane<2 * g>((uint32_t)m1, bk1);
                        bk1 = wrlane<2 * g + 1>((uint32_t)(m1 >> 32), bk1);
                        if constexpr (MODE == kOutlier) {
                            bk2 = wrlane<2 * g>((uint32_t)m2, bk2);
                            bk2 = wrlane<2 * g + 1>((uint32_t)(m2 >> 32), bk2);
                        }
                    } else {
                        bk1 = wrlane<g>((uint32_t)m1 | (uint32_t)(m1 >> 32), bk1);
                        if constexpr (MODE == kOutlier) bk2 = wrlane<g>((uint32_t)m2 | (uint32_t)(m2 >> 32), bk2);
                    }
                });
                return anywide;
            }
            static_for<0, G>([&](auto gc) {
                constexpr int g = decltype(gc)::value;
                const uint32_t q = quantize<T>(xval(g), r);   // x = 0 outside the array (q = -1 if r = inf: masked below)
                roll_fn(g);
                uint32_t d = residual(q);
                d = ((vm >> g) & 1) ? d : 0u;   // codes and residuals of padding are 0
                const uint32_t a = ((int32_t)d < 0) ? 0u - d : d;
                uint64_t m1, m2 = 0;
                const bool wide = ballot((int32_t)a < 0) != 0;
                if (!wide) {
                    W[g] = xpose32(bfi(0x80000000u, d, a), xc);
                    m1 = ballot(W[g] != 0) & 0x7fffffff7fffffffull;
                    if constexpr (MODE == kOutlier) m2 = ballot((W[g] & not_e0) != 0) & 0x7fffffff7fffffffull;
                } else {
                    W[g] = xpose32(a, xc);
                    m1 = ballot(W[g] != 0);
                    if constexpr (MODE == kOutlier) m2 = ballot((W[g] & not_e0) != 0);
                    const uint64_t sg = stream_order64(ballot((int32_t)d < 0));   // the sign words
                    if (lane == 0) {
                        sm.aux[wave][2 * g] = (uint32_t)sg;
                        sm.aux[wave][2 * g + 1] = (uint32_t)(sg >> 32);
                    }
                    big |= 1u << g;
                }
                bk1 = wrlane<g>((uint32_t)m1 | (uint32_t)(m1 >> 32), bk1);
                if constexpr (MODE == kOutlier) {
                    bk2 = wrlane<g>((uint32_t)m2 | (uint32_t)(m2 >> 32), bk2);
                    bko = wrlane<g>(rdlane(a, 0), bko);
                }
            });
            return anywide;
        };
        const uint64_t anywide = pass1(std::false_type{}, [&](int g) -> T { return xv[g]; }, vmask, roll);
        if constexpr (DIM != 1) {
            if (anywide) {   // rare: rerun the chunk with 32-plane blocks
                T xr[G];
                uint64_t vr;
                GroupLoads<DIM, T, G>::template run<false>(in, geo, lg, grp0, lane, xr, vr);
                pass1(std::true_type{}, [&](int g) -> T { return xr[g]; }, vr, [](int) {});
            }
        }

        // ---- bookkeeping: lane j <-> block j of this chunk ---------------------
        const bool blk = lane < NB;
        uint32_t byte;
        BlockView v;
        if constexpr (MODE == kOutlier) {
            if (DIM == 1 || !anywide) {
                v = outlier_code<S>(bitlen(bk1), bitlen(bk2), bitlen(bko), byte);
            } else {
                byte = rate_byte<MODE, S>(bitlen(bk1), bitlen(bk2), bitlen(bko));
                v = view_of<MODE, S>(byte);
                // R1 = 32 (a residual of magnitude 2^31, only from saturated or
                // wrapped codes) is kept as the plain byte 32, which every
                // outlier decoder reads as rate 0 (written by
                // src/cuSZp_kernels_2D_f32.cu:1494-1505, masked with 0x1f by the
                // decoder at :2320): the block decodes to zeros, so the call
                // reports it
                lossy |= ballot(blk && byte == 32u) != 0;
            }
        } else {
            byte = bitlen(bk1);
            v = view_of<MODE, S>(byte);
        }
        const uint32_t sz = blk ? v.size : 0u;
        const uint32_t incl = scan_row16(sz);   // the chunk's NB <= 16 blocks sit in lanes 0..NB-1
        const uint32_t ctot = rdlane(incl, NB - 1);
        const uint32_t off = wave_tot + incl - sz;   // block offset within the wave's payload
        const Meta* const mrow = &sm.meta[wave][DIM == 1 ? h : 0];
        if constexpr (!DIRECT) {
            // header bytes (lane j: block j) go to the stream after the wave's
            // chunks: no global store inside the chunk loop, whose waits for
            // the rolling loads would otherwise also wait for those stores
            if (blk) sm.hdr[wave][c * NB + lane] = (uint8_t)byte;
            const bool put = !over && wave_tot + ctot <= kStageBytes;
            over |= !put;
            if (put) {
                if (blk) sm.meta[wave][lane] = make_meta<MODE>(off, v, sz != 0);
                if constexpr (MODE == kOutlier) {
                    // |d| of element 0 (4 bytes: the sign word written next
                    // covers the bytes past the prefix)
                    if (blk && v.sel) *(lds_u32u*)(sbase + off) = bko;
                }
                // staging: every lane writes every group (lanes without a word
                // of the block write the spare word past the area), so the
                // metadata reads need no branches
                if (DIM == 1 || big == 0) {
                    Meta mnext = mrow[0];
                    static_for<0, G>([&](auto gc) {
                        constexpr int g = decltype(gc)::value;
                        const Meta m = mnext;
                        if constexpr (g + 1 < G) mnext = mrow[BPG * (g + 1)];   // one group ahead
                        uint32_t w = W[g];
                        if constexpr (MODE == kOutlier) w &= m.c | keep_e0;   // element 0 travels in the prefix
                        const uint32_t o = m.has(key) ? m.off() + so : spare;
                        *(lds_u32u*)(size_t)o = w;
                    });
                } else {   // a 2D/3D group of 32 planes in the chunk
                    static_for<0, G>([&](auto gc) {
                        constexpr int g = decltype(gc)::value;
                        const Meta m = mrow[g];
                        const bool wideg = (big >> g) & 1;
                        uint32_t w = W[g];
                        if constexpr (MODE == kOutlier) w &= m.c | keep_e0;
                        const uint32_t o = m.has(key) ? m.off() + (wideg ? so_wide : so) : spare;
                        *(lds_u32u*)(size_t)o = w;
                        if (wideg && m.nz() && lane < 2) *(lds_u32u*)(sbase + m.off() + 4 * lane) = sm.aux[wave][2 * g + lane];
                    });
                }
            }
        } else {
            if (blk) sm.meta[wave][lane] = make_meta<MODE>(off, v, sz != 0);
            if constexpr (MODE == kOutlier) {
                if (blk && v.sel) store_le(gbase + off, bko, v.nb);   // |d| of element 0, little-endian
            }
            static_for<0, G>([&](auto gc) {
                constexpr int g = decltype(gc)::value;
                const Meta m = mrow[BPG * g];
                const bool wideg = (DIM != 1) && ((big >> g) & 1);
                uint32_t w = W[g];
                if constexpr (MODE == kOutlier) w &= m.c | keep_e0;
                if (m.has(key)) store_u32u(gbase + (m.off() + (wideg ? wo_wide : wo)), w);
                if constexpr (DIM != 1) {
                    if (wideg && m.nz() && lane < 2) store_u32u(gbase + (m.off() + 4 * lane), sm.aux[wave][2 * g + lane]);
                }
            });
        }
        wave_tot += ctot;
    };
    // The chunks one by one with bounds-checked loads: tiles with elements
    // past the array, and the direct pass. A chunk past the array's groups is
    // all zero: rate bytes 0 in every mode and no payload, so it only writes
    // its header bytes (1D header padding, 2D/3D pad bytes below H).
    auto checked_chunks = [&](auto directc, uint8_t* const gbase) {
        constexpr bool DIRECT = decltype(directc)::value;
        const uint64_t gvalid = (DIM == 1) ? (geo.n + 63) / 64 : geo.nblocks;
        for (int c = 0; c < kChunks; ++c) {
            const uint64_t grp0 = wgrp0 + (uint64_t)c * G;
            if (grp0 >= gvalid) {
                if (!DIRECT && lane < NB) sm.hdr[wave][c * NB + lane] = 0;
                continue;
            }
            T xv[G];
            uint64_t vmask;
            GroupLoads<DIM, T, G>::template run<false>(in, geo, lg, grp0, lane, xv, vmask);
            process(directc, gbase, c, xv, vmask, [](int) {});
        }
    };

    // ---- pass 0: header bytes and staging ------------------------------------
    // first element of chunk c's first group for this lane; with a fast tile,
    // group g of the chunk starts `step` * g elements later
    constexpr uint64_t step = (DIM == 1) ? 64 : (DIM == 2 ? 8 : 4);
    // fast tile: every element inside the array and (2D/3D) each chunk's
    // blocks consecutive along x
    const bool fast = groups_full<DIM, kChunks * G>(geo, wgrp0) && (DIM == 1 || geo.Bx % G == 0);
    if (fast) {
        // wkn: the chunk whose loads are issued next (2D/3D; a run of G blocks
        // along x, so the walker advances by G)
        Walker wkn;
        if constexpr (DIM != 1) wkn.init(wgrp0, geo);
        auto chunk_ptr = [&](int c) -> const T* {
            if constexpr (DIM == 1) return in + (wgrp0 + (uint64_t)c * G) * 64 + (uint64_t)lane;
            else return in + block_origin<DIM>(geo, wkn) + lg.off;
        };
        T xv[G];
        const T* p = chunk_ptr(0);
#pragma unroll
        for (int g = 0; g < G; ++g) xv[g] = load_elem<DIM>(p + step * g);
        for (int c = 0; c + 1 < kChunks; ++c) {
            // group g of the next chunk loads as soon as group g of this one is coded
            if constexpr (DIM != 1) wkn.step(geo, G);
            const T* const pn = chunk_ptr(c + 1);
            process(std::false_type{}, nullptr, c, xv, ~0ull, [&](int g) { xv[g] = load_elem<DIM>(pn + step * g); });
        }
        process(std::false_type{}, nullptr, kChunks - 1, xv, ~0ull, [](int) {});   // the last chunk loads nothing more
    } else {
        checked_chunks(std::false_type{}, nullptr);
    }

    // ---- the wave's header bytes (their positions do not depend on the
    // offset); 2D/3D pad bytes [nblocks, H) are 0 -------------------------------
    {
        constexpr int NBW = kChunks * NB;
        constexpr int BPL = (NBW + kWaveSize - 1) / kWaveSize;
        static_assert(NBW <= 128, "header bytes of a wave");
        __builtin_amdgcn_wave_barrier();
        const uint64_t wb0 = (DIM == 1) ? 2 * wgrp0 : wgrp0;
#pragma unroll
        for (int k = 0; k < BPL; ++k) {
            const int j = BPL * lane + k;
            const uint64_t b = wb0 + (uint64_t)j;
            if (j < NBW && (DIM == 1 || b < geo.H)) dst[b] = sm.hdr[wave][j];
        }
    }

    // ---- the tile's offset (wave 0) ---------------------------------------------
    if (lane == 0) {
        sm.wtot[wave] = wave_tot;
        sm.over[wave] = over;
        if (lossy) atomicOr(cur.err, kWarnLossy);
    }
    __syncthreads();
#pragma unroll
    for (int w = 0; w < WAVES; ++w) {
        const uint64_t t = sm.wtot[w];
        wexcl += (w < wave) ? t : 0;
        ttot += t;
    }
    if (wave == 0) {
        const uint64_t e = wave_lookback<LB>(sc, tile, cur.tile0, ttot, lane, [&](uint64_t p) {
            return compress_tile_bytes<DIM, MODE, T, kChunks * G, WAVES>((const T*)cur.src, geo, r,
                                                                        p - cur.tile0, lane);
        });
        if (lane == 0) {
            sm.excl = e;
            if (lt == geo.ntiles - 1 && cur.total) *cur.total = geo.H + e + ttot;
        }
    }
    __syncthreads();
    excl = sm.excl;
    if (geo.H + excl + ttot > cur.limit) {   // never write past the caller's buffer
        if (wave == 0 && lane == 0) atomicOr(cur.err, kErrOverflow);
        return;
    }
    if (over) {   // pass 1: run the chunks again, storing directly
        wave_tot = 0;
        checked_chunks(std::true_type{}, dst + uniform64(geo.H + excl + wexcl));
        return;
    }

    // ---- copy the staged payload: 16 bytes per lane, then the tail bytes ---------
    uint8_t* const obase = dst + uniform64(geo.H + excl + wexcl);
    typedef uint32_t v4uu __attribute__((ext_vector_type(4), aligned(1)));
    typedef uint32_t v4ul __attribute__((ext_vector_type(4)));
    const uint32_t n16 = wave_tot & ~15u;
    if (cur.flags & kItemStreamNt) {
        for (uint32_t i = 16u * (uint32_t)lane; i < n16; i += 16u * kWaveSize)
            __builtin_nontemporal_store(*(const v4ul*)&sm.stage[wave][i / 16], (v4uu*)(obase + i));
    } else {
        for (uint32_t i = 16u * (uint32_t)lane; i < n16; i += 16u * kWaveSize)
            *(v4uu*)(obase + i) = *(const v4ul*)&sm.stage[wave][i / 16];
    }
    if ((uint32_t)lane < (wave_tot & 15u)) obase[n16 + lane] = ((const uint8_t*)sm.stage[wave])[n16 + lane];
}

// Decompresses tile `tile` of array `cur` (every wave of the workgroup calls
// it); the tile has the compress tile's shape (kChunks chunks of G groups per
// wave):
//   1. each wave reads the header bytes of its blocks (one or two per lane),
//      scans their sizes once and writes every block's metadata (offset within
//      the wave's payload, lane mask) to LDS;
//   2. the wave's payload offset: from the offset table (OFFS), else the wave
//      totals combine in LDS and wave 0 runs the lookback;
//   3. each wave works through its chunks: every lane loads its stream words
//      (the next chunk's loads are issued while this chunk decodes), transposes
//      them back, rebuilds q along the Lorenzo chains and stores (T)q * eb * 2.
//
// OFFS (single-array launches): the wave's payload offset comes from the
// offset table `offs` written by unit_offsets_kernel: no lookback and no
// workgroup barrier, every wave runs on its own.
template <int DIM, int MODE, class T, int WAVES, int LB = kLookbackWords, bool OFFS = false>
__device__ __forceinline__ void decompress_tile(const Item& cur, const uint64_t tile, const ScanCtx& sc,
                                                SmemD<WAVES>& sm,
                                                const unsigned long long* __restrict__ offs = nullptr) {
    constexpr int G = ShapeC<T>::G;
    constexpr int S = (DIM == 1) ? 4 : 8;
    constexpr int BPG = (DIM == 1) ? 2 : 1;
    constexpr int NB = BPG * G;                         // blocks per chunk
    constexpr int NBW = kChunks * NB;                   // blocks per wave: 128, 64 or 32
    constexpr int BPL = (NBW + kWaveSize - 1) / kWaveSize;   // blocks per bookkeeping lane
    static_assert(NBW <= 2 * kWaveSize && NBW <= kMetaPerWave, "a wave's blocks: at most two per lane");

    const int lane = (int)__lane_id();
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const uint32_t l = (uint32_t)lane & 31, h = (uint32_t)lane >> 5;
    const Geom& geo = cur.geo;
    const uint8_t* __restrict__ cmp = (const uint8_t*)cur.src;
    const uint64_t lt = tile - cur.tile0;
    const uint64_t wgrp0 = (lt * WAVES + (uint64_t)wave) * (kChunks * G);   // this wave's first group
    const uint64_t wb0 = (DIM == 1) ? 2 * wgrp0 : wgrp0;                     // and first block

    // ---- bookkeeping: lane j holds blocks BPL*j .. BPL*j + BPL-1 of the wave ------
    BlockView v[BPL];
    uint32_t sz[BPL];
    bool okb[BPL];
    uint32_t lsum = 0, r32 = 0;
#pragma unroll
    for (int k = 0; k < BPL; ++k) {
        const int j = BPL * lane + k;
        const uint64_t b = wb0 + (uint64_t)j;
        okb[k] = j < NBW && (DIM == 1 || b < geo.nblocks);
        const uint32_t byte = okb[k] ? (uint32_t)cmp[b] : 0u;
        v[k] = view_of<MODE, S>(byte);
        sz[k] = okb[k] ? v[k].size : 0u;
        lsum += sz[k];
        r32 |= (uint32_t)(byte >= 32);
    }
    const uint32_t incl = scan_wave(lsum);                 // < 2^32; a cross-lane scan: all lanes
    const uint32_t wtot = rdlane(incl, kWaveSize - 1);
    const uint32_t excl0 = incl - lsum;                    // offset of this lane's first block
    {
        uint32_t off = excl0;                              // offsets within the wave's payload
#pragma unroll
        for (int k = 0; k < BPL; ++k) {
            const int j = BPL * lane + k;
            // every entry is written: blocks past the array get an empty lane mask
            if (j < NBW) sm.meta[wave][j] = make_dmeta<DIM>(off, v[k], okb[k] && sz[k] != 0);
            off += sz[k];
        }
    }
    // 2D/3D blocks of 32 planes (not in outlier streams) take the general loop
    bool wide_blocks = false;
    if constexpr (DIM != 1 && MODE != kOutlier) wide_blocks = ballot(r32 != 0) != 0;

    uint64_t wave_base;
    if constexpr (OFFS) {
        (void)sc;
        wave_base = uniform64(offs[lt * WAVES + (uint64_t)wave]);
        __builtin_amdgcn_wave_barrier();   // the metadata in LDS is this wave's own
    } else {
        if (lane == 0) sm.wtot[wave] = wtot;
        __syncthreads();
        uint64_t wexcl = 0, ttot = 0;
#pragma unroll
        for (int w = 0; w < WAVES; ++w) {
            const uint64_t t = sm.wtot[w];
            wexcl += (w < wave) ? t : 0;
            ttot += t;
        }
        // wave-uniform: kept in scalar registers across the lookback
        wexcl = uniform64(wexcl);
        ttot = uniform64(ttot);
        if (wave == 0) {
            const uint64_t e = wave_lookback<LB>(sc, tile, cur.tile0, ttot, lane, [&](uint64_t p) {
                return header_tile_bytes<DIM, MODE, kChunks * G, WAVES>(cmp, geo, p - cur.tile0, lane);
            });
            if (lane == 0) sm.excl = e;
        }
        __syncthreads();
        wave_base = geo.H + sm.excl + wexcl;
    }
    const uint8_t* const wbase = cmp + uniform64(wave_base);
    const bool intact = wave_base + wtot <= cur.limit;
    if (!intact && lane == 0) atomicOr(cur.err, kErrTruncated);
    // The wave's stream words go through a buffer resource over its payload:
    // 32-bit offsets, and an offset past the range (bit 31 set) reads 0, so
    // lanes without a word need no address select and no masking; a truncated
    // stream gets an empty range (every word reads 0).
    const uint64_t span = cur.limit - wave_base;
    const int nrec = __builtin_amdgcn_readfirstlane(intact ? (int)(span < 0x7fffffffull ? span : 0x7fffffffull) : 0);
    const __amdgpu_buffer_rsrc_t rsrc = __builtin_amdgcn_make_buffer_rsrc((void*)wbase, 0, nrec, kBufferWord3);

    // outlier streams: |d0| prefix word of each selected block of this lane
    // (masked with PK at the use, so nothing waits for the load here); a
    // selected block starts with nb <= 4 prefix bytes and then a 4-byte sign
    // word, so one u32 read at the block start stays inside the block
    uint32_t PR[BPL], PK[BPL], PV[BPL];   // PV = PR & PK, formed once the first chunk's loads are issued
#pragma unroll
    for (int k = 0; k < BPL; ++k) PR[k] = PK[k] = PV[k] = 0;
    auto prefixes_ready = [&]() {
#pragma unroll
        for (int k = 0; k < BPL; ++k) PV[k] = PR[k] & PK[k];
    };
    if constexpr (MODE == kOutlier) {
        uint32_t off = incl - lsum;
#pragma unroll
        for (int k = 0; k < BPL; ++k) {
            const bool pre = okb[k] && v[k].sel;
            PR[k] = (uint32_t)__builtin_amdgcn_raw_buffer_load_b32(rsrc, pre ? off : kAbsent, 0, 0);
            PK[k] = v[k].nb < 4 ? (1u << (8 * v[k].nb)) - 1 : ~0u;
            off += sz[k];
        }
    }

    XposeConsts xc;
    xc.init(lane, true);
    LaneGeo<DIM> lg;
    lg.init(lane, geo);
    T* __restrict__ out = (T*)cur.dst;
    const T eb = (T)cur.param;
    const uint32_t wo = word_offset<DIM>(l, h, true);
    const uint32_t wo_wide = word_offset<DIM>(l, h, false);
    const uint32_t keyx = lane_keyx(l);   // see DMeta
    constexpr uint64_t step = (DIM == 1) ? 64 : (DIM == 2 ? 8 : 4);
    // fast tile: every element inside the array and (2D/3D) each chunk's blocks
    // consecutive along x
    const bool fast = groups_full<DIM, kChunks * G>(geo, wgrp0) && (DIM == 1 || geo.Bx % G == 0);

    // metadata of this lane's block in group g of chunk c
    auto meta = [&](int c, int g) -> DMeta { return sm.meta[wave][c * NB + BPG * g + (DIM == 1 ? (int)h : 0)]; };
    // stream word of this lane in group g of chunk c (0 for a lane without a
    // word and for every lane of a truncated stream); no 32-plane blocks
    auto word = [&](int c, int g) -> uint32_t {
        return (uint32_t)__builtin_amdgcn_raw_buffer_load_b32(rsrc, meta(c, g).word_at(keyx, wo), 0, 0);
    };
    // the same where 2D/3D blocks of 32 planes may occur
    auto word_any = [&](int c, int g) -> uint32_t {
        const DMeta m = meta(c, g);
        const bool wideg = (DIM != 1) && (MODE != kOutlier) && m.wide();
        const uint32_t o = m.off() + (wideg ? wo_wide : wo);
        return (uint32_t)__builtin_amdgcn_raw_buffer_load_b32(rsrc, m.has(keyx) ? o : kAbsent, 0, 0);
    };
    // residuals of group g of chunk c from its word after the transpose (lane
    // e: element e)
    auto residual_x = [&](int c, int g, uint32_t x, auto widec) -> uint32_t {
        constexpr bool WIDE = decltype(widec)::value;   // 2D/3D 32-plane blocks possible
        const DMeta m = meta(c, g);
        uint32_t a = x & 0x7fffffffu, neg = x >> 31;   // lane l: sign in bit 31 over the magnitude bits
        if constexpr (WIDE) {
            if (m.wide()) {   // 32 planes: the sign words sit in their own two words
                const uint32_t sw = (intact && m.nz()) ? load_u32u(wbase + (m.off() + 4 * h)) : 0u;
                neg = (sw >> (l ^ 7)) & 1;   // element l's sign: stream bit l^7
                a = x;
            }
        }
        if constexpr (MODE == kOutlier) {   // element 0 of a selected block (0 when not intact)
            // block c*NB + BPG*g (+ h in 1D) is held by bookkeeping lane
            // (block / BPL), slot (block % BPL)
            uint32_t e0;
            if constexpr (BPL == 2) {
                static_assert(DIM == 1 && NB == 2 * G, "two blocks per lane: 1D f32");
                const int ln = c * G + g;
                e0 = h ? rdlane(PV[BPL - 1], ln) : rdlane(PV[0], ln);
            } else if constexpr (DIM == 1) {
                e0 = h ? rdlane(PV[0], c * NB + 2 * g + 1) : rdlane(PV[0], c * NB + 2 * g);
            } else {
                e0 = rdlane(PV[0], c * NB + g);
            }
            a = (m.sel() && l == 0 && (DIM == 1 || h == 0)) ? e0 : a;
        }
        return neg ? 0u - a : a;
    };
    // the value of code q (a block without payload, or of a truncated stream,
    // decoded all-zero words: q = 0)
    auto value_of = [&](uint32_t q) -> T { return (T)(int32_t)q * eb * T(2); };   // src/cuSZp_kernels_1D_f32.cu:490
    auto decode_x = [&](int c, int g, uint32_t x, auto widec) -> T {
        const uint32_t d = residual_x(c, g, x, widec);
        if constexpr (MODE == kFixed) return value_of(d);
        else return value_of(inverse_lorenzo<DIM>(d, lane));
    };
    // decodes group g of chunk c from its loaded word w
    auto decode = [&](int c, int g, uint32_t w, auto widec) -> T { return decode_x(c, g, xpose32(w, xc), widec); };

    if (fast && !wide_blocks) {
        // ---- every element inside the array, chunks along x: no branches ----------
        // Chunks go in pairs through two word buffers: a chunk's loads are
        // issued before the stores of the chunk decoded meanwhile (loads and
        // stores retire in order on this GPU, so a load issued after a store
        // would also wait for that store). The values are stored nontemporally:
        // the output is written once, and dirty lines of it left in L2 and the
        // last-level cache would cost the next kernel their write-back.
        Walker wkc;   // 2D/3D: the chunk being decoded
        if constexpr (DIM != 1) wkc.init(wgrp0, geo);
        // 2D/3D output staging: block width BW, rows of RL values, row stride RS
        // (values) in LDS; lane e of group g goes to row ROW, column g * BW + COL.
        // Read back, lane j (and j + 64) takes 16 bytes: row j / VPR, vector j % VPR.
        constexpr int BW = (DIM == 3) ? 4 : 8;
        constexpr int RL = G * BW;
        constexpr int RS = RL + kOutPad / (int)sizeof(T);
        constexpr int VPR = RL * (int)sizeof(T) / 16;          // 16-byte vectors per row
        constexpr int NR = 64 / BW;                            // rows: 16 (3D) or 8 (2D)
        static_assert(NR * RS * (int)sizeof(T) <= kOutStage, "output staging too small");
        static_assert(2 * 64 * 16 == NR * RL * (int)sizeof(T), "two 16-byte vectors per lane");
        uint32_t st_w = 0, st_r = 0;   // LDS byte offsets: this lane's value / first vector
        uint64_t roff = 0, rdelta = 0;  // global element offsets of the rows read back
        if constexpr (DIM != 1) {
            uint32_t rw, cw;
            if constexpr (DIM == 3) {
                rw = lg.zz * 4 + lg.yy;
                cw = lg.xx;
            } else {
                rw = lg.r;
                cw = lg.c;
            }
            st_w = (rw * RS + cw) * (uint32_t)sizeof(T);
            const uint32_t rr = (uint32_t)lane / VPR, cv = (uint32_t)lane % VPR;
            st_r = rr * RS * (uint32_t)sizeof(T) + cv * 16;
            // the second vector is 64 / VPR rows further down
            if constexpr (DIM == 3) {
                roff = (uint64_t)(rr >> 2) * geo.XY + (uint64_t)(rr & 3) * geo.X + cv * (16 / sizeof(T));
                rdelta = (uint64_t)(64 / VPR / 4) * geo.XY;
            } else {
                roff = (uint64_t)rr * geo.X + cv * (16 / sizeof(T));
                rdelta = (uint64_t)(64 / VPR) * geo.X;
            }
        }
        auto decode_chunk = [&](int c, const uint32_t (&W)[G]) {
            uint32_t X[G];
#pragma unroll
            for (int g = 0; g < G; ++g) X[g] = W[g];
            constexpr int XB = kXposeBatch < G ? kXposeBatch : G;
#pragma unroll
            for (int g = 0; g < G; g += XB) xpose32_n<XB>(X + g, xc);
            // X[g]: residuals, then (batch by batch) codes
#pragma unroll
            for (int g = 0; g < G; ++g) X[g] = residual_x(c, g, X[g], std::false_type{});
            if constexpr (MODE != kFixed) {
#pragma unroll
                for (int g = 0; g < G; g += XB) inverse_lorenzo_n<DIM, XB>(X + g, lane);
            }
            if constexpr (DIM == 1) {
                T* const orun = out + (wgrp0 + (uint64_t)c * G) * 64 + (uint64_t)lane;
#pragma unroll
                for (int g = 0; g < G; ++g) __builtin_nontemporal_store(value_of(X[g]), orun + step * g);
            } else {
                lds_u8* const stg = (lds_u8*)sm.ostage[wave];
#pragma unroll
                for (int g = 0; g < G; ++g)
                    *(__attribute__((address_space(3))) T*)(stg + st_w + g * BW * sizeof(T)) = value_of(X[g]);
                __builtin_amdgcn_wave_barrier();
                typedef uint32_t v4u __attribute__((ext_vector_type(4)));
                typedef uint32_t v4uu __attribute__((ext_vector_type(4), aligned(1)));
                const v4u v0 = *(__attribute__((address_space(3))) v4u*)(stg + st_r);
                const v4u v1 = *(__attribute__((address_space(3))) v4u*)(stg + st_r + (64 / VPR) * RS * sizeof(T));
                __builtin_amdgcn_wave_barrier();
                T* const ob = out + block_origin<DIM>(geo, wkc);
                __builtin_nontemporal_store(v0, (v4uu*)(ob + roff));
                __builtin_nontemporal_store(v1, (v4uu*)(ob + roff + rdelta));
                wkc.step(geo, G);
            }
        };
        // Stream words. A chunk's payload is contiguous: where every chunk of the
        // wave spans at most kDenseBytes from its 4-byte-aligned start, it is read
        // with one or two dense 16-byte-per-lane loads into the wave's LDS buffer
        // and each lane reads its words from there (lanes without a word read a
        // zero word). Loads with some lanes past the buffer range or 14-of-64
        // active lanes cost far more memory-pipeline time than dense ones.
        // Otherwise each lane loads its own word from global memory.
        uint32_t cs[kChunks + 1];   // chunk starts within the wave's payload (wave-uniform)
#pragma unroll
        for (int c = 0; c < kChunks; ++c) cs[c] = rdlane(excl0, c * NB / BPL);
        cs[kChunks] = wtot;
        // Buffer loads check their range per dword, so the dwords holding the
        // payload's last bytes must lie inside it: an outlier stream whose length
        // is not a multiple of 4 ends inside a dword, and its last wave keeps the
        // per-lane loads (a truncated stream has an empty range: the same).
        const bool whole = ((wtot + 3u) & ~3u) <= (uint32_t)nrec;   // every payload dword inside the range
        bool dense = whole;
#pragma unroll
        for (int c = 0; c < kChunks; ++c) dense &= cs[c + 1] - (cs[c] & ~3u) <= (uint32_t)kDenseBytes;
        // 1D: the wave's whole payload in one burst of loads before its first
        // chunk, where it fits the wave's buffer. A compute-free kernel with the
        // 1D decode's traffic (2^30 f32 written, 0.94 GB read) ran 2.5-4 % faster
        // with each wave's reads in one burst than with a read per chunk.
        const bool burst = DIM == 1 && whole && wtot <= (uint32_t)kWaveBuf;
        if (burst) {
            typedef uint32_t v4u __attribute__((ext_vector_type(4)));
            constexpr int NL = (kWaveBuf + 1023) / 1024;
            lds_u8* const buf = (lds_u8*)sm.ostage[wave];
            if (lane < 2) sm.zero[wave][lane] = 0;
            const lds_u8* const zw = (const lds_u8*)&sm.zero[wave][0];
            v4u P[NL];
#pragma unroll
            for (int k = 0; k < NL; ++k) {
                // 1 KiB per load; none past the payload; lanes past the buffer read nothing
                const uint32_t o = 1024u * (uint32_t)k + 16u * (uint32_t)lane;
                P[k] = __builtin_amdgcn_raw_buffer_load_b128(
                    rsrc, (1024u * (uint32_t)k < wtot && o < (uint32_t)kWaveBuf) ? o : kAbsent, 0, 0);
            }
            prefixes_ready();
#pragma unroll
            for (int k = 0; k < NL; ++k)
                if (1024 * k + 16 * lane < kWaveBuf) ((__attribute__((address_space(3))) v4u*)buf)[64 * k + lane] = P[k];
            __builtin_amdgcn_wave_barrier();
            // this lane's word of group g of chunk c (block offsets are payload offsets)
            auto word_buf = [&](int c, int g) -> uint32_t {
                const DMeta m = meta(c, g);
                typedef __attribute__((address_space(3))) const uint32_t lds_cu32;
                if constexpr (MODE != kOutlier) {
                    return *(lds_cu32*)(m.has(keyx) ? buf + (m.a + wo) : zw);
                } else {
                    // (v_alignbyte uses the low two bits of its shift operand)
                    const uint32_t o = m.a + wo;
                    lds_cu32* const q = (lds_cu32*)(m.has(keyx) ? buf + (o & ~3u) : zw);
                    return __builtin_amdgcn_alignbyte(q[1], q[0], o);
                }
            };
            static_for<0, kChunks>([&](auto cc) {
                constexpr int c = decltype(cc)::value;
                uint32_t W[G];
#pragma unroll
                for (int g = 0; g < G; ++g) W[g] = word_buf(c, g);
                decode_chunk(c, W);
            });
        } else if (dense) {
            typedef uint32_t v4u __attribute__((ext_vector_type(4)));
            lds_u8* const buf = (lds_u8*)sm.ostage[wave];
            if (lane < 2) sm.zero[wave][lane] = 0;
            const lds_u8* const zw = (const lds_u8*)&sm.zero[wave][0];
            // the chunk's payload from its aligned start; the second KiB only when
            // needed (else an offset past every range: no memory access)
            auto load_chunk = [&](int c, v4u (&P)[2]) {
                const uint32_t a0 = cs[c] & ~3u;
                P[0] = __builtin_amdgcn_raw_buffer_load_b128(rsrc, a0 + 16u * (uint32_t)lane, 0, 0);
                P[1] = __builtin_amdgcn_raw_buffer_load_b128(
                    rsrc, cs[c + 1] - a0 > 1024u ? a0 + 1024u + 16u * (uint32_t)lane : kAbsent, 0, 0);
            };
            auto stage_chunk = [&](const v4u (&P)[2]) {
                __builtin_amdgcn_wave_barrier();   // the previous chunk's reads of the buffer come first
                ((__attribute__((address_space(3))) v4u*)buf)[lane] = P[0];
                ((__attribute__((address_space(3))) v4u*)buf)[64 + lane] = P[1];
                __builtin_amdgcn_wave_barrier();
            };
            auto word_lds = [&](int c, int g) -> uint32_t {
                const DMeta m = meta(c, g);
                typedef __attribute__((address_space(3))) const uint32_t lds_cu32;
                if constexpr (MODE != kOutlier) {
                    // fixed/plain blocks are whole 4-byte words: an aligned read
                    const uint32_t o = m.a + wo - (cs[c] & ~3u);
                    return *(lds_cu32*)(m.has(keyx) ? buf + o : zw);
                } else {
                    // outlier prefixes shift later blocks by 1-3 bytes: the two
                    // aligned words around the word, then a byte funnel shift
                    // (v_alignbyte uses the low two bits of its shift operand)
                    const uint32_t o = m.a + wo - (cs[c] & ~3u);
                    lds_cu32* const q = (lds_cu32*)(m.has(keyx) ? buf + (o & ~3u) : zw);
                    return __builtin_amdgcn_alignbyte(q[1], q[0], o);
                }
            };
            v4u PA[2], PB[2];
            load_chunk(0, PA);
            load_chunk(1, PB);
            prefixes_ready();
            static_for<0, kChunks>([&](auto cc) {
                constexpr int c = decltype(cc)::value;
                v4u (&P)[2] = (c % 2 == 0) ? PA : PB;
                stage_chunk(P);
                if constexpr (c + 2 < kChunks) load_chunk(c + 2, P);
                uint32_t W[G];
#pragma unroll
                for (int g = 0; g < G; ++g) W[g] = word_lds(c, g);
                decode_chunk(c, W);
            });
        } else {
            static_assert(kChunks % 2 == 0, "chunks go in pairs");
            // Fully unrolled: in a loop, the wait for a chunk's words at the loop
            // head would be merged with the path from before the loop (where no
            // stores follow the loads) and so count the previous chunk's stores as
            // well, putting their completion on the critical path.
            uint32_t WA[G], WB[G];
#pragma unroll
            for (int g = 0; g < G; ++g) WA[g] = word(0, g);
            prefixes_ready();
#pragma unroll
            for (int c = 0; c + 2 < kChunks; c += 2) {
#pragma unroll
                for (int g = 0; g < G; ++g) WB[g] = word(c + 1, g);
                decode_chunk(c, WA);
#pragma unroll
                for (int g = 0; g < G; ++g) WA[g] = word(c + 2, g);
                decode_chunk(c + 1, WB);
            }
            // the last pair: nothing more to load after it
#pragma unroll
            for (int g = 0; g < G; ++g) WB[g] = word(kChunks - 1, g);
            decode_chunk(kChunks - 2, WA);
            decode_chunk(kChunks - 1, WB);
        }
    } else {
        // ---- general: per-element checks, 32-plane blocks -------------------------
        Walker wk;
        if constexpr (DIM != 1) wk.init(wgrp0, geo);
        const uint64_t gvalid = (DIM == 1) ? (geo.n + 63) / 64 : geo.nblocks;   // groups holding elements
        prefixes_ready();
        for (int c = 0; c < kChunks; ++c) {
            if (wgrp0 + (uint64_t)c * G >= gvalid) break;   // the rest of the wave writes nothing
            uint32_t W[G];
#pragma unroll
            for (int g = 0; g < G; ++g) W[g] = word_any(c, g);
            const uint64_t grp0 = wgrp0 + (uint64_t)c * G;
#pragma unroll
            for (int g = 0; g < G; ++g) {
                const T val = decode(c, g, W[g], std::true_type{});
                uint64_t idx;
                if (elem_of<DIM>(geo, lg, wk, grp0 + g, lane, idx)) out[idx] = val;
                if constexpr (DIM != 1) wk.next(geo);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// single-array kernels: one workgroup per tile, tile = base + blockIdx.x (tiles
// are numbered in launch order; arrays of more than one grid's tiles take
// several launches)
// ---------------------------------------------------------------------------
template <int DIM, int MODE, class T, int WAVES>
__global__ void __launch_bounds__(WAVES * kWaveSize, kCompressWavesPerSimd)
compress_kernel(const Item item, const ScanCtx sc, const uint64_t base) {
    __shared__ SmemC<WAVES> sm;
    // a compress tile's predecessors have mostly published by the time its
    // payload is staged: the 128-tile window (2 words per lane) is enough
    compress_tile<DIM, MODE, T, WAVES, 2>(item, base + blockIdx.x, sc, sm);
}

// Waves take their payload offsets from `offs` (unit_offsets_kernel, launched
// before on the same stream); sc is unused.
template <int DIM, int MODE, class T, int WAVES>
__global__ void __launch_bounds__(WAVES * kWaveSize, kDecompressWavesPerSimd)
decompress_kernel(const Item item, const ScanCtx sc, const unsigned long long* __restrict__ offs,
                  const uint64_t base) {
    __shared__ SmemD<WAVES> sm;
    decompress_tile<DIM, MODE, T, WAVES, kLookbackWords, true>(item, base + blockIdx.x, sc, sm, offs);
}

// ---------------------------------------------------------------------------
// batched kernels: the arrays of a table, of any variant, in one launch with
// one workgroup per tile of the concatenated tile space. A tile's array is
// found by binary search on tile0, and each array's tiles form their own
// lookback segment, so every array's stream is exactly its single-array
// stream.
// ---------------------------------------------------------------------------
#define CUSZP_BATCH_CASES(RUN)                                                        \
    case 0: RUN(1, kFixed, float); break;                                             \
    case 1: RUN(1, kPlain, float); break;                                             \
    case 2: RUN(1, kOutlier, float); break;                                           \
    case 3: RUN(2, kFixed, float); break;                                             \
    case 4: RUN(2, kPlain, float); break;                                             \
    case 5: RUN(2, kOutlier, float); break;                                           \
    case 6: RUN(3, kFixed, float); break;                                             \
    case 7: RUN(3, kPlain, float); break;                                             \
    case 8: RUN(3, kOutlier, float); break;                                           \
    case 9: RUN(1, kFixed, double); break;                                            \
    case 10: RUN(1, kPlain, double); break;                                           \
    case 11: RUN(1, kOutlier, double); break;                                         \
    case 12: RUN(2, kFixed, double); break;                                           \
    case 13: RUN(2, kPlain, double); break;                                           \
    case 14: RUN(2, kOutlier, double); break;                                         \
    case 15: RUN(3, kFixed, double); break;                                           \
    case 16: RUN(3, kPlain, double); break;                                           \
    case 17: RUN(3, kOutlier, double); break;                                         \
    default: break;

template <int WAVES, int LB>
__global__ void __launch_bounds__(WAVES * kWaveSize, kCompressWavesPerSimd)
compress_batch_kernel(const Item* __restrict__ items, uint32_t nitems, ScanCtx sc,
                      const unsigned long long* __restrict__ epoch, const uint64_t base) {
    __shared__ SmemC<WAVES> sm;
    sc.tag = (epoch[0] & kEpochMask) << kEpochShift;
    const ItemTable src{items, nitems};
    const uint64_t tile = base + blockIdx.x;
    const Item cur = src.get(src.locate(tile, 0));
#define CUSZP_RUN_C(D, M, T) compress_tile<D, M, T, WAVES, LB>(cur, tile, sc, sm)
    switch (cur.variant) { CUSZP_BATCH_CASES(CUSZP_RUN_C) }
#undef CUSZP_RUN_C
}

template <int WAVES, int LB>
__global__ void __launch_bounds__(WAVES * kWaveSize, kDecompressWavesPerSimd)
decompress_batch_kernel(const Item* __restrict__ items, uint32_t nitems, ScanCtx sc,
                        const unsigned long long* __restrict__ epoch, const uint64_t base) {
    __shared__ SmemD<WAVES> sm;
    sc.tag = (epoch[0] & kEpochMask) << kEpochShift;
    const ItemTable src{items, nitems};
    const uint64_t tile = base + blockIdx.x;
    const Item cur = src.get(src.locate(tile, 0));
#define CUSZP_RUN_D(D, M, T) decompress_tile<D, M, T, WAVES, LB>(cur, tile, sc, sm)
    switch (cur.variant) { CUSZP_BATCH_CASES(CUSZP_RUN_D) }
#undef CUSZP_RUN_D
}

}  // namespace cuszp
