// Kernel instantiations for float: {1D,2D,3D} x {fixed,plain,outlier}, compress and decompress.
#include "launch.hpp"

namespace cuszp {

template <int D, int M, int W>
static hipError_t launch_c(const Item& item, const ScanCtx& sc, hipStream_t st) {
    return launch_tiles(compress_kernel<D, M, float, W>, W * kWaveSize, item.geo.ntiles, st,
                             item, sc);
}
template <int D, int M, int W>
static hipError_t launch_d(const Item& item, const ScanCtx& sc, const unsigned long long* offs, hipStream_t st) {
    return launch_tiles(decompress_kernel<D, M, float, W>, W * kWaveSize, item.geo.ntiles, st, item, sc, offs);
}

// offset pass: S bytes per plane (1D 4, 2D/3D 8), UB blocks per decode wave
template <int M, int S, int UB>
static hipError_t launch_o(const Item& item, uint64_t nunits, unsigned long long* offs, const ScanCtx& sc,
                           hipStream_t st) {
    const uint64_t limit = S == 4 ? item.geo.H : item.geo.nblocks;   // 2D/3D header pad bytes hold no block
    const uint64_t nseg = (nunits * UB + kPrefixSegBytes - 1) / kPrefixSegBytes;
    return launch_tiles(unit_offsets_kernel<M, S, UB>, kPrefixThreads, nseg, st, (const uint8_t*)item.src, limit,
                        item.geo.H, nunits, offs, (unsigned long long*)nullptr, sc, nseg);
}

#include "dispatch.inc"

hipError_t launch_compress_f32(int dim, int mode, int waves, const Item& item, const ScanCtx& sc, hipStream_t st) {
    CUSZP_DISPATCH(launch_c, item, sc, st)
}

hipError_t launch_decompress_f32(int dim, int mode, int waves, const Item& item, const ScanCtx& sc,
                                 const unsigned long long* offs, hipStream_t st) {
    CUSZP_DISPATCH(launch_d, item, sc, offs, st)
}

hipError_t launch_unit_offsets_f32(int dim, int mode, const Item& item, uint64_t nunits, unsigned long long* offs,
                                   const ScanCtx& sc, hipStream_t st) {
    static_assert(128 == kChunks * ShapeC<float>::G * 2 && 64 == kChunks * ShapeC<float>::G, "decode wave units");
    switch ((dim == 1 ? 0 : 3) + mode) {
        case 0: return launch_o<kFixed, 4, 128>(item, nunits, offs, sc, st);
        case 1: return launch_o<kPlain, 4, 128>(item, nunits, offs, sc, st);
        case 2: return launch_o<kOutlier, 4, 128>(item, nunits, offs, sc, st);
        case 3: return launch_o<kFixed, 8, 64>(item, nunits, offs, sc, st);
        case 4: return launch_o<kPlain, 8, 64>(item, nunits, offs, sc, st);
        case 5: return launch_o<kOutlier, 8, 64>(item, nunits, offs, sc, st);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace cuszp
