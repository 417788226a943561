// Kernel instantiations for float: {1D,2D,3D} x {fixed,plain,outlier}, compress and decompress.
#include "launch.hpp"

namespace cuszp {

template <int D, int M, int W>
static hipError_t launch_c(const Item& item, const ScanCtx& sc, hipStream_t st) {
    return launch_tiles(compress_kernel<D, M, float, W>, W * kWaveSize, item.geo.ntiles, st,
                             item, sc);
}
template <int D, int M, int W>
static hipError_t launch_d(const Item& item, const ScanCtx& sc, hipStream_t st) {
    return launch_tiles(decompress_kernel<D, M, float, W>, W * kWaveSize, item.geo.ntiles, st,
                             item, sc);
}

#include "dispatch.inc"

hipError_t launch_compress_f32(int dim, int mode, int waves, const Item& item, const ScanCtx& sc, hipStream_t st) {
    CUSZP_DISPATCH(launch_c, item, sc, st)
}

hipError_t launch_decompress_f32(int dim, int mode, int waves, const Item& item, const ScanCtx& sc, hipStream_t st) {
    CUSZP_DISPATCH(launch_d, item, sc, st)
}

}  // namespace cuszp
