// Kernel instantiations for float: {1D,2D,3D} x {fixed,plain,outlier}, compress and decompress.
#include "launch.hpp"

namespace cuszp {

template <int D, int M, int W>
static hipError_t launch_c(const float* in, uint8_t* out, const Geom& g, float r, const ScanArgs& sa,
                           uint64_t cap, hipStream_t st) {
    return launch_persistent(compress_kernel<D, M, float, kGroupsPerWave, W>, W * kWaveSize, g.ntiles, st,
                             in, out, g, r, sa, cap);
}
template <int D, int M, int W>
static hipError_t launch_d(float* out, const uint8_t* cmp, const Geom& g, float eb, const ScanArgs& sa,
                           uint64_t cmp_size, hipStream_t st) {
    return launch_persistent(decompress_kernel<D, M, float, kGroupsPerWave, W>, W * kWaveSize, g.ntiles, st,
                             out, cmp, g, eb, sa, cmp_size);
}

#define CUSZP_DISPATCH_W(FN, D, M, ...)                                  \
    switch (waves) {                                                     \
        case 4: return FN<D, M, 4>(__VA_ARGS__);                         \
        case 8: return FN<D, M, 8>(__VA_ARGS__);                         \
        case 16: return FN<D, M, 16>(__VA_ARGS__);                       \
        default: return hipErrorInvalidValue;                            \
    }
#define CUSZP_DISPATCH(FN, ...)                                          \
    switch (dim * 3 + mode) {                                            \
        case 3: CUSZP_DISPATCH_W(FN, 1, kFixed, __VA_ARGS__) \
        case 4: CUSZP_DISPATCH_W(FN, 1, kPlain, __VA_ARGS__) \
        case 5: CUSZP_DISPATCH_W(FN, 1, kOutlier, __VA_ARGS__) \
        case 6: CUSZP_DISPATCH_W(FN, 2, kFixed, __VA_ARGS__) \
        case 7: CUSZP_DISPATCH_W(FN, 2, kPlain, __VA_ARGS__) \
        case 8: CUSZP_DISPATCH_W(FN, 2, kOutlier, __VA_ARGS__) \
        case 9: CUSZP_DISPATCH_W(FN, 3, kFixed, __VA_ARGS__) \
        case 10: CUSZP_DISPATCH_W(FN, 3, kPlain, __VA_ARGS__) \
        case 11: CUSZP_DISPATCH_W(FN, 3, kOutlier, __VA_ARGS__) \
        default: return hipErrorInvalidValue;                            \
    }

hipError_t launch_compress_f32(int dim, int mode, int waves, const float* in, uint8_t* out, const Geom& g,
                               float eb, const ScanArgs& sa, uint64_t cap, hipStream_t st) {
    const float r = float(0.5f) / eb;   // 0.5f/eb rounded on the host (src/cuSZp_kernels_1D_f32.cu:48)
    CUSZP_DISPATCH(launch_c, in, out, g, r, sa, cap, st)
}

hipError_t launch_decompress_f32(int dim, int mode, int waves, float* out, const uint8_t* cmp, const Geom& g,
                                 float eb, const ScanArgs& sa, uint64_t cmp_size, hipStream_t st) {
    CUSZP_DISPATCH(launch_d, out, cmp, g, eb, sa, cmp_size, st)
}

}  // namespace cuszp
