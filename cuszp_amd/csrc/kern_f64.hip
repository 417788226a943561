// Kernel instantiations for double: {1D,2D,3D} x {fixed,plain,outlier}, compress and decompress.
#include "launch.hpp"

namespace cuszp {

template <int D, int M>
static void launch_c(const double* in, uint8_t* out, const Geom& g, double r, const ScanArgs& sa,
                     uint64_t cap, hipStream_t st) {
    hipLaunchKernelGGL((compress_kernel<D, M, double, kGroupsPerWave>), dim3((unsigned)g.ntiles),
                       dim3(kThreads), 0, st, in, out, g, r, sa, cap);
}
template <int D, int M>
static void launch_d(double* out, const uint8_t* cmp, const Geom& g, double eb, const ScanArgs& sa,
                     uint64_t cmp_size, hipStream_t st) {
    hipLaunchKernelGGL((decompress_kernel<D, M, double, kGroupsPerWave>), dim3((unsigned)g.ntiles),
                       dim3(kThreads), 0, st, out, cmp, g, eb, sa, cmp_size);
}

#define CUSZP_DISPATCH(FN, ...)                                   \
    switch (dim * 3 + mode) {                                     \
        case 3: FN<1, kFixed>(__VA_ARGS__); break;                \
        case 4: FN<1, kPlain>(__VA_ARGS__); break;                \
        case 5: FN<1, kOutlier>(__VA_ARGS__); break;              \
        case 6: FN<2, kFixed>(__VA_ARGS__); break;                \
        case 7: FN<2, kPlain>(__VA_ARGS__); break;                \
        case 8: FN<2, kOutlier>(__VA_ARGS__); break;              \
        case 9: FN<3, kFixed>(__VA_ARGS__); break;                \
        case 10: FN<3, kPlain>(__VA_ARGS__); break;               \
        case 11: FN<3, kOutlier>(__VA_ARGS__); break;             \
        default: return hipErrorInvalidValue;                     \
    }

hipError_t launch_compress_f64(int dim, int mode, const double* in, uint8_t* out, const Geom& g,
                               double eb, const ScanArgs& sa, uint64_t cap, hipStream_t st) {
    const double r = double(0.5f) / eb;   // 0.5f/eb in double on the host (src/cuSZp_kernels_1D_f64.cu:33)
    CUSZP_DISPATCH(launch_c, in, out, g, r, sa, cap, st)
    return hipGetLastError();
}

hipError_t launch_decompress_f64(int dim, int mode, double* out, const uint8_t* cmp, const Geom& g,
                                 double eb, const ScanArgs& sa, uint64_t cmp_size, hipStream_t st) {
    CUSZP_DISPATCH(launch_d, out, cmp, g, eb, sa, cmp_size, st)
    return hipGetLastError();
}

}  // namespace cuszp
