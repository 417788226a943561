// Command-line front end with the options and output lines of the reference's
// CLI (examples/cuSZp.cpp:9-330): reads a raw little-endian f32/f64 file,
// compresses and decompresses it on the GPU, prints the speeds, the ratio and
// the error check, and optionally writes the stream (-x, a raw header-less
// cuSZp stream) and the reconstruction (-o, raw values).
//
// Differences: the relative bound is computed on the device with the same
// float arithmetic (cuSZp_rel_error_bound); the stream buffer is sized with
// cuSZp_max_compressed_size instead of nbEle * sizeof(value), which can be too
// small; the timings are host wall time around each blocking call after 10
// warm-up compressions, in the reference's units (MiB per ms, printed as GB/s).
//
// Two options the reference does not have: -X writes the stream inside the
// self-describing container of include/cuSZp.h (type, mode, dims, bound and
// length in a 64-byte header), and -r reads such a container back and
// decompresses it to -o with no other arguments (with -i it also runs the
// error check against the original).
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../include/cuSZp.h"

namespace {

void usage() {
    std::printf(
        "cuSZp (MI355X build): a GPU error-bounded lossy compressor for floating-point data\n\n"
        "Usage:\n"
        "  cuszp -i [oriFilePath] -t [dataType] -m [encodingMode] -d [dim] [dim_z] [dim_y] [dim_x] "
        "-eb [errorBoundMode] [errorBound] -x [cmpFilePath] -o [decFilePath]\n\n"
        "Options:\n"
        "  -i  [oriFilePath]    raw input file (little-endian values)\n"
        "  -t  [dataType]       f32 | f64\n"
        "  -m  [encodingMode]   fixed | plain | outlier\n"
        "  -d  [dim]            1 | 2 [dim_z] [dim_y] [dim_x] | 3 [dim_z] [dim_y] [dim_x] (dim_x fastest)\n"
        "  -eb [errorBoundMode] [errorBound]   abs | rel, e.g. -eb rel 1e-4\n"
        "  -x  [cmpFilePath]    write the compressed stream (optional)\n"
        "  -o  [decFilePath]    write the decompressed values (optional)\n"
        "  -X  [cszpFilePath]   write the stream in a self-describing container (optional)\n\n"
        "  cuszp -r [cszpFilePath] -o [decFilePath] [-i oriFilePath]\n"
        "  -r  [cszpFilePath]   decompress a container; type, mode, dims and bound come from its header\n");
}

#define CLI_HIP(call)                                                            \
    do {                                                                         \
        hipError_t e_ = (call);                                                  \
        if (e_ != hipSuccess) {                                                  \
            std::fprintf(stderr, "Error: %s: %s\n", #call, hipGetErrorString(e_)); \
            return 1;                                                            \
        }                                                                        \
    } while (0)

bool read_file(const char* path, std::vector<unsigned char>& out) {
    FILE* f = std::fopen(path, "rb");
    if (!f) return false;
    std::fseek(f, 0, SEEK_END);
    const long len = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    out.resize(len > 0 ? (size_t)len : 0);
    const bool ok = len >= 0 && std::fread(out.data(), 1, out.size(), f) == out.size();
    std::fclose(f);
    return ok;
}

bool write_file(const char* path, const void* p, size_t n) {
    FILE* f = std::fopen(path, "wb");
    if (!f) return false;
    const bool ok = std::fwrite(p, 1, n, f) == n;
    std::fclose(f);
    return ok;
}

double now_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// Decompresses the container at `path` to decPath (if given); with `ori`, runs
// the reference's 1.1 * eb error check against the original values.
int read_container(const char* path, const char* decPath, const char* ori) {
    std::vector<unsigned char> box;
    if (!read_file(path, box)) {
        std::printf("Error: cannot read %s\n", path);
        return 1;
    }
    cuszp_container_info_t info;
    int rc = cuSZp_container_read_header(box.data(), box.size(), &info);
    if (rc) {
        std::printf("Error: %s is not a cuSZp container: %s\n", path, cuSZp_status_string(rc));
        return 1;
    }
    if (box.size() < CUSZP_CONTAINER_HEADER_BYTES + info.streamBytes) {
        std::printf("Error: %s is truncated (%zu of %llu bytes)\n", path, box.size(),
                    (unsigned long long)(CUSZP_CONTAINER_HEADER_BYTES + info.streamBytes));
        return 1;
    }
    const size_t s = info.type == CUSZP_TYPE_FLOAT ? 4 : 8;
    const size_t n = (size_t)info.nbEle;
    static const char* const kModes[] = {"fixed", "plain", "outlier"};
    std::printf("container: %s %s %dD", info.type == CUSZP_TYPE_FLOAT ? "f32" : "f64", kModes[info.mode],
                (int)info.dim);
    if (info.dim != CUSZP_DIM_1D) std::printf(" [z y x] = %u %u %u", info.dims.z, info.dims.y, info.dims.x);
    std::printf(", %zu values, error bound %.9g", n, info.errorBound);
    if (info.relative) std::printf(" (rel %.9g)", info.relBound);
    std::printf(", stream %llu bytes\n", (unsigned long long)info.streamBytes);
    void* d_dec = nullptr;
    unsigned char* d_box = nullptr;
    CLI_HIP(hipMalloc(&d_dec, n * s > 0 ? n * s : 1));
    CLI_HIP(hipMalloc((void**)&d_box, box.size()));
    CLI_HIP(hipMemcpy(d_box, box.data(), box.size(), hipMemcpyHostToDevice));
    hipStream_t st;
    CLI_HIP(hipStreamCreate(&st));
    const double t0 = now_ms();
    rc = cuSZp_decompress_container(d_dec, n, d_box, box.size(), nullptr, st);
    const double decTime = now_ms() - t0;
    if (rc) {
        std::printf("Error: decompression failed: %s\n", cuSZp_status_string(rc));
        return 1;
    }
    std::vector<unsigned char> dec(n * s);
    CLI_HIP(hipMemcpy(dec.data(), d_dec, n * s, hipMemcpyDeviceToHost));
    std::printf("cuSZp decompression end-to-end speed: %f GB/s\n", n * s / 1024.0 / 1024.0 / decTime);
    if (ori) {
        std::vector<unsigned char> raw;
        if (!read_file(ori, raw) || raw.size() != n * s) {
            std::printf("Error: %s does not hold the container's %zu values\n", ori, n);
            return 1;
        }
        size_t not_bound = 0;
        for (size_t i = 0; i < n; ++i) {
            double e;
            if (info.type == CUSZP_TYPE_FLOAT)
                e = std::fabs(((const float*)raw.data())[i] - ((const float*)dec.data())[i]);
            else
                e = std::fabs(((const double*)raw.data())[i] - ((const double*)dec.data())[i]);
            if (e > info.errorBound * 1.1) ++not_bound;
        }
        if (!not_bound) std::printf("Pass error check!\n");
        else std::printf("Fail error check! Exceeding data count: %zu\n", not_bound);
    }
    if (decPath && !write_file(decPath, dec.data(), n * s)) std::printf("Error: cannot write %s\n", decPath);
    hipFree(d_dec);
    hipFree(d_box);
    hipStreamDestroy(st);
    return 0;
}

}  // namespace

int main(int argc, char* argv[]) {
    if (argc == 1 || (argc == 2 && (!std::strcmp(argv[1], "-h") || !std::strcmp(argv[1], "--help")))) {
        usage();
        return 1;
    }
    const char *ori = nullptr, *cmpPath = nullptr, *decPath = nullptr, *boxPath = nullptr, *readPath = nullptr;
    int dim = 1, type = -1, mode = -1;
    cuszp_dims_t dims = {0, 0, 0};
    bool rel = false, have_eb = false;
    float eb = 0.0f;
    for (int i = 1; i < argc; ++i) {
        const char* a = argv[i];
        if (!std::strcmp(a, "-i") && i + 1 < argc) {
            ori = argv[++i];
        } else if (!std::strcmp(a, "-t") && i + 1 < argc) {
            const char* t = argv[++i];
            type = !std::strcmp(t, "f32") ? CUSZP_TYPE_FLOAT : !std::strcmp(t, "f64") ? CUSZP_TYPE_DOUBLE : -2;
            if (type < 0) {
                std::printf("Error: Unsupported data type. Use 'f32' or 'f64'.\n");
                return 1;
            }
        } else if (!std::strcmp(a, "-m") && i + 1 < argc) {
            const char* m = argv[++i];
            mode = !std::strcmp(m, "fixed") ? CUSZP_MODE_FIXED
                 : !std::strcmp(m, "plain") ? CUSZP_MODE_PLAIN
                 : !std::strcmp(m, "outlier") ? CUSZP_MODE_OUTLIER : -2;
            if (mode < 0) {
                std::printf("Error: Unsupported encoding mode. Use 'fixed', 'plain', or 'outlier'.\n");
                return 1;
            }
        } else if (!std::strcmp(a, "-d") && i + 1 < argc) {
            dim = std::atoi(argv[++i]);
            if (dim == 2 || dim == 3) {
                if (i + 3 >= argc) {
                    std::printf("Error: Missing dimension sizes for 2D/3D processing.\n");
                    return 1;
                }
                dims.z = (unsigned)std::atoi(argv[i + 1]);
                dims.y = (unsigned)std::atoi(argv[i + 2]);
                dims.x = (unsigned)std::atoi(argv[i + 3]);
                i += 3;
            } else if (dim != 1) {
                std::printf("Error: Unsupported processing dimension. Use 1, 2, or 3.\n");
                return 1;
            }
        } else if (!std::strcmp(a, "-eb") && i + 2 < argc) {
            const char* m = argv[++i];
            if (std::strcmp(m, "rel") && std::strcmp(m, "abs")) {
                std::printf("Error: Unsupported error bound mode. Use 'rel' or 'abs'.\n");
                return 1;
            }
            rel = !std::strcmp(m, "rel");
            eb = (float)std::atof(argv[++i]);
            have_eb = true;
        } else if (!std::strcmp(a, "-x") && i + 1 < argc) {
            cmpPath = argv[++i];
        } else if (!std::strcmp(a, "-o") && i + 1 < argc) {
            decPath = argv[++i];
        } else if (!std::strcmp(a, "-X") && i + 1 < argc) {
            boxPath = argv[++i];
        } else if (!std::strcmp(a, "-r") && i + 1 < argc) {
            readPath = argv[++i];
        } else {
            std::printf("Error: Unrecognized option '%s'.\n", a);
            return 1;
        }
    }
    if (readPath) return read_container(readPath, decPath, ori);
    if (!ori || !have_eb || eb == 0.0f || type < 0 || mode < 0) {
        std::printf("Error: Missing required arguments.\n");
        return 1;
    }
    std::vector<unsigned char> raw;
    if (!read_file(ori, raw)) {
        std::printf("Error: cannot read %s\n", ori);
        return 1;
    }
    const size_t s = type == CUSZP_TYPE_FLOAT ? 4 : 8;
    const size_t n = raw.size() / s;
    if (dim != 1 && n != (size_t)dims.x * dims.y * dims.z) {
        std::fprintf(stderr, "Error: The number of elements in the original data does not match the dimensions\n");
        return 1;
    }
    if (n == 0) {
        std::printf("Error: empty input\n");
        return 1;
    }
    const size_t cap = cuSZp_max_compressed_size(n, (cuszp_dim_t)dim, dims, (cuszp_type_t)type, (cuszp_mode_t)mode);
    if (!cap) {
        std::printf("Error: invalid dimensions for this processing manner\n");
        return 1;
    }
    void *d_ori = nullptr, *d_dec = nullptr;
    unsigned char* d_cmp = nullptr;
    CLI_HIP(hipMalloc(&d_ori, n * s));
    CLI_HIP(hipMalloc(&d_dec, n * s));
    CLI_HIP(hipMalloc((void**)&d_cmp, cap));
    CLI_HIP(hipMemcpy(d_ori, raw.data(), n * s, hipMemcpyHostToDevice));
    hipStream_t st;
    CLI_HIP(hipStreamCreate(&st));
    const float rel_eb = eb;
    if (rel) {
        float abs_eb = 0.0f;
        const int rc = cuSZp_rel_error_bound(d_ori, n, (cuszp_type_t)type, eb, &abs_eb, st);
        if (rc) {
            std::printf("Error: %s\n", cuSZp_status_string(rc));
            return 1;
        }
        eb = abs_eb;
    }
    size_t cmpSize = 0;
    int rc = 0;
    for (int w = 0; w < 10 && !rc; ++w)
        rc = cuSZp_compress_ex(d_ori, d_cmp, cap, n, &cmpSize, eb, (cuszp_dim_t)dim, dims, (cuszp_type_t)type,
                               (cuszp_mode_t)mode, st);
    double t0 = now_ms();
    if (!rc)
        rc = cuSZp_compress_ex(d_ori, d_cmp, cap, n, &cmpSize, eb, (cuszp_dim_t)dim, dims, (cuszp_type_t)type,
                               (cuszp_mode_t)mode, st);
    const double cmpTime = now_ms() - t0;
    if (rc) {
        std::printf("Error: compression failed: %s\n", cuSZp_status_string(rc));
        return 1;
    }
    t0 = now_ms();
    rc = cuSZp_decompress_ex(d_dec, d_cmp, n, cmpSize, eb, (cuszp_dim_t)dim, dims, (cuszp_type_t)type,
                             (cuszp_mode_t)mode, st);
    const double decTime = now_ms() - t0;
    if (rc) {
        std::printf("Error: decompression failed: %s\n", cuSZp_status_string(rc));
        return 1;
    }
    std::vector<unsigned char> dec(n * s), cmp(cmpSize);
    CLI_HIP(hipMemcpy(dec.data(), d_dec, n * s, hipMemcpyDeviceToHost));
    CLI_HIP(hipMemcpy(cmp.data(), d_cmp, cmpSize, hipMemcpyDeviceToHost));
    const double mib = n * s / 1024.0 / 1024.0;
    std::printf("cuSZp finished!\n");
    std::printf("cuSZp compression   end-to-end speed: %f GB/s\n", mib / cmpTime);
    std::printf("cuSZp decompression end-to-end speed: %f GB/s\n", mib / decTime);
    std::printf("cuSZp compression ratio: %f\n\n", mib / (cmpSize / 1024.0 / 1024.0));
    std::printf("error bound: %.9g\n", (double)eb);
    size_t not_bound = 0;   // examples/cuSZp.cpp:272-295: |ori - dec| > 1.1 * eb
    for (size_t i = 0; i < n; ++i) {
        double e;
        if (type == CUSZP_TYPE_FLOAT)
            e = std::fabs(((const float*)raw.data())[i] - ((const float*)dec.data())[i]);
        else
            e = std::fabs(((const double*)raw.data())[i] - ((const double*)dec.data())[i]);
        if (e > eb * 1.1) ++not_bound;
    }
    if (!not_bound) std::printf("Pass error check!\n");
    else std::printf("Fail error check! Exceeding data count: %zu\n", not_bound);
    if (cmpPath && !write_file(cmpPath, cmp.data(), cmpSize)) std::printf("Error: cannot write %s\n", cmpPath);
    if (boxPath) {
        cuszp_container_info_t info;
        std::memset(&info, 0, sizeof(info));
        info.dim = (cuszp_dim_t)dim;
        info.type = (cuszp_type_t)type;
        info.mode = (cuszp_mode_t)mode;
        info.relative = rel ? 1 : 0;
        info.dims = dim == 1 ? cuszp_dims_t{0, 0, 0} : dims;
        info.nbEle = n;
        info.errorBound = eb;
        info.relBound = rel ? (double)rel_eb : 0.0;
        info.streamBytes = cmpSize;
        std::vector<unsigned char> box(CUSZP_CONTAINER_HEADER_BYTES + cmpSize);
        rc = cuSZp_container_write_header(&info, box.data());
        if (rc) std::printf("Error: container header: %s\n", cuSZp_status_string(rc));
        else {
            std::memcpy(box.data() + CUSZP_CONTAINER_HEADER_BYTES, cmp.data(), cmpSize);
            if (!write_file(boxPath, box.data(), box.size())) std::printf("Error: cannot write %s\n", boxPath);
        }
    }
    if (decPath && !write_file(decPath, dec.data(), n * s)) std::printf("Error: cannot write %s\n", decPath);
    hipFree(d_ori);
    hipFree(d_dec);
    hipFree(d_cmp);
    hipStreamDestroy(st);
    return 0;   // like the reference, the error check is reported, not returned
}
