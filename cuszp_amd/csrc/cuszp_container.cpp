// Self-describing container around the raw stream: a 64-byte header (layout
// in include/cuSZp.h) that records what the raw stream alone does not carry,
// i.e. the type, mode, dims, bound and length. The reference has no such
// format: its CLI writes the raw stream only (examples/cuSZp.cpp:306-317), and
// the caller must remember every argument needed to read it back.
//
// The raw stream inside is byte for byte the output of cuSZp_compress_ex, so
// the kernels are the single-array ones and nothing here touches the device
// except one 64-byte copy each way.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstring>

#include "../../include/cuSZp.h"
#include "common.hpp"

namespace cuszp {
namespace {

constexpr unsigned char kMagic[8] = {0x89, 'C', 'S', 'Z', 'P', 0x0D, 0x0A, 0x1A};

// CRC-32 with the reflected IEEE 802.3 polynomial (the checksum of zlib's
// crc32 and of PNG), bit by bit: it runs over 64 bytes per call.
uint32_t crc32_ieee(const unsigned char* p, size_t n) {
    uint32_t c = 0xffffffffu;
    for (size_t i = 0; i < n; ++i) {
        c ^= p[i];
        for (int k = 0; k < 8; ++k) c = (c >> 1) ^ (0xedb88320u & (0u - (c & 1u)));
    }
    return ~c;
}

void put_u16(unsigned char* p, uint16_t v) { p[0] = (unsigned char)v; p[1] = (unsigned char)(v >> 8); }
void put_u32(unsigned char* p, uint32_t v) {
    for (int i = 0; i < 4; ++i) p[i] = (unsigned char)(v >> (8 * i));
}
void put_u64(unsigned char* p, uint64_t v) {
    for (int i = 0; i < 8; ++i) p[i] = (unsigned char)(v >> (8 * i));
}
uint16_t get_u16(const unsigned char* p) { return (uint16_t)(p[0] | (p[1] << 8)); }
uint32_t get_u32(const unsigned char* p) {
    uint32_t v = 0;
    for (int i = 3; i >= 0; --i) v = (v << 8) | p[i];
    return v;
}
uint64_t get_u64(const unsigned char* p) {
    uint64_t v = 0;
    for (int i = 7; i >= 0; --i) v = (v << 8) | p[i];
    return v;
}
double get_f64(const unsigned char* p) {
    const uint64_t b = get_u64(p);
    double d;
    std::memcpy(&d, &b, 8);
    return d;
}
void put_f64(unsigned char* p, double d) {
    uint64_t b;
    std::memcpy(&b, &d, 8);
    put_u64(p, b);
}

// The checks a header must pass in both directions: the compressor's own
// argument rules (enums, geometry, bound) plus a stream that is at least the
// header the geometry implies and at most the largest stream it allows.
int check_info(const cuszp_container_info_t& in) {
    if (bad_enums(in.dim, in.type, in.mode)) return CUSZP_ERROR_INVALID_ARGUMENT;
    if (bad_eb(in.errorBound, in.type)) return CUSZP_ERROR_INVALID_ARGUMENT;
    if (in.relative ? !(in.relBound > 0.0 && std::isfinite(in.relBound)) : in.relBound != 0.0)
        return CUSZP_ERROR_INVALID_ARGUMENT;
    const cuszp_dims_t d = in.dim == CUSZP_DIM_1D ? cuszp_dims_t{0, 0, 0} : in.dims;
    if (in.dim == CUSZP_DIM_1D && (in.dims.x || in.dims.y || in.dims.z)) return CUSZP_ERROR_INVALID_DIMS;
    Geom g;
    const int rc = make_geom(in.dim, (size_t)in.nbEle, d, g);
    if (rc) return rc;
    if (in.nbEle == 0) return in.streamBytes == 0 ? CUSZP_SUCCESS : CUSZP_ERROR_CORRUPT_STREAM;
    if (in.streamBytes < g.H || in.streamBytes > max_stream_bytes(in.dim, g)) return CUSZP_ERROR_CORRUPT_STREAM;
    return CUSZP_SUCCESS;
}

}  // namespace
}  // namespace cuszp

using namespace cuszp;

extern "C" {

int cuSZp_container_write_header(const cuszp_container_info_t* info, unsigned char* header) {
    if (!info || !header) return CUSZP_ERROR_INVALID_ARGUMENT;
    const int rc = check_info(*info);
    if (rc) return rc;
    unsigned char h[CUSZP_CONTAINER_HEADER_BYTES] = {};
    std::memcpy(h, kMagic, 8);
    put_u16(h + 8, CUSZP_CONTAINER_VERSION);
    put_u16(h + 10, CUSZP_CONTAINER_HEADER_BYTES);
    h[12] = (unsigned char)info->dim;
    h[13] = (unsigned char)info->type;
    h[14] = (unsigned char)info->mode;
    h[15] = info->relative ? 1 : 0;
    put_u32(h + 16, info->dims.x);
    put_u32(h + 20, info->dims.y);
    put_u32(h + 24, info->dims.z);
    put_u64(h + 32, info->nbEle);
    put_f64(h + 40, info->type == CUSZP_TYPE_FLOAT ? (double)(float)info->errorBound : info->errorBound);
    put_f64(h + 48, info->relative ? info->relBound : 0.0);
    put_u64(h + 56, info->streamBytes);
    put_u32(h + 28, crc32_ieee(h, sizeof(h)));
    std::memcpy(header, h, sizeof(h));
    return CUSZP_SUCCESS;
}

int cuSZp_container_read_header(const unsigned char* header, size_t len, cuszp_container_info_t* info) {
    if (!header || !info) return CUSZP_ERROR_INVALID_ARGUMENT;
    std::memset(info, 0, sizeof(*info));
    if (len < CUSZP_CONTAINER_HEADER_BYTES) return CUSZP_ERROR_CORRUPT_STREAM;
    if (std::memcmp(header, kMagic, 8)) return CUSZP_ERROR_CORRUPT_STREAM;
    if (get_u16(header + 8) != CUSZP_CONTAINER_VERSION || get_u16(header + 10) != CUSZP_CONTAINER_HEADER_BYTES)
        return CUSZP_ERROR_CORRUPT_STREAM;
    unsigned char h[CUSZP_CONTAINER_HEADER_BYTES];
    std::memcpy(h, header, sizeof(h));
    const uint32_t crc = get_u32(h + 28);
    put_u32(h + 28, 0);
    if (crc32_ieee(h, sizeof(h)) != crc) return CUSZP_ERROR_CORRUPT_STREAM;
    if (h[15] & ~1u) return CUSZP_ERROR_CORRUPT_STREAM;
    cuszp_container_info_t in;
    std::memset(&in, 0, sizeof(in));
    in.dim = (cuszp_dim_t)h[12];
    in.type = (cuszp_type_t)h[13];
    in.mode = (cuszp_mode_t)h[14];
    in.relative = h[15] & 1;
    in.dims.x = get_u32(h + 16);
    in.dims.y = get_u32(h + 20);
    in.dims.z = get_u32(h + 24);
    in.nbEle = get_u64(h + 32);
    in.errorBound = get_f64(h + 40);
    in.relBound = get_f64(h + 48);
    in.streamBytes = get_u64(h + 56);
    if (check_info(in)) return CUSZP_ERROR_CORRUPT_STREAM;
    *info = in;
    return CUSZP_SUCCESS;
}

int cuSZp_compress_container(const void* d_oriData, unsigned char* d_container, size_t capacity, size_t nbEle,
                             size_t* containerSize, double errorBound, double relBound, cuszp_dim_t dim,
                             cuszp_dims_t dims, cuszp_type_t type, cuszp_mode_t mode, hipStream_t stream) {
    if (containerSize) *containerSize = 0;
    if (!d_container || capacity < CUSZP_CONTAINER_HEADER_BYTES) return CUSZP_ERROR_OUTPUT_TOO_SMALL;
    if (!(relBound >= 0.0) || !std::isfinite(relBound)) return CUSZP_ERROR_INVALID_ARGUMENT;
    cuszp_container_info_t info;
    std::memset(&info, 0, sizeof(info));
    info.dim = dim;
    info.type = type;
    info.mode = mode;
    info.relative = relBound > 0.0;
    info.dims = dim == CUSZP_DIM_1D ? cuszp_dims_t{0, 0, 0} : dims;
    info.nbEle = nbEle;
    info.errorBound = errorBound;
    info.relBound = relBound;
    size_t stream_bytes = 0;
    int rc = cuSZp_compress_ex(d_oriData, d_container + CUSZP_CONTAINER_HEADER_BYTES,
                               capacity - CUSZP_CONTAINER_HEADER_BYTES, nbEle, &stream_bytes, errorBound, dim, dims,
                               type, mode, stream);
    if (!status_delivers(rc)) return rc;
    info.streamBytes = stream_bytes;
    unsigned char h[CUSZP_CONTAINER_HEADER_BYTES];
    const int hrc = cuSZp_container_write_header(&info, h);
    if (hrc) return hrc;
    // h lives on this stack frame: the copy completes before returning
    if (hipMemcpyAsync(d_container, h, sizeof(h), hipMemcpyHostToDevice, stream) != hipSuccess ||
        hipStreamSynchronize(stream) != hipSuccess)
        return CUSZP_ERROR_HIP;
    if (containerSize) *containerSize = CUSZP_CONTAINER_HEADER_BYTES + stream_bytes;
    return rc;
}

int cuSZp_decompress_container(void* d_decData, size_t decCapacity, const unsigned char* d_container,
                               size_t containerSize, cuszp_container_info_t* info, hipStream_t stream) {
    if (info) std::memset(info, 0, sizeof(*info));
    if (!d_container) return CUSZP_ERROR_INVALID_ARGUMENT;
    if (containerSize < CUSZP_CONTAINER_HEADER_BYTES) return CUSZP_ERROR_CORRUPT_STREAM;
    unsigned char h[CUSZP_CONTAINER_HEADER_BYTES];
    if (hipMemcpyAsync(h, d_container, sizeof(h), hipMemcpyDeviceToHost, stream) != hipSuccess ||
        hipStreamSynchronize(stream) != hipSuccess)
        return CUSZP_ERROR_HIP;
    cuszp_container_info_t in;
    int rc = cuSZp_container_read_header(h, sizeof(h), &in);
    if (rc) return rc;
    if (info) *info = in;
    if (containerSize - CUSZP_CONTAINER_HEADER_BYTES < in.streamBytes) return CUSZP_ERROR_CORRUPT_STREAM;
    if (in.nbEle > decCapacity) return CUSZP_ERROR_OUTPUT_TOO_SMALL;
    return cuSZp_decompress_ex(d_decData, d_container + CUSZP_CONTAINER_HEADER_BYTES, (size_t)in.nbEle,
                               (size_t)in.streamBytes, in.errorBound, in.dim, in.dims, in.type, in.mode, stream);
}

}  // extern "C"
