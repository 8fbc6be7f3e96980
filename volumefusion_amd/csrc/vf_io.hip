// vf_io.hip -- TSDF volumes on disk: checkpoint and resume of resident volumes.
//
// Two layouts:
//   VF_TSDF_FORMAT_NATIVE  a 64-byte header (magic "VFTSDF", version 1,
//       header size, dims, size, centre, delta, max weight, flags) followed by
//       tsdf[N] and weight[N] as little-endian float32, z-major as in memory.
//   VF_TSDF_FORMAT_LEGACY  the reference's TSDFVolume::save layout
//       (src/TSDFVolume.cpp:339-358): int32 dims[3], a 24-byte size field,
//       float delta, tsdf[N], weight[N], colour R/G/B (uint8 each) and colour
//       weight (float) [N].  The reference writes its Vec3f size with
//       sizeof(double) * 3 (:344, :324 on load), i.e. 12 bytes of size and the
//       12 bytes of the voxel size that follows it in the object; here the
//       field holds the float size and size / dim, and loading reads only its
//       first 12 bytes.  Colours are written as zeros (the device volumes
//       carry none) and skipped on load.
// Device volumes are streamed through a bounded host buffer.
#include <cstdio>
#include <cstring>
#include <vector>

#include "vf_host.hpp"

namespace vf {
namespace {

constexpr char kMagic[8] = {'V', 'F', 'T', 'S', 'D', 'F', 0, 1};
constexpr uint32_t kHeaderBytes = 64;
constexpr uint64_t kChunk = 16ull << 20;  // floats per host staging chunk (64 MB)

struct Header {
    char magic[8];
    uint32_t header_bytes;
    int32_t dim[3];
    float size[3];
    float center[3];
    float delta;
    float max_weight;
    uint32_t flags;
    uint32_t reserved;
};
static_assert(sizeof(Header) == kHeaderBytes, "header is 64 bytes");

struct File {
    FILE* f = nullptr;
    ~File() {
        if (f) fclose(f);
    }
};

int io_fail(const char* what, const char* path) { return fail(VF_ERR_INVALID_ARGUMENT, "%s: %s", what, path); }

uint64_t voxels_of(const vf_int3& d) { return (uint64_t)d.x * (uint64_t)d.y * (uint64_t)d.z; }

// Copies n floats between the file and a host or device array.
int write_floats(FILE* f, const float* src, uint64_t n, bool device, std::vector<float>& buf) {
    for (uint64_t o = 0; o < n; o += kChunk) {
        const uint64_t m = n - o < kChunk ? n - o : kChunk;
        const float* p = src + o;
        if (device) {
            buf.resize(m);
            VF_HIP(hipMemcpy(buf.data(), src + o, m * 4, hipMemcpyDeviceToHost));
            p = buf.data();
        }
        if (fwrite(p, 4, m, f) != m) return fail(VF_ERR_INVALID_ARGUMENT, "short write");
    }
    return VF_OK;
}

int read_floats(FILE* f, float* dst, uint64_t n, bool device, std::vector<float>& buf) {
    for (uint64_t o = 0; o < n; o += kChunk) {
        const uint64_t m = n - o < kChunk ? n - o : kChunk;
        float* p = device ? nullptr : dst + o;
        if (device) {
            buf.resize(m);
            p = buf.data();
        }
        if (fread(p, 4, m, f) != m) return fail(VF_ERR_INVALID_ARGUMENT, "file ends inside the volume data");
        if (device) VF_HIP(hipMemcpy(dst + o, buf.data(), m * 4, hipMemcpyHostToDevice));
    }
    return VF_OK;
}

int write_zeros(FILE* f, uint64_t bytes) {
    static const char z[4096] = {0};
    while (bytes) {
        const uint64_t m = bytes < sizeof(z) ? bytes : sizeof(z);
        if (fwrite(z, 1, m, f) != m) return fail(VF_ERR_INVALID_ARGUMENT, "short write");
        bytes -= m;
    }
    return VF_OK;
}

int save(const char* path, int format, const vf_tsdf_info* info, const float* tsdf, const float* weights,
         bool device) {
    if (!path || !info || !tsdf || !weights) return fail(VF_ERR_INVALID_ARGUMENT, "vf_tsdf_save: NULL argument");
    if (info->dim.x < 0 || info->dim.y < 0 || info->dim.z < 0)
        return fail(VF_ERR_INVALID_ARGUMENT, "vf_tsdf_save: negative dims");
    const uint64_t n = voxels_of(info->dim);
    File fh;
    fh.f = fopen(path, "wb");
    if (!fh.f) return io_fail("cannot open for writing", path);
    std::vector<float> buf;
    int st;
    if (format == VF_TSDF_FORMAT_NATIVE) {
        Header h;
        memset(&h, 0, sizeof(h));
        memcpy(h.magic, kMagic, 8);
        h.header_bytes = kHeaderBytes;
        h.dim[0] = info->dim.x;
        h.dim[1] = info->dim.y;
        h.dim[2] = info->dim.z;
        h.size[0] = info->size.x;
        h.size[1] = info->size.y;
        h.size[2] = info->size.z;
        h.center[0] = info->center.x;
        h.center[1] = info->center.y;
        h.center[2] = info->center.z;
        h.delta = info->delta;
        h.max_weight = info->max_weight;
        h.flags = 1;  // weights present
        if (fwrite(&h, sizeof(h), 1, fh.f) != 1) return io_fail("short write", path);
        st = write_floats(fh.f, tsdf, n, device, buf);
        if (!st) st = write_floats(fh.f, weights, n, device, buf);
    } else if (format == VF_TSDF_FORMAT_LEGACY) {
        const int32_t dim[3] = {info->dim.x, info->dim.y, info->dim.z};
        const float size[3] = {info->size.x, info->size.y, info->size.z};
        // The 24-byte size field is sizeof(double) * 3 bytes of the Vec3f
        // m_size and of the Vec3f m_voxelSize after it (include/af/TSDFVolume.h:93-94):
        // TSDFVolume::load reads both from it and does not recompute the
        // voxel size, so the second half carries size / dim as the
        // constructor and setSize compute it (src/TSDFVolume.cpp:11, 391).
        const float vsz[3] = {size[0] / (float)dim[0], size[1] / (float)dim[1], size[2] / (float)dim[2]};
        if (fwrite(dim, 4, 3, fh.f) != 3 || fwrite(size, 4, 3, fh.f) != 3 || fwrite(vsz, 4, 3, fh.f) != 3)
            return io_fail("short write", path);
        st = VF_OK;
        if (fwrite(&info->delta, 4, 1, fh.f) != 1) return io_fail("short write", path);
        if (!st) st = write_floats(fh.f, tsdf, n, device, buf);
        if (!st) st = write_floats(fh.f, weights, n, device, buf);
        if (!st) st = write_zeros(fh.f, 3 * n + 4 * n);  // colour R, G, B and colour weights
    } else {
        return fail(VF_ERR_INVALID_ARGUMENT, "vf_tsdf_save: unknown format %d", format);
    }
    if (st) return st;
    if (fflush(fh.f) != 0) return io_fail("write failed", path);
    return VF_OK;
}

// Reads the header of either layout; leaves the file at the volume data.
int read_header(FILE* f, const char* path, vf_tsdf_info* info, int* format) {
    char magic[8];
    if (fread(magic, 1, 8, f) != 8) return io_fail("file too short", path);
    memset(info, 0, sizeof(*info));
    if (memcmp(magic, kMagic, 6) == 0) {
        if (magic[6] != 0 || magic[7] != 1) return io_fail("unsupported format version", path);
        Header h;
        memcpy(h.magic, magic, 8);
        if (fread((char*)&h + 8, sizeof(h) - 8, 1, f) != 1) return io_fail("file too short", path);
        if (h.header_bytes < kHeaderBytes) return io_fail("bad header size", path);
        if (h.header_bytes > kHeaderBytes && fseek(f, (long)(h.header_bytes - kHeaderBytes), SEEK_CUR) != 0)
            return io_fail("file too short", path);
        info->dim = vf_int3{h.dim[0], h.dim[1], h.dim[2]};
        info->size = vf_float3{h.size[0], h.size[1], h.size[2]};
        info->center = vf_float3{h.center[0], h.center[1], h.center[2]};
        info->delta = h.delta;
        info->max_weight = h.max_weight;
        *format = VF_TSDF_FORMAT_NATIVE;
    } else {
        // legacy: dims (12 B; the first 8 already read), size field (24 B), delta
        int32_t dim[3];
        memcpy(dim, magic, 8);
        float size[3];
        char pad[12];
        if (fread(&dim[2], 4, 1, f) != 1 || fread(size, 4, 3, f) != 3 || fread(pad, 1, 12, f) != 12 ||
            fread(&info->delta, 4, 1, f) != 1)
            return io_fail("file too short", path);
        info->dim = vf_int3{dim[0], dim[1], dim[2]};
        info->size = vf_float3{size[0], size[1], size[2]};
        info->max_weight = 0.0f;
        *format = VF_TSDF_FORMAT_LEGACY;
    }
    if (info->dim.x < 0 || info->dim.y < 0 || info->dim.z < 0 || voxels_of(info->dim) > (1ull << 40))
        return io_fail("implausible dims in", path);
    return VF_OK;
}

int load(const char* path, vf_tsdf_info* info, float* tsdf, float* weights, uint64_t capacity, bool device) {
    if (!path || !info) return fail(VF_ERR_INVALID_ARGUMENT, "vf_tsdf_load: NULL argument");
    File fh;
    fh.f = fopen(path, "rb");
    if (!fh.f) return io_fail("cannot open", path);
    int format = 0;
    int st = read_header(fh.f, path, info, &format);
    if (st) return st;
    info->format = format;
    const uint64_t n = voxels_of(info->dim);
    if (!tsdf && !weights) return VF_OK;  // header only
    if (!tsdf || !weights) return fail(VF_ERR_INVALID_ARGUMENT, "vf_tsdf_load: tsdf and weights both or neither");
    if (n > capacity)
        return fail(VF_ERR_CAPACITY, "%s holds %llu voxels; buffers hold %llu", path, (unsigned long long)n,
                    (unsigned long long)capacity);
    std::vector<float> buf;
    st = read_floats(fh.f, tsdf, n, device, buf);
    if (!st) st = read_floats(fh.f, weights, n, device, buf);
    return st;
}

}  // namespace
}  // namespace vf

using namespace vf;

extern "C" {

int vf_tsdf_save(const char* path, int format, const vf_tsdf_info* info, const float* tsdf, const float* weights) {
    return save(path, format, info, tsdf, weights, true);
}

int vf_tsdf_load(const char* path, vf_tsdf_info* info, float* tsdf, float* weights, uint64_t capacity) {
    return load(path, info, tsdf, weights, capacity, true);
}

int vf_tsdf_save_host(const char* path, int format, const vf_tsdf_info* info, const float* tsdf,
                      const float* weights) {
    return save(path, format, info, tsdf, weights, false);
}

int vf_tsdf_load_host(const char* path, vf_tsdf_info* info, float* tsdf, float* weights, uint64_t capacity) {
    return load(path, info, tsdf, weights, capacity, false);
}

}  // extern "C"
