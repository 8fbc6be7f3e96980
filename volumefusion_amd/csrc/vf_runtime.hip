// vf_runtime.hip -- status reporting, shared host helpers and the device
// plumbing part of the C ABI (memory, streams, events).
#include <cstdio>
#include <cstring>

#include "vf_host.hpp"

namespace vf {

static thread_local char g_last_error[1024] = "";

int fail(int status, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_last_error, sizeof(g_last_error), fmt, ap);
    va_end(ap);
    return status;
}

void mat3_inverse_eigen(const float* m, float* out) {
    auto at = [m](int i, int j) { return m[j * 3 + i]; };
    auto cof = [&](int i, int j) {
        const int i1 = (i + 1) % 3, i2 = (i + 2) % 3, j1 = (j + 1) % 3, j2 = (j + 2) % 3;
        const float a = at(i1, j1) * at(i2, j2);
        const float b = at(i1, j2) * at(i2, j1);
        return a - b;
    };
    const float c0 = cof(0, 0), c1 = cof(1, 0), c2 = cof(2, 0);
    const float det = c0 * at(0, 0) + (c1 * at(1, 0) + c2 * at(2, 0));
    const float invdet = 1.0f / det;
    auto set = [out](int i, int j, float v) { out[j * 3 + i] = v; };
    set(0, 0, c0 * invdet);
    set(0, 1, c1 * invdet);
    set(0, 2, c2 * invdet);
    set(1, 0, cof(0, 1) * invdet);
    set(1, 1, cof(1, 1) * invdet);
    set(1, 2, cof(2, 1) * invdet);
    set(2, 0, cof(0, 2) * invdet);
    set(2, 1, cof(1, 2) * invdet);
    set(2, 2, cof(2, 2) * invdet);
}

int check_volume(vf_int3 dim, vf_float3 vs, uint64_t* n_out) {
    if (dim.x < 1 || dim.y < 1 || dim.z < 1)
        return fail(VF_ERR_INVALID_ARGUMENT, "volume dims (%d, %d, %d) must be positive", dim.x, dim.y, dim.z);
    if (dim.y > 65535 || dim.z > 65535)
        return fail(VF_ERR_INVALID_ARGUMENT, "volume dims (%d, %d, %d): y and z must be <= 65535", dim.x, dim.y,
                    dim.z);
    const uint64_t n = (uint64_t)dim.x * (uint64_t)dim.y * (uint64_t)dim.z;
    // The reference indexes voxels with int (src/TSDFIntegration.cu:218).
    if (n > 0x7fffffffull)
        return fail(VF_ERR_INVALID_ARGUMENT, "volume of %llu voxels exceeds 2^31-1", (unsigned long long)n);
    if (!(vs.x > 0.0f) || !(vs.y > 0.0f) || !(vs.z > 0.0f))
        return fail(VF_ERR_INVALID_ARGUMENT, "voxel size must be positive");
    *n_out = n;
    return VF_OK;
}

}  // namespace vf

extern "C" {

const char* vf_status_string(int status) {
    switch (status) {
        case VF_OK: return "ok";
        case VF_ERR_INVALID_ARGUMENT: return "invalid argument";
        case VF_ERR_UNSUPPORTED_K: return "unsupported knn slot count";
        case VF_ERR_HIP: return "HIP runtime error";
        case VF_ERR_OUT_OF_MEMORY: return "out of device memory";
        case VF_ERR_WORKSPACE: return "workspace too small";
        case VF_ERR_CAPACITY: return "capacity exceeded";
        case VF_ERR_SINGULAR: return "singular linear system";
        default: return "unknown status";
    }
}

const char* vf_last_error(void) { return vf::g_last_error; }

const char* vf_version(void) { return "volumefusion_amd 0.1 gfx950"; }

int vf_device_count(int* count) {
    VF_HIP(hipGetDeviceCount(count));
    return VF_OK;
}

int vf_set_device(int device) {
    VF_HIP(hipSetDevice(device));
    return VF_OK;
}

int vf_malloc(void** ptr, uint64_t bytes) {
    *ptr = nullptr;
    if (bytes == 0) return VF_OK;
    hipError_t e = hipMalloc(ptr, bytes);
    if (e == hipErrorOutOfMemory)
        return vf::fail(VF_ERR_OUT_OF_MEMORY, "hipMalloc(%llu) failed: out of memory", (unsigned long long)bytes);
    if (e != hipSuccess) return vf::fail(VF_ERR_HIP, "hipMalloc(%llu): %s", (unsigned long long)bytes, hipGetErrorString(e));
    return VF_OK;
}

int vf_free(void* ptr) {
    if (ptr) VF_HIP(hipFree(ptr));
    return VF_OK;
}

int vf_memcpy_h2d(void* dst, const void* src, uint64_t bytes, vf_stream stream) {
    if (bytes) VF_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, (hipStream_t)stream));
    return VF_OK;
}

int vf_memcpy_d2h(void* dst, const void* src, uint64_t bytes, vf_stream stream) {
    if (bytes) {
        VF_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, (hipStream_t)stream));
        VF_HIP(hipStreamSynchronize((hipStream_t)stream));
    }
    return VF_OK;
}

int vf_memcpy_d2d(void* dst, const void* src, uint64_t bytes, vf_stream stream) {
    if (bytes) VF_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, (hipStream_t)stream));
    return VF_OK;
}

int vf_memset(void* dst, int value, uint64_t bytes, vf_stream stream) {
    if (bytes) VF_HIP(hipMemsetAsync(dst, value, bytes, (hipStream_t)stream));
    return VF_OK;
}

int vf_stream_create(vf_stream* stream) {
    hipStream_t s;
    VF_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    *stream = (vf_stream)s;
    return VF_OK;
}

int vf_stream_destroy(vf_stream stream) {
    if (stream) VF_HIP(hipStreamDestroy((hipStream_t)stream));
    return VF_OK;
}

int vf_stream_synchronize(vf_stream stream) {
    VF_HIP(hipStreamSynchronize((hipStream_t)stream));
    return VF_OK;
}

int vf_device_synchronize(void) {
    VF_HIP(hipDeviceSynchronize());
    return VF_OK;
}

int vf_host_alloc(void** ptr, uint64_t bytes) {
    *ptr = nullptr;
    if (bytes) VF_HIP(hipHostMalloc(ptr, bytes, hipHostMallocDefault));
    return VF_OK;
}

int vf_host_free(void* ptr) {
    if (ptr) VF_HIP(hipHostFree(ptr));
    return VF_OK;
}

int vf_event_create(void** event) {
    hipEvent_t e;
    VF_HIP(hipEventCreate(&e));
    *event = (void*)e;
    return VF_OK;
}

int vf_event_destroy(void* event) {
    if (event) VF_HIP(hipEventDestroy((hipEvent_t)event));
    return VF_OK;
}

int vf_event_record(void* event, vf_stream stream) {
    VF_HIP(hipEventRecord((hipEvent_t)event, (hipStream_t)stream));
    return VF_OK;
}

int vf_event_elapsed_ms(float* ms, void* start, void* stop) {
    VF_HIP(hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)stop));
    return VF_OK;
}

int vf_event_synchronize(void* event) {
    VF_HIP(hipEventSynchronize((hipEvent_t)event));
    return VF_OK;
}

}  // extern "C"
