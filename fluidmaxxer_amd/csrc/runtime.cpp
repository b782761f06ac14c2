// runtime.cpp -- C ABI runtime layer: errors, device memory, streams, events.
//
// Replaces the reference's CUDA helper layer (CudaHelpers.hpp:5-23,
// CudaHelpers.cu:7-59).  Errors are returned as status codes with a
// thread-local message instead of printed and thrown; the reference's
// synchronise-after-every-launch behaviour is available as a debug switch
// (fmx_set_sync_check or FMX_SYNC_CHECK=1 in the environment).
#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "fmx_internal.hpp"

namespace fmx {

namespace {
thread_local std::string g_last_error;
std::atomic<int> g_sync_check{-1};
}  // namespace

void set_error_msg(const char *msg) { g_last_error = msg ? msg : ""; }

int set_error(int code, const char *fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

bool sync_check_enabled() {
    int v = g_sync_check.load(std::memory_order_relaxed);
    if (v < 0) {
        const char *e = std::getenv("FMX_SYNC_CHECK");
        v = (e && std::atoi(e) != 0) ? 1 : 0;
        g_sync_check.store(v, std::memory_order_relaxed);
    }
    return v != 0;
}

}  // namespace fmx

using namespace fmx;

extern "C" {

int fmx_abi_version(void) { return FMX_ABI_VERSION; }

const char *fmx_last_error(void) { return g_last_error.c_str(); }

int fmx_device_count(int *count) {
    if (!count) return set_error(FMX_ERR_ARG, "fmx_device_count: null pointer");
    *count = 0;
    const hipError_t e = hipGetDeviceCount(count);
    if (e == hipErrorNoDevice) {
        *count = 0;
        return FMX_OK;
    }
    return check_hip(e, "hipGetDeviceCount");
}

int fmx_device_info(int device, char *name, int name_len, int64_t *total_mem, int *cu_count, char *arch,
                    int arch_len) {
    hipDeviceProp_t prop;
    int rc = check_hip(hipGetDeviceProperties(&prop, device), "hipGetDeviceProperties");
    if (rc) return rc;
    if (name && name_len > 0) {
        std::strncpy(name, prop.name, size_t(name_len - 1));
        name[name_len - 1] = 0;
    }
    if (arch && arch_len > 0) {
        std::strncpy(arch, prop.gcnArchName, size_t(arch_len - 1));
        arch[arch_len - 1] = 0;
    }
    if (total_mem) *total_mem = int64_t(prop.totalGlobalMem);
    if (cu_count) *cu_count = prop.multiProcessorCount;
    return FMX_OK;
}

int fmx_set_device(int device) { return check_hip(hipSetDevice(device), "hipSetDevice"); }

int fmx_alloc(void **ptr, size_t bytes) {
    if (!ptr) return set_error(FMX_ERR_ARG, "fmx_alloc: null pointer");
    *ptr = nullptr;
    if (bytes == 0) return FMX_OK;
    return check_hip(hipMalloc(ptr, bytes), "hipMalloc");
}

int fmx_free(void *ptr) {
    if (!ptr) return FMX_OK;
    return check_hip(hipFree(ptr), "hipFree");
}

static int copy(void *dst, const void *src, size_t bytes, hipMemcpyKind kind, void *stream, const char *what) {
    if (bytes == 0) return FMX_OK;
    if (!dst || !src) return set_error(FMX_ERR_ARG, "%s: null pointer", what);
    int rc = check_hip(hipMemcpyAsync(dst, src, bytes, kind, as_stream(stream)), what);
    if (rc) return rc;
    if (sync_check_enabled()) return check_hip(hipStreamSynchronize(as_stream(stream)), "hipStreamSynchronize");
    return FMX_OK;
}

int fmx_memcpy_h2d(void *dst, const void *src, size_t bytes, void *stream) {
    return copy(dst, src, bytes, hipMemcpyHostToDevice, stream, "fmx_memcpy_h2d");
}
int fmx_memcpy_d2h(void *dst, const void *src, size_t bytes, void *stream) {
    return copy(dst, src, bytes, hipMemcpyDeviceToHost, stream, "fmx_memcpy_d2h");
}
int fmx_memcpy_d2d(void *dst, const void *src, size_t bytes, void *stream) {
    return copy(dst, src, bytes, hipMemcpyDeviceToDevice, stream, "fmx_memcpy_d2d");
}

int fmx_stream_create(void **stream) {
    if (!stream) return set_error(FMX_ERR_ARG, "fmx_stream_create: null pointer");
    hipStream_t s = nullptr;
    int rc = check_hip(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreateWithFlags");
    *stream = s;
    return rc;
}
int fmx_stream_destroy(void *stream) {
    if (!stream) return FMX_OK;
    return check_hip(hipStreamDestroy(as_stream(stream)), "hipStreamDestroy");
}
int fmx_stream_sync(void *stream) { return check_hip(hipStreamSynchronize(as_stream(stream)), "hipStreamSynchronize"); }

int fmx_device_sync(void) {
    int rc = check_hip(hipGetLastError(), "launch");
    if (rc) return rc;
    return check_hip(hipDeviceSynchronize(), "hipDeviceSynchronize");
}

int fmx_set_sync_check(int on) {
    g_sync_check.store(on ? 1 : 0, std::memory_order_relaxed);
    return FMX_OK;
}

int fmx_event_create(void **event) {
    if (!event) return set_error(FMX_ERR_ARG, "fmx_event_create: null pointer");
    hipEvent_t e = nullptr;
    int rc = check_hip(hipEventCreate(&e), "hipEventCreate");
    *event = e;
    return rc;
}
int fmx_event_destroy(void *event) {
    if (!event) return FMX_OK;
    return check_hip(hipEventDestroy(static_cast<hipEvent_t>(event)), "hipEventDestroy");
}
int fmx_event_record(void *event, void *stream) {
    return check_hip(hipEventRecord(static_cast<hipEvent_t>(event), as_stream(stream)), "hipEventRecord");
}
int fmx_event_elapsed_ms(void *start, void *stop, float *ms) {
    if (!ms) return set_error(FMX_ERR_ARG, "fmx_event_elapsed_ms: null pointer");
    int rc = check_hip(hipEventSynchronize(static_cast<hipEvent_t>(stop)), "hipEventSynchronize");
    if (rc) return rc;
    return check_hip(hipEventElapsedTime(ms, static_cast<hipEvent_t>(start), static_cast<hipEvent_t>(stop)),
                     "hipEventElapsedTime");
}

}  // extern "C"
