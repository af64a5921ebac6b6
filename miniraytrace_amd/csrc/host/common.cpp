// Logging, last-error storage and the HIP device layer.
// Logging and last-error semantics follow the reference
// core/src/gpu/GpuContext.cpp:13-45: a process-wide, mutex-guarded log
// callback that falls back to stderr, and a thread-local last-error string.
#include "common.h"

#include <cstdio>
#include <mutex>

namespace mrt {

namespace {
std::mutex g_logMutex;
mrtLogFn g_logFn = nullptr;
void* g_logUser = nullptr;
thread_local std::string g_lastError;
} // namespace

void setLogCallback(mrtLogFn fn, void* user) {
    std::lock_guard<std::mutex> lock(g_logMutex);
    g_logFn = fn;
    g_logUser = user;
}

void logMessage(LogLevel level, const std::string& msg) {
    std::lock_guard<std::mutex> lock(g_logMutex);
    if (g_logFn) {
        g_logFn(static_cast<int>(level), msg.c_str(), g_logUser);
        return;
    }
    if (level >= LogLevel::Warn) std::fprintf(stderr, "[mrt] %s\n", msg.c_str());
}

const std::string& lastErrorMessage() { return g_lastError; }
void setLastErrorMessage(const std::string& msg) { g_lastError = msg; }

const char* resultToString(mrtResult r) {
    switch (r) {
        case MRT_SUCCESS: return "Success";
        case MRT_ERROR_VULKAN_INIT: return "ErrorVulkanInit";
        case MRT_ERROR_OUT_OF_MEMORY: return "ErrorOutOfMemory";
        case MRT_ERROR_INVALID_ARGUMENT: return "ErrorInvalidArgument";
        case MRT_ERROR_INVALID_HANDLE: return "ErrorInvalidHandle";
        case MRT_ERROR_SHADER_LOAD: return "ErrorShaderLoad";
        case MRT_ERROR_DEVICE_LOST: return "ErrorDeviceLost";
        default: return "ErrorUnknown";
    }
}

void throwHip(hipError_t e, const char* what, const char* file, int line) {
    std::string msg = std::string("HIP error ") + hipGetErrorName(e) + " (" +
                      hipGetErrorString(e) + ") in " + what + " at " + file + ":" +
                      std::to_string(line);
    mrtResult code = MRT_ERROR_DEVICE_LOST;
    if (e == hipErrorOutOfMemory) code = MRT_ERROR_OUT_OF_MEMORY;
    else if (e == hipErrorNoDevice || e == hipErrorInvalidDevice || e == hipErrorNotInitialized)
        code = MRT_ERROR_VULKAN_INIT;
    else if (e == hipErrorInvalidValue || e == hipErrorInvalidConfiguration)
        code = MRT_ERROR_UNKNOWN;
    throw Error(code, msg);
}

int selectDevice(uint32_t gpuIndex1) {
    int count = 0;
    const hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count <= 0)
        throw Error(MRT_ERROR_VULKAN_INIT, "no HIP device available");
    if (gpuIndex1 == 0) {
        int cur = 0;
        MRT_HIP(hipGetDevice(&cur));
        return cur;
    }
    const int idx = int(gpuIndex1 - 1);
    if (idx >= count)
        throw Error(MRT_ERROR_VULKAN_INIT,
                    "gpuIndex1 " + std::to_string(gpuIndex1) + " is out of range (" +
                        std::to_string(count) + " HIP device(s) present)");
    return idx;
}

DeviceScope::DeviceScope(int device) {
    if (hipGetDevice(&prev_) != hipSuccess) prev_ = -1;
    if (prev_ != device) {
        MRT_HIP(hipSetDevice(device));
        changed_ = true;
    }
}
DeviceScope::~DeviceScope() {
    if (changed_ && prev_ >= 0) (void)hipSetDevice(prev_);
}

void DeviceBuffer::allocate(size_t bytes) {
    release();
    if (bytes == 0) return;
    void* p = nullptr;
    MRT_HIP(hipMalloc(&p, bytes));
    ptr_ = p;
    bytes_ = bytes;
}

void DeviceBuffer::ensure(size_t bytes) {
    if (bytes <= bytes_ && ptr_) return;
    size_t cap = bytes_ + bytes_ / 2;
    if (cap < bytes) cap = bytes;
    allocate(cap);
}

void DeviceBuffer::release() {
    if (ptr_) (void)hipFree(ptr_);
    ptr_ = nullptr;
    bytes_ = 0;
}

void DeviceBuffer::upload(const void* src, size_t bytes, size_t offset, hipStream_t s) {
    if (bytes == 0) return;
    if (offset + bytes > bytes_) throw Error(MRT_ERROR_UNKNOWN, "device upload out of range");
    MRT_HIP(hipMemcpyAsync(static_cast<char*>(ptr_) + offset, src, bytes, hipMemcpyHostToDevice, s));
    MRT_HIP(hipStreamSynchronize(s));
}

void DeviceBuffer::download(void* dst, size_t bytes, size_t offset, hipStream_t s) const {
    if (bytes == 0) return;
    if (offset + bytes > bytes_) throw Error(MRT_ERROR_UNKNOWN, "device download out of range");
    MRT_HIP(hipMemcpyAsync(dst, static_cast<const char*>(ptr_) + offset, bytes, hipMemcpyDeviceToHost, s));
    MRT_HIP(hipStreamSynchronize(s));
}

void DeviceBuffer::zero(hipStream_t s) {
    if (bytes_) MRT_HIP(hipMemsetAsync(ptr_, 0, bytes_, s));
}

PinnedBuffer::~PinnedBuffer() {
    if (ptr_) (void)hipHostFree(ptr_);
}

void PinnedBuffer::ensure(size_t bytes) {
    if (bytes <= bytes_ && ptr_) return;
    if (ptr_) (void)hipHostFree(ptr_);
    ptr_ = nullptr;
    bytes_ = 0;
    void* p = nullptr;
    MRT_HIP(hipHostMalloc(&p, bytes, hipHostMallocDefault));
    ptr_ = p;
    bytes_ = bytes;
}

} // namespace mrt
