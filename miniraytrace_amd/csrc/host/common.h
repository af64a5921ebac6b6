// Host-side shared pieces: error type, logging, HIP error mapping, device
// buffers.  Errors inside the library are C++ exceptions carrying an
// mrtResult; the C entry points catch them and return the code, recording the
// message for mrtGetLastErrorMessage (reference capi/src/capi.cpp:93-316
// convention: nothing crosses the C boundary as an exception).
#pragma once

#include "mini_raytrace.h"

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <utility>

namespace mrt {

enum class LogLevel : int { Debug = 0, Info = 1, Warn = 2, Error = 3 };

void logMessage(LogLevel level, const std::string& msg);
void setLogCallback(mrtLogFn fn, void* user);
const std::string& lastErrorMessage();
void setLastErrorMessage(const std::string& msg);
const char* resultToString(mrtResult r);

struct Error : std::runtime_error {
    mrtResult code;
    Error(mrtResult c, const std::string& m) : std::runtime_error(m), code(c) {}
};

[[noreturn]] void throwHip(hipError_t e, const char* what, const char* file, int line);

#define MRT_HIP(expr)                                                          \
    do {                                                                       \
        hipError_t mrt_hip_err_ = (expr);                                      \
        if (mrt_hip_err_ != hipSuccess) ::mrt::throwHip(mrt_hip_err_, #expr,  \
                                                        __FILE__, __LINE__);   \
    } while (0)

// Runs `f`, converting exceptions into an mrtResult and a last-error message.
template <typename F>
mrtResult guarded(F&& f) {
    try {
        return f();
    } catch (const Error& e) {
        setLastErrorMessage(e.what());
        logMessage(LogLevel::Error, e.what());
        return e.code;
    } catch (const std::bad_alloc&) {
        setLastErrorMessage("host allocation failed");
        return MRT_ERROR_OUT_OF_MEMORY;
    } catch (const std::exception& e) {
        setLastErrorMessage(e.what());
        logMessage(LogLevel::Error, e.what());
        return MRT_ERROR_UNKNOWN;
    } catch (...) {
        setLastErrorMessage("unknown error");
        return MRT_ERROR_UNKNOWN;
    }
}

// Selects the HIP device for an engine: gpuIndex1 == 0 keeps the current
// device, N > 0 selects device N-1; out of range throws
// MRT_ERROR_VULKAN_INIT with "out of range" in the message (reference
// core/src/gpu/GpuContext.cpp:335-341 behaviour).
int selectDevice(uint32_t gpuIndex1);

// Scoped "make this device current" guard.
class DeviceScope {
public:
    explicit DeviceScope(int device);
    ~DeviceScope();
    DeviceScope(const DeviceScope&) = delete;
    DeviceScope& operator=(const DeviceScope&) = delete;

private:
    int prev_ = -1;
    bool changed_ = false;
};

// Owning device allocation.
class DeviceBuffer {
public:
    DeviceBuffer() = default;
    explicit DeviceBuffer(size_t bytes) { allocate(bytes); }
    ~DeviceBuffer() { release(); }
    DeviceBuffer(DeviceBuffer&& o) noexcept { swap(o); }
    DeviceBuffer& operator=(DeviceBuffer&& o) noexcept {
        if (this != &o) {
            release();
            swap(o);
        }
        return *this;
    }
    DeviceBuffer(const DeviceBuffer&) = delete;
    DeviceBuffer& operator=(const DeviceBuffer&) = delete;

    void allocate(size_t bytes);           // exact size, contents undefined
    void ensure(size_t bytes);             // grow (x1.5) when smaller
    void release();
    void upload(const void* src, size_t bytes, size_t offset = 0, hipStream_t s = nullptr);
    void download(void* dst, size_t bytes, size_t offset = 0, hipStream_t s = nullptr) const;
    void zero(hipStream_t s = nullptr);

    void* get() const { return ptr_; }
    template <typename T> T* as() const { return static_cast<T*>(ptr_); }
    size_t size() const { return bytes_; }

private:
    void swap(DeviceBuffer& o) noexcept {
        std::swap(ptr_, o.ptr_);
        std::swap(bytes_, o.bytes_);
    }
    void* ptr_ = nullptr;
    size_t bytes_ = 0;
};

// Pinned host staging memory (hipHostMalloc), grown on demand.
class PinnedBuffer {
public:
    PinnedBuffer() = default;
    ~PinnedBuffer();
    PinnedBuffer(const PinnedBuffer&) = delete;
    PinnedBuffer& operator=(const PinnedBuffer&) = delete;
    void ensure(size_t bytes);
    void* get() const { return ptr_; }
    size_t size() const { return bytes_; }

private:
    void* ptr_ = nullptr;
    size_t bytes_ = 0;
};

} // namespace mrt
