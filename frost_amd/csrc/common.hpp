// Shared host-side helpers of libfrost_amd: HIP error handling, device
// buffers and the C-ABI error boundary.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <utility>

#include "frost_amd.h"

namespace fa {

struct OutOfMemory : std::runtime_error {
    using std::runtime_error::runtime_error;
};

inline void hip_check(hipError_t e, const char* what) {
    if (e == hipSuccess) return;
    const std::string msg = std::string(what) + ": " + hipGetErrorString(e);
    if (e == hipErrorOutOfMemory) throw OutOfMemory(msg);
    throw std::runtime_error(msg);
}

#define FA_HIP(expr) ::fa::hip_check((expr), #expr)

// Sizing without a device: while a SizingScope is alive on this thread,
// DeviceBuffer::resize and PinnedBuffer::reserve record their sizes without
// allocating, upload() copies nothing and DeviceGuard leaves the current
// device alone. Constructors run under it report what they would allocate
// (the memory planner) and make no HIP call.
inline thread_local bool t_sizing = false;
inline bool sizing() { return t_sizing; }

class SizingScope {
public:
    SizingScope() : prev_(t_sizing) { t_sizing = true; }
    ~SizingScope() { t_sizing = prev_; }
    SizingScope(const SizingScope&)            = delete;
    SizingScope& operator=(const SizingScope&) = delete;

private:
    bool prev_;
};

// Host-to-device copy of a table (skipped while sizing).
inline void upload(void* dst, const void* src, std::size_t bytes) {
    if (!t_sizing && bytes) FA_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
}

// Owning device allocation (hipMalloc / hipFree) on the current device.
template <typename T> class DeviceBuffer {
public:
    DeviceBuffer() = default;
    explicit DeviceBuffer(std::size_t n) { resize(n); }
    DeviceBuffer(const DeviceBuffer&)            = delete;
    DeviceBuffer& operator=(const DeviceBuffer&) = delete;
    DeviceBuffer(DeviceBuffer&& o) noexcept { swap(o); }
    DeviceBuffer& operator=(DeviceBuffer&& o) noexcept {
        swap(o);
        return *this;
    }
    ~DeviceBuffer() { reset(); }

    void resize(std::size_t n) {
        if (n == n_) return;
        reset();
        if (n == 0) return;
        if (t_sizing) {
            n_ = n;
            return;
        }
        void* p = nullptr;
        FA_HIP(hipMalloc(&p, n * sizeof(T)));
        p_ = static_cast<T*>(p);
        n_ = n;
    }
    // Grows only.
    void reserve(std::size_t n) {
        if (n > n_) resize(n);
    }
    void reset() noexcept {
        if (p_) (void)hipFree(p_);
        p_ = nullptr;
        n_ = 0;
    }
    void swap(DeviceBuffer& o) noexcept {
        std::swap(p_, o.p_);
        std::swap(n_, o.n_);
    }
    T* get() const { return p_; }
    std::size_t size() const { return n_; }
    std::size_t bytes() const { return n_ * sizeof(T); }

private:
    T* p_          = nullptr;
    std::size_t n_ = 0;
};

// Pinned host allocation for asynchronous copies.
template <typename T> class PinnedBuffer {
public:
    PinnedBuffer() = default;
    PinnedBuffer(const PinnedBuffer&)            = delete;
    PinnedBuffer& operator=(const PinnedBuffer&) = delete;
    ~PinnedBuffer() { reset(); }
    void reserve(std::size_t n, unsigned flags = hipHostMallocDefault) {
        if (n <= n_) return;
        reset();
        if (t_sizing) {
            n_ = n;
            return;
        }
        void* p = nullptr;
        FA_HIP(hipHostMalloc(&p, n * sizeof(T), flags));
        p_ = static_cast<T*>(p);
        n_ = n;
    }
    void reset() noexcept {
        if (p_) (void)hipHostFree(p_);
        p_ = nullptr;
        n_ = 0;
    }
    T* get() const { return p_; }
    std::size_t size() const { return n_; }

private:
    T* p_          = nullptr;
    std::size_t n_ = 0;
};

// Makes `device` current for the scope and restores the previous one.
class DeviceGuard {
public:
    explicit DeviceGuard(int device) {
        if (t_sizing) return;
        active_ = true;
        FA_HIP(hipGetDevice(&prev_));
        if (device != prev_) FA_HIP(hipSetDevice(device));
    }
    ~DeviceGuard() {
        if (active_) (void)hipSetDevice(prev_);
    }

private:
    int prev_    = 0;
    bool active_ = false;
};

void set_last_error(const std::string& msg);

// Runs `f`, translating exceptions into the C ABI's status codes.
template <typename F> int guard(F&& f) noexcept {
    try {
        f();
        return FA_OK;
    } catch (const std::invalid_argument& e) {
        set_last_error(e.what());
        return FA_INVALID_ARGUMENT;
    } catch (const OutOfMemory& e) {
        set_last_error(e.what());
        return FA_OUT_OF_MEMORY;
    } catch (const std::bad_alloc& e) {
        set_last_error(std::string("host allocation failed: ") + e.what());
        return FA_OUT_OF_MEMORY;
    } catch (const std::exception& e) {
        set_last_error(e.what());
        return FA_RUNTIME_ERROR;
    } catch (...) {
        set_last_error("unknown error");
        return FA_RUNTIME_ERROR;
    }
}

inline void require(bool ok, const char* msg) {
    if (!ok) throw std::invalid_argument(msg);
}

inline hipStream_t as_stream(void* s) { return static_cast<hipStream_t>(s); }

inline unsigned ceil_div(std::size_t a, std::size_t b) {
    return static_cast<unsigned>((a + b - 1) / b);
}

// Elapsed seconds between two recorded events (0 if either is missing).
double event_seconds(hipEvent_t a, hipEvent_t b);

} // namespace fa
