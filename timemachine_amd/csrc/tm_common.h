// Common host/device utilities for the MI355X (gfx950) MD hot path.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>
#include <stdexcept>
#include <string>

namespace tmamd {

// Validation failure raised by host code; the C-ABI turns it into a status
// code plus tm_last_error() text (the same text the reference puts into its
// std::runtime_error, see timemachine/cpp/src/nonbonded_all_pairs.cu:253-263).
struct Error : public std::runtime_error {
    explicit Error(const std::string &msg) : std::runtime_error(msg) {}
};

// HIP runtime failure (launch error, fault, out of memory).
struct HipError : public std::runtime_error {
    explicit HipError(const std::string &msg) : std::runtime_error(msg) {}
};

#define TM_HIP_CHECK(expr)                                                                                             \
    do {                                                                                                               \
        hipError_t _e = (expr);                                                                                        \
        if (_e != hipSuccess) {                                                                                        \
            throw ::tmamd::HipError(std::string(#expr) + " failed: " + hipGetErrorString(_e) + " at " + __FILE__ +     \
                                    ":" + std::to_string(__LINE__));                                                   \
        }                                                                                                              \
    } while (0)

#define TM_LAUNCH_CHECK() TM_HIP_CHECK(hipGetLastError())

inline int ceil_div(long long a, long long b) { return static_cast<int>((a + b - 1) / b); }

// Several host threads may drive engines at once (a serving thread stepping
// a batch while another prepares the next job on its own stream). HIP
// invalidates a stream capture in one thread when another thread allocates,
// frees, synchronises the device or uses the null stream meanwhile, so those
// operations and every capture hold this lock (recursive: a holder may nest).
inline std::recursive_mutex &device_mutex() {
    static std::recursive_mutex m;
    return m;
}
struct DeviceLock {
    std::lock_guard<std::recursive_mutex> guard{device_mutex()};
};

inline void device_synchronize() {
    DeviceLock lock;
    TM_HIP_CHECK(hipDeviceSynchronize());
}

// Tile width of the neighborlist and of the pair kernel: one wave64 per tile.
constexpr int TILE = 64;

// Coulomb constant used by the reference's unit system (timemachine/constants.py:7).
constexpr double ONE_4PI_EPS0 = 138.935456;

// RAII device allocation.
template <typename T> struct DeviceArray {
    T *ptr = nullptr;
    size_t n = 0;
    bool owned = true; // false: a view of memory this array does not free
    DeviceArray() = default;
    explicit DeviceArray(size_t count) { alloc(count); }
    DeviceArray(const DeviceArray &) = delete;
    DeviceArray &operator=(const DeviceArray &) = delete;
    DeviceArray(DeviceArray &&o) noexcept : ptr(o.ptr), n(o.n), owned(o.owned) {
        o.ptr = nullptr;
        o.n = 0;
        o.owned = true;
    }
    DeviceArray &operator=(DeviceArray &&o) noexcept {
        if (this != &o) {
            release();
            ptr = o.ptr;
            n = o.n;
            owned = o.owned;
            o.ptr = nullptr;
            o.n = 0;
            o.owned = true;
        }
        return *this;
    }
    ~DeviceArray() { release(); }
    void alloc(size_t count) {
        DeviceLock lock;
        release();
        n = count;
        if (count > 0) {
            TM_HIP_CHECK(hipMalloc(reinterpret_cast<void **>(&ptr), count * sizeof(T)));
        }
    }
    // Points at count elements of existing device memory (no allocation;
    // release() forgets the pointer without freeing it).
    void view(T *p, size_t count) {
        release();
        ptr = count > 0 ? p : nullptr;
        n = count;
        owned = false;
    }
    void release() {
        if (ptr && owned) {
            DeviceLock lock;
            (void)hipFree(ptr);
        }
        ptr = nullptr;
        n = 0;
        owned = true;
    }
    size_t bytes() const { return n * sizeof(T); }
    void upload(const T *host, size_t count, hipStream_t s = 0) {
        if (count > n) {
            throw Error("DeviceArray::upload overflow");
        }
        if (count) {
            if (s == nullptr) {
                DeviceLock lock; // null stream
                TM_HIP_CHECK(hipMemcpyAsync(ptr, host, count * sizeof(T), hipMemcpyHostToDevice, s));
            } else {
                TM_HIP_CHECK(hipMemcpyAsync(ptr, host, count * sizeof(T), hipMemcpyHostToDevice, s));
            }
        }
    }
    void download(T *host, size_t count, hipStream_t s = 0) const {
        if (count > n) {
            throw Error("DeviceArray::download overflow");
        }
        if (count) {
            if (s == nullptr) {
                DeviceLock lock; // null stream
                TM_HIP_CHECK(hipMemcpyAsync(host, ptr, count * sizeof(T), hipMemcpyDeviceToHost, s));
            } else {
                TM_HIP_CHECK(hipMemcpyAsync(host, ptr, count * sizeof(T), hipMemcpyDeviceToHost, s));
            }
        }
    }
    void zero(hipStream_t s = 0) {
        if (n) {
            if (s == nullptr) {
                DeviceLock lock; // null stream
                TM_HIP_CHECK(hipMemsetAsync(ptr, 0, n * sizeof(T), s));
            } else {
                TM_HIP_CHECK(hipMemsetAsync(ptr, 0, n * sizeof(T), s));
            }
        }
    }
};

} // namespace tmamd
