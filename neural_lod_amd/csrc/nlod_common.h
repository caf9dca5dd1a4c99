// Shared host-side plumbing of libnlod: error capture for the C ABI and HIP
// call checking.  Device helpers live next to the kernels that use them.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>

#include "nlod.h"

namespace nlod {

// Error thrown inside the library and turned into a status code at the ABI.
struct Error : std::runtime_error {
    int status;
    Error(int s, const std::string& msg) : std::runtime_error(msg), status(s) {}
};

void set_last_error(const std::string& msg);
void clear_last_error();

[[noreturn]] inline void fail(int status, const std::string& msg) { throw Error(status, msg); }

inline void check_hip(hipError_t e, const char* what, const char* file, int line) {
    if (e != hipSuccess) {
        fail(e == hipErrorOutOfMemory ? NLOD_ERR_OUT_OF_MEMORY : NLOD_ERR_HIP,
             std::string(what) + " failed: " + hipGetErrorString(e) + " (" + file + ":" +
                 std::to_string(line) + ")");
    }
}

#define NLOD_HIP(call) ::nlod::check_hip((call), #call, __FILE__, __LINE__)
#define NLOD_CHECK(cond, msg)                                                     \
    do {                                                                          \
        if (!(cond)) ::nlod::fail(NLOD_ERR_INVALID_ARGUMENT, std::string(msg));  \
    } while (0)
#define NLOD_LAUNCH_CHECK() NLOD_HIP(hipGetLastError())

// Runs f() and converts any exception into an ABI status.
template <typename F>
int guarded(F&& f) {
    try {
        clear_last_error();
        f();
        return NLOD_OK;
    } catch (const Error& e) {
        set_last_error(e.what());
        return e.status;
    } catch (const std::bad_alloc&) {
        set_last_error("host allocation failed");
        return NLOD_ERR_OUT_OF_MEMORY;
    } catch (const std::exception& e) {
        set_last_error(e.what());
        return NLOD_ERR_INVALID_ARGUMENT;
    } catch (...) {
        set_last_error("unknown error");
        return NLOD_ERR_INVALID_ARGUMENT;
    }
}

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

inline uint32_t div_round_up(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

// Grow-only device allocation (never reallocated while a stream may use it
// within one call; growth happens before launches).  An owner that captures
// the buffer's address into a hipGraph attaches an epoch counter: every free
// (including the free of a growth) increments it, so a graph stamped with an
// older epoch may hold a dangling address and must be captured again.
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    uint64_t* epoch = nullptr;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    void release() {
        if (p) {
            (void)hipFree(p);
            if (epoch) ++*epoch;
        }
        p = nullptr;
        bytes = 0;
    }
    void ensure(size_t b) {
        if (b <= bytes) return;
        release();
        NLOD_HIP(hipMalloc(&p, b));
        bytes = b;
    }
    template <typename T>
    T* as() const { return static_cast<T*>(p); }
};

// Compute units of the current device.
int cu_count();
// Resident blocks of `kernel` per CU at `block` threads and `smem` dynamic LDS
// bytes (also raises the kernel's dynamic-LDS limit to smem); cached per
// (device, kernel, block, smem).
int occupancy_blocks(const void* kernel, int block, size_t smem);

}  // namespace nlod
