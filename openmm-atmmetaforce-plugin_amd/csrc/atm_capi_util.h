// Error handling and device buffers shared by the host files of the C ABI
// (no exceptions cross the ABI: every entry point runs under guarded()).
#pragma once
#include <hip/hip_runtime.h>
#include <algorithm>
#include <stdexcept>
#include <string>

#include "../../include/atm_mi355x.h"

namespace atm {

// The message atm_last_error() returns (one per calling thread).
std::string& last_error();

struct AtmError : std::runtime_error {
    int code;
    AtmError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define HIPCHK(expr)                                                                                   \
    do {                                                                                               \
        hipError_t e_ = (expr);                                                                        \
        if (e_ != hipSuccess)                                                                          \
            throw ::atm::AtmError(ATM_ERR_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_));  \
    } while (0)

[[noreturn]] inline void fail(int code, const std::string& msg) { throw AtmError(code, msg); }

template <typename F>
int guarded(F&& f) {
    try {
        f();
        return ATM_OK;
    } catch (const AtmError& e) {
        last_error() = e.what();
        return e.code;
    } catch (const std::exception& e) {
        last_error() = e.what();
        return ATM_ERR_INVALID_ARGUMENT;
    }
}

// Device allocation owned by a host object.
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    void alloc(size_t n) {
        release();
        bytes = std::max<size_t>(n, 256);
        HIPCHK(hipMalloc(&p, bytes));
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    template <typename T> T* as() const { return static_cast<T*>(p); }
};

}  // namespace atm
