// Shared infrastructure of the fvm_amd HIP library: error state, device arrays,
// launch helpers and device-wide scans (hipCUB).
#pragma once

#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

namespace fvm {

// Internal errors are C++ exceptions that the C ABI turns into status codes.
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

void set_last_error(const std::string &msg);

#define FVM_HIP(call)                                                                        \
    do {                                                                                     \
        hipError_t e_ = (call);                                                              \
        if (e_ != hipSuccess)                                                                \
            throw ::fvm::Error(-2, std::string(#call " failed: ") + hipGetErrorString(e_) + \
                                       " (" + __FILE__ + ":" + std::to_string(__LINE__) + ")"); \
    } while (0)

#define FVM_CHECK(cond, code, msg)                                                           \
    do {                                                                                     \
        if (!(cond)) throw ::fvm::Error((code), (msg));                                      \
    } while (0)

// Owning device array (hipMalloc'd). Not copyable; movable.
template <class T>
struct DArr {
    T *p = nullptr;
    size_t n = 0;
    DArr() = default;
    explicit DArr(size_t count) { alloc(count); }
    DArr(const DArr &) = delete;
    DArr &operator=(const DArr &) = delete;
    DArr(DArr &&o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
    DArr &operator=(DArr &&o) noexcept {
        if (this != &o) { release(); p = o.p; n = o.n; o.p = nullptr; o.n = 0; }
        return *this;
    }
    ~DArr() { release(); }
    void alloc(size_t count) {
        release();
        n = count;
        if (count) {
            hipError_t e = hipMalloc((void **)&p, count * sizeof(T));
            if (e != hipSuccess) {
                p = nullptr; n = 0;
                throw Error(-5, std::string("hipMalloc of ") + std::to_string(count * sizeof(T)) +
                                    " bytes failed: " + hipGetErrorString(e));
            }
        }
    }
    void ensure(size_t count) { if (n < count) alloc(count); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr; n = 0;
    }
    size_t bytes() const { return n * sizeof(T); }
};

inline unsigned grid_for(long long n, int block, unsigned cap = 1u << 20) {
    long long g = (n + block - 1) / block;
    if (g < 1) g = 1;
    if (g > (long long)cap) g = cap;
    return (unsigned)g;
}

// Device-wide exclusive scan of int32 with a reusable temp buffer.
struct Scanner {
    DArr<unsigned char> tmp;
    void exclusive(const int *in, int *out, int n, hipStream_t s) {
        size_t bytes = 0;
        FVM_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, in, out, n, s));
        tmp.ensure(bytes + 16);
        FVM_HIP(hipcub::DeviceScan::ExclusiveSum(tmp.p, bytes, in, out, n, s));
    }
    void inclusive(const int *in, int *out, int n, hipStream_t s) {
        size_t bytes = 0;
        FVM_HIP(hipcub::DeviceScan::InclusiveSum(nullptr, bytes, in, out, n, s));
        tmp.ensure(bytes + 16);
        FVM_HIP(hipcub::DeviceScan::InclusiveSum(tmp.p, bytes, in, out, n, s));
    }
};

template <class T>
inline T read_scalar(const T *dptr, hipStream_t s) {
    T v;
    FVM_HIP(hipMemcpyAsync(&v, dptr, sizeof(T), hipMemcpyDeviceToHost, s));
    FVM_HIP(hipStreamSynchronize(s));
    return v;
}

}  // namespace fvm
