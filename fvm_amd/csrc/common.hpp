// Shared infrastructure of the fvm_amd HIP library: error state, device arrays,
// launch helpers and device-wide scans (hipCUB).
#pragma once

#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <map>
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

// Caching device allocator of one handle. All work of a handle runs on its one
// stream, so a block returned here can be handed out again at once: later
// kernels on that stream are ordered after the earlier users of the block.
// Avoids hipMalloc/hipFree (which synchronize and map pages) on every setup.
struct Pool {
    std::multimap<size_t, void *> free_blocks;
    size_t cached = 0;
    static size_t round(size_t b) {
        if (b <= (1u << 20)) {
            size_t r = 256;
            while (r < b) r <<= 1;
            return r;
        }
        const size_t g = 2u << 20;
        return (b + g - 1) / g * g;
    }
    void *get(size_t bytes, size_t *cap) {
        const size_t r = round(bytes);
        auto it = free_blocks.lower_bound(r);
        if (it != free_blocks.end() && it->first <= r + r / 4) {
            void *p = it->second;
            *cap = it->first;
            cached -= it->first;
            free_blocks.erase(it);
            return p;
        }
        void *p = nullptr;
        hipError_t e = hipMalloc(&p, r);
        if (e != hipSuccess) {
            trim();   // give cached blocks back and retry once
            e = hipMalloc(&p, r);
        }
        if (e != hipSuccess)
            throw Error(-5, std::string("hipMalloc of ") + std::to_string(r) + " bytes failed: " +
                                hipGetErrorString(e));
        *cap = r;
        return p;
    }
    void put(void *p, size_t cap) {
        free_blocks.emplace(cap, p);
        cached += cap;
    }
    void trim() {
        for (auto &kv : free_blocks) (void)hipFree(kv.second);
        free_blocks.clear();
        cached = 0;
    }
    ~Pool() { trim(); }
};

// The pool that DArr allocations of the current thread draw from (nullptr:
// plain hipMalloc). Set for the duration of an entry point by PoolScope.
inline Pool *&current_pool() {
    static thread_local Pool *p = nullptr;
    return p;
}
struct PoolScope {
    Pool *prev;
    explicit PoolScope(Pool *p) : prev(current_pool()) { current_pool() = p; }
    ~PoolScope() { current_pool() = prev; }
};

// Owning device array. Not copyable; movable.
template <class T>
struct DArr {
    T *p = nullptr;
    size_t n = 0;
    size_t cap = 0;
    Pool *pool = nullptr;
    DArr() = default;
    explicit DArr(size_t count) { alloc(count); }
    DArr(const DArr &) = delete;
    DArr &operator=(const DArr &) = delete;
    DArr(DArr &&o) noexcept : p(o.p), n(o.n), cap(o.cap), pool(o.pool) { o.p = nullptr; o.n = 0; o.cap = 0; o.pool = nullptr; }
    DArr &operator=(DArr &&o) noexcept {
        if (this != &o) {
            release();
            p = o.p; n = o.n; cap = o.cap; pool = o.pool;
            o.p = nullptr; o.n = 0; o.cap = 0; o.pool = nullptr;
        }
        return *this;
    }
    ~DArr() { release(); }
    void alloc(size_t count) {
        release();
        n = count;
        if (!count) return;
        pool = current_pool();
        if (pool) {
            p = (T *)pool->get(count * sizeof(T), &cap);
        } else {
            hipError_t e = hipMalloc((void **)&p, count * sizeof(T));
            if (e != hipSuccess) {
                p = nullptr; n = 0;
                throw Error(-5, std::string("hipMalloc of ") + std::to_string(count * sizeof(T)) +
                                    " bytes failed: " + hipGetErrorString(e));
            }
            cap = count * sizeof(T);
        }
    }
    void ensure(size_t count) { if (n < count) alloc(count); }
    void release() {
        if (p) {
            if (pool) pool->put(p, cap);
            else (void)hipFree(p);
        }
        p = nullptr; n = 0; cap = 0; pool = nullptr;
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

// Pinned staging word of the calling host thread (a copy into pageable memory
// goes through a bounce buffer); held until the process ends.
inline void *pinned_word() {
    static thread_local void *p = nullptr;
    if (!p) FVM_HIP(hipHostMalloc(&p, 64, hipHostMallocDefault));
    return p;
}

template <class T>
inline T read_scalar(const T *dptr, hipStream_t s) {
    static_assert(sizeof(T) <= 64, "scalar too large");
    T *v = (T *)pinned_word();
    FVM_HIP(hipMemcpyAsync(v, dptr, sizeof(T), hipMemcpyDeviceToHost, s));
    FVM_HIP(hipStreamSynchronize(s));
    return *v;
}

// Pipelined reads of a device counter for loops that run batches of kernels
// until the counter reaches zero: the counter after batch b is copied to a
// pinned slot and read while batch b + 1 is already queued, so the device never
// waits for the host.
struct CounterProbe {
    int *hv = nullptr;
    hipEvent_t ev[2] = {nullptr, nullptr};
    void init() {
        if (hv) return;
        FVM_HIP(hipHostMalloc((void **)&hv, 2 * sizeof(int), hipHostMallocDefault));
        FVM_HIP(hipEventCreateWithFlags(&ev[0], hipEventDisableTiming));
        FVM_HIP(hipEventCreateWithFlags(&ev[1], hipEventDisableTiming));
    }
    void post(int slot, const int *d, hipStream_t s) {
        FVM_HIP(hipMemcpyAsync(&hv[slot & 1], d, sizeof(int), hipMemcpyDeviceToHost, s));
        FVM_HIP(hipEventRecord(ev[slot & 1], s));
    }
    int wait(int slot) {
        FVM_HIP(hipEventSynchronize(ev[slot & 1]));
        return hv[slot & 1];
    }
    void release() {   // with the device of the events current
        for (auto &e : ev)
            if (e) { (void)hipEventDestroy(e); e = nullptr; }
        if (hv) { (void)hipHostFree(hv); hv = nullptr; }
    }
};

}  // namespace fvm
