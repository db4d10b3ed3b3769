// Device runtime helpers: status codes, HIP error checks, RAII device buffers.
// Role of mas_utils/CudaUtils.cuh:22-148 and the CHECK_* macros of
// GPUHybrid.cu:34-63, except that errors are reported as status codes, never exit().
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <utility>

#include "lss.h"

namespace lss {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

[[noreturn]] inline void fail(int code, const std::string &msg) { throw Error(code, msg); }

#define LSS_HIP(call)                                                                                       \
    do {                                                                                                    \
        hipError_t e_ = (call);                                                                             \
        if (e_ != hipSuccess)                                                                               \
            ::lss::fail(e_ == hipErrorOutOfMemory ? LSS_ERR_OOM : LSS_ERR_HIP,                              \
                        std::string(#call) + " failed: " + hipGetErrorString(e_) + " (" __FILE__ ":" +      \
                            std::to_string(__LINE__) + ")");                                                \
    } while (0)

#define LSS_LAUNCH_CHECK() LSS_HIP(hipGetLastError())

// Owning device allocation of T[n].
template <class T>
class DBuf {
public:
    DBuf() = default;
    explicit DBuf(size_t n) { resize(n); }
    DBuf(const DBuf &) = delete;
    DBuf &operator=(const DBuf &) = delete;
    DBuf(DBuf &&o) noexcept : p_(o.p_), n_(o.n_) { o.p_ = nullptr, o.n_ = 0; }
    DBuf &operator=(DBuf &&o) noexcept
    {
        if (this != &o) {
            release();
            p_ = o.p_, n_ = o.n_;
            o.p_ = nullptr, o.n_ = 0;
        }
        return *this;
    }
    ~DBuf() { release(); }

    void resize(size_t n)
    {
        if (n == n_ && p_) return;
        release();
        if (n) {
            void *q = nullptr;
            LSS_HIP(hipMalloc(&q, n * sizeof(T)));
            p_ = static_cast<T *>(q);
        }
        n_ = n;
    }
    void release()
    {
        if (p_) (void)hipFree(p_);
        p_ = nullptr;
        n_ = 0;
    }
    void zero(hipStream_t s)
    {
        if (n_) LSS_HIP(hipMemsetAsync(p_, 0, n_ * sizeof(T), s));
    }
    void upload(const T *h, size_t n, hipStream_t s)
    {
        resize(n);
        if (n) LSS_HIP(hipMemcpyAsync(p_, h, n * sizeof(T), hipMemcpyHostToDevice, s));
    }
    void download(T *h, size_t n, hipStream_t s) const
    {
        if (n) LSS_HIP(hipMemcpyAsync(h, p_, n * sizeof(T), hipMemcpyDeviceToHost, s));
    }
    T *get() { return p_; }
    const T *get() const { return p_; }
    size_t size() const { return n_; }
    size_t bytes() const { return n_ * sizeof(T); }
    explicit operator bool() const { return p_ != nullptr; }

private:
    T *p_ = nullptr;
    size_t n_ = 0;
};

inline int div_up(int64_t a, int64_t b) { return static_cast<int>((a + b - 1) / b); }

}  // namespace lss
