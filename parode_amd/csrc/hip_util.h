// Thin HIP helpers: error checks and an owning device array.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <stdexcept>
#include <string>

namespace parode {

struct HipError : std::runtime_error {
    hipError_t code;
    HipError(hipError_t c, const std::string& what)
        : std::runtime_error(what + ": " + hipGetErrorString(c)), code(c) {}
};

inline void hip_check(hipError_t e, const char* what)
{
    if (e != hipSuccess) throw HipError(e, what);
}

#define PARODE_HIP(call) ::parode::hip_check((call), #call)

// Device allocation owned by its holder; moves, never copies.
template <typename T>
class DevArray {
public:
    DevArray() = default;
    explicit DevArray(size_t count) { alloc(count); }
    ~DevArray() { release(); }
    DevArray(const DevArray&) = delete;
    DevArray& operator=(const DevArray&) = delete;
    DevArray(DevArray&& o) noexcept : p_(o.p_), n_(o.n_) { o.p_ = nullptr; o.n_ = 0; }
    DevArray& operator=(DevArray&& o) noexcept
    {
        if (this != &o) {
            release();
            p_ = o.p_; n_ = o.n_;
            o.p_ = nullptr; o.n_ = 0;
        }
        return *this;
    }
    void alloc(size_t count)
    {
        if (count == n_ && p_) return;
        release();
        if (count == 0) return;
        PARODE_HIP(hipMalloc(reinterpret_cast<void**>(&p_), count * sizeof(T)));
        n_ = count;
    }
    void release()
    {
        if (p_) (void)hipFree(p_);
        p_ = nullptr;
        n_ = 0;
    }
    T* get() const { return p_; }
    size_t size() const { return n_; }
    size_t bytes() const { return n_ * sizeof(T); }

private:
    T* p_ = nullptr;
    size_t n_ = 0;
};

inline long next_pow2(long v)
{
    long p = 1;
    while (p < v) p <<= 1;
    return p;
}

inline int ilog2(long v)
{
    int l = 0;
    while ((1L << l) < v) ++l;
    return l;
}

}  // namespace parode
