// Thin HIP helpers: error checks and an owning device array.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <iterator>
#include <map>
#include <mutex>
#include <new>
#include <stdexcept>
#include <string>
#include <unordered_map>

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

// First-fit sub-allocator over one region (device or pinned host memory),
// 256-byte aligned blocks, neighbours coalesced on release.  Thread-safe.
class Arena {
public:
    static constexpr size_t kAlign = 256;
    enum Kind { kDevice, kPinnedHost, kBorrowed };
    Arena(size_t bytes, Kind kind) : kind_(kind)
    {
        cap_ = round(bytes);
        if (cap_ == 0) return;
        if (kind == kDevice)
            PARODE_HIP(hipMalloc(reinterpret_cast<void**>(&base_), cap_));
        else if (kind == kPinnedHost)
            PARODE_HIP(hipHostMalloc(reinterpret_cast<void**>(&base_), cap_, hipHostMallocDefault));
        else
            throw std::invalid_argument("Arena: borrowed memory needs a base address");
        free_[0] = cap_;
    }
    // Carves blocks out of memory the caller owns (256-byte aligned base);
    // allocates and frees no device memory.
    Arena(void* base, size_t bytes) : kind_(kBorrowed)
    {
        if (reinterpret_cast<uintptr_t>(base) % kAlign)
            throw std::invalid_argument("Arena: borrowed base not 256-byte aligned");
        base_ = static_cast<char*>(base);
        cap_ = bytes / kAlign * kAlign;
        if (cap_) free_[0] = cap_;
    }
    ~Arena()
    {
        if (!base_) return;
        if (kind_ == kDevice)
            (void)hipFree(base_);
        else if (kind_ == kPinnedHost)
            (void)hipHostFree(base_);
    }
    Arena(const Arena&) = delete;
    Arena& operator=(const Arena&) = delete;

    static size_t round(size_t b) { return (b + kAlign - 1) / kAlign * kAlign; }
    // a block of at least `bytes`, or nullptr when no free range fits
    void* take(size_t bytes)
    {
        const size_t b = round(bytes ? bytes : 1);
        std::lock_guard<std::mutex> g(mu_);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < b) continue;
            const size_t off = it->first, rest = it->second - b;
            free_.erase(it);
            if (rest) free_[off + b] = rest;
            live_[off] = b;
            used_ += b;
            return base_ + off;
        }
        return nullptr;
    }
    void give(void* p)
    {
        if (!p) return;
        std::lock_guard<std::mutex> g(mu_);
        const size_t off = static_cast<char*>(p) - base_;
        auto lv = live_.find(off);
        if (lv == live_.end()) return;
        size_t start = off, size = lv->second;
        used_ -= size;
        live_.erase(lv);
        auto next = free_.lower_bound(start);
        if (next != free_.end() && next->first == start + size) {
            size += next->second;
            next = free_.erase(next);
        }
        if (next != free_.begin()) {
            auto prev = std::prev(next);
            if (prev->first + prev->second == start) {
                start = prev->first;
                size += prev->second;
                free_.erase(prev);
            }
        }
        free_[start] = size;
    }
    size_t capacity() const { return cap_; }
    size_t in_use()
    {
        std::lock_guard<std::mutex> g(mu_);
        return used_;
    }
    // the largest block take() could return now
    size_t largest_free()
    {
        std::lock_guard<std::mutex> g(mu_);
        size_t best = 0;
        for (auto& f : free_) best = f.second > best ? f.second : best;
        return best;
    }

private:
    Kind kind_;
    char* base_ = nullptr;
    size_t cap_ = 0, used_ = 0;
    std::map<size_t, size_t> free_;             // offset -> size
    std::unordered_map<size_t, size_t> live_;   // offset -> size
    std::mutex mu_;
};

// Device allocation owned by its holder (hipMalloc, or a block of an Arena);
// moves, never copies.
template <typename T>
class DevArray {
public:
    DevArray() = default;
    explicit DevArray(size_t count, Arena* arena = nullptr) { alloc(count, arena); }
    ~DevArray() { release(); }
    DevArray(const DevArray&) = delete;
    DevArray& operator=(const DevArray&) = delete;
    DevArray(DevArray&& o) noexcept : p_(o.p_), n_(o.n_), arena_(o.arena_)
    {
        o.p_ = nullptr; o.n_ = 0; o.arena_ = nullptr;
    }
    DevArray& operator=(DevArray&& o) noexcept
    {
        if (this != &o) {
            release();
            p_ = o.p_; n_ = o.n_; arena_ = o.arena_;
            o.p_ = nullptr; o.n_ = 0; o.arena_ = nullptr;
        }
        return *this;
    }
    void alloc(size_t count, Arena* arena = nullptr)
    {
        if (count == n_ && p_ && arena == arena_) return;
        release();
        if (count == 0) return;
        if (arena) {
            p_ = static_cast<T*>(arena->take(count * sizeof(T)));
            if (!p_) throw std::bad_alloc();
            arena_ = arena;
        } else {
            PARODE_HIP(hipMalloc(reinterpret_cast<void**>(&p_), count * sizeof(T)));
        }
        n_ = count;
    }
    void release()
    {
        if (p_) {
            if (arena_)
                arena_->give(p_);
            else
                (void)hipFree(p_);
        }
        p_ = nullptr;
        n_ = 0;
        arena_ = nullptr;
    }
    T* get() const { return p_; }
    size_t size() const { return n_; }
    size_t bytes() const { return n_ * sizeof(T); }

private:
    T* p_ = nullptr;
    size_t n_ = 0;
    Arena* arena_ = nullptr;
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
