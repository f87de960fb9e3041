// Device memory arena: bump allocation over cached hipMalloc chunks.
//
// A partition call allocates its hierarchy and scratch from the arena and
// rewinds it at the end; chunks stay cached in the context, so repeated calls
// of the same size make no hipMalloc/hipFree calls.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstddef>
#include <vector>

#include "jet_types.h"

namespace jet {

class Arena {
  public:
    struct Mark {
        size_t ci, off;
    };

    Arena() = default;
    Arena(const Arena&) = delete;
    Arena& operator=(const Arena&) = delete;
    ~Arena() { free_all(); }

    template <class T>
    T* alloc(int64_t count) {
        return static_cast<T*>(alloc_bytes(sizeof(T) * (size_t)std::max<int64_t>(count, 1)));
    }

    void* alloc_bytes(size_t b) {
        b = (b + 255) & ~size_t(255);
        while (ci_ < chunks_.size()) {
            if (off_ + b <= chunks_[ci_].size) {
                void* p = chunks_[ci_].base + off_;
                off_ += b;
                return p;
            }
            ci_++;
            off_ = 0;
        }
        size_t total = 0;
        for (auto& c : chunks_) total += c.size;
        size_t sz = std::max<size_t>({b, total, size_t(64) << 20});
        char* base = nullptr;
        hipError_t e = hipMalloc(&base, sz);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            sz = b;
            e = hipMalloc(&base, sz);
            if (e != hipSuccess) {
                (void)hipGetLastError();
                throw std::bad_alloc();
            }
        }
        chunks_.push_back({base, sz});
        ci_ = chunks_.size() - 1;
        off_ = b;
        return base;
    }

    Mark mark() const { return {ci_, off_}; }
    void release(Mark m) {
        ci_ = m.ci;
        off_ = m.off;
    }
    void reset() {
        ci_ = 0;
        off_ = 0;
    }
    size_t reserved_bytes() const {
        size_t t = 0;
        for (auto& c : chunks_) t += c.size;
        return t;
    }
    void free_all() {
        for (auto& c : chunks_) (void)hipFree(c.base);
        chunks_.clear();
        ci_ = off_ = 0;
    }

  private:
    struct Chunk {
        char* base;
        size_t size;
    };
    std::vector<Chunk> chunks_;
    size_t ci_ = 0, off_ = 0;
};

// Scoped rewind of an arena.
struct ArenaScope {
    Arena& a;
    Arena::Mark m;
    explicit ArenaScope(Arena& ar) : a(ar), m(ar.mark()) {}
    ~ArenaScope() { a.release(m); }
};

}  // namespace jet
