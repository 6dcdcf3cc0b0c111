// Allocator returning storage aligned to a fixed byte boundary (same name and
// template parameters as the reference's include/lqcs/core/aligned_allocator.hpp:6-39,
// so callers that name it compile unchanged). Host mirrors of device registers
// use it with 64-byte alignment.
#pragma once

#include <cstddef>
#include <new>

namespace lqcs {

template <class T, std::size_t Align>
struct AlignedAllocator {
    static_assert((Align & (Align - 1)) == 0, "alignment must be a power of two");
    static_assert(Align >= alignof(T), "alignment must be at least alignof(T)");
    using value_type = T;
    template <class U>
    struct rebind {
        using other = AlignedAllocator<U, Align>;
    };
    AlignedAllocator() noexcept = default;
    template <class U>
    AlignedAllocator(const AlignedAllocator<U, Align>&) noexcept {}
    T* allocate(std::size_t n) {
        return static_cast<T*>(::operator new(n * sizeof(T), std::align_val_t{Align}));
    }
    void deallocate(T* p, std::size_t) noexcept { ::operator delete(p, std::align_val_t{Align}); }
    friend bool operator==(const AlignedAllocator&, const AlignedAllocator&) noexcept { return true; }
};

}  // namespace lqcs
