// Scalar types of the lqcs API (same names as the reference's
// include/lqcs/core/types.hpp:11-16 so callers compile unchanged).
#pragma once

#include <complex>
#include <cstddef>
#include <cstdint>
#include <new>
#include <vector>

namespace lqcs {

using complex_t = std::complex<double>;
using qubit_t = std::uint32_t;
using clbit_t = std::uint32_t;

// Allocator returning storage aligned to `Align` bytes (host mirror of a
// device register; 64-byte aligned like the reference's aligned_vector).
template <class T, std::size_t Align>
struct AlignedAllocator {
    static_assert((Align & (Align - 1)) == 0 && Align >= alignof(T));
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

template <class T>
using aligned_vector = std::vector<T, AlignedAllocator<T, 64>>;

}  // namespace lqcs
