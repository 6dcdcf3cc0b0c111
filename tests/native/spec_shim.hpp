// Host stand-ins for the HIP device names that the specialized-pass kernels
// use, so that a generated kernel compiles as host C++ (clang) and runs one
// workgroup at a time: one std::thread per GPU thread, a std::barrier for
// __syncthreads, and one static LDS array per kernel shared by the threads of
// the workgroup running. Test infrastructure: tests/test_spec_pass.py.
#pragma once

#include <barrier>
#include <bit>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

struct double2 {
    double x, y;
};
inline double2 make_double2(double x, double y) { return double2{x, y}; }

struct shim_dim3 {
    unsigned x = 0, y = 0, z = 0;
};
inline thread_local shim_dim3 threadIdx;
inline thread_local shim_dim3 blockIdx;
inline shim_dim3 gridDim;
inline std::barrier<>* shim_barrier = nullptr;

#define __device__
#define __forceinline__ inline
#define __global__
#define __launch_bounds__(n)
#define __shared__ static

inline void __syncthreads() { shim_barrier->arrive_and_wait(); }
inline int __builtin_amdgcn_readfirstlane(int v) { return v; }
inline int __popc(unsigned v) { return __builtin_popcount(v); }
inline long long __double_as_longlong(double v) { return std::bit_cast<long long>(v); }
inline double __longlong_as_double(long long v) { return std::bit_cast<double>(v); }

// Runs kernel k on every workgroup of a (tiles x batch) grid of `threads`
// threads each.
template <class K>
void shim_launch(K k, unsigned tiles, unsigned batch, unsigned threads = 256) {
    gridDim.x = tiles;
    gridDim.y = batch;
    for (unsigned by = 0; by < batch; ++by)
        for (unsigned bx = 0; bx < tiles; ++bx) {
            std::barrier<> bar(threads);
            shim_barrier = &bar;
            std::vector<std::thread> pool;
            pool.reserve(threads);
            for (unsigned t = 0; t < threads; ++t)
                pool.emplace_back([&, t] {
                    threadIdx.x = t;
                    blockIdx.x = bx;
                    blockIdx.y = by;
                    k();
                });
            for (auto& th : pool) th.join();
        }
}
