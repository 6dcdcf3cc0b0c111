// Shared device-side definitions of the gfx950 statevector engine.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

namespace lqcs_dev {

using amp_t = double2;  // complex128, (re, im) interleaved

// Error carrying an lqcs_status code up to the C-ABI boundary.
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

void check_hip(hipError_t e, const char* what);
#define LQCS_HIP(call) ::lqcs_dev::check_hip((call), #call)

hipStream_t stream();
int num_cus();

// ---- complex arithmetic on double2 ------------------------------------------
__host__ __device__ __forceinline__ amp_t cmul(amp_t a, amp_t b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__host__ __device__ __forceinline__ amp_t cadd(amp_t a, amp_t b) {
    return make_double2(a.x + b.x, a.y + b.y);
}
__host__ __device__ __forceinline__ amp_t cscale(amp_t a, double s) { return make_double2(a.x * s, a.y * s); }
__device__ __forceinline__ amp_t cconj(amp_t a) { return make_double2(a.x, -a.y); }
// m0*a0 + m1*a1
__device__ __forceinline__ amp_t cmac2(amp_t m0, amp_t a0, amp_t m1, amp_t a1) {
    return make_double2(m0.x * a0.x - m0.y * a0.y + m1.x * a1.x - m1.y * a1.y,
                        m0.x * a0.y + m0.y * a0.x + m1.x * a1.y + m1.y * a1.x);
}
__host__ __device__ __forceinline__ double cnorm(amp_t a) { return a.x * a.x + a.y * a.y; }

__host__ __device__ __forceinline__ uint64_t insert_zero_bit(uint64_t v, unsigned pos) {
    const uint64_t low = (uint64_t{1} << pos) - 1;
    return ((v & ~low) << 1) | (v & low);
}
__host__ __device__ __forceinline__ uint64_t insert_two_zero_bits(uint64_t v, unsigned lo,
                                                                   unsigned hi) {
    return insert_zero_bit(insert_zero_bit(v, lo), hi);
}

// Non-temporal-free plain 16-byte loads/stores of one amplitude.
__device__ __forceinline__ amp_t ld(const amp_t* p) { return *p; }
__device__ __forceinline__ void st(amp_t* p, amp_t v) { *p = v; }

// ---- launch geometry ---------------------------------------------------------
constexpr int kBlock = 256;
#define LQCS_LAYOUT_MAX 40     // controls + targets of one operation
#define LQCS_CU_LARGE_ROWS 16  // rows per thread of the LDS-staged wide unitary

inline unsigned grid_for(uint64_t items, int per_thread = 1) {
    const uint64_t per_block = uint64_t(kBlock) * per_thread;
    const uint64_t g = (items + per_block - 1) / per_block;
    return static_cast<unsigned>(g ? g : 1);
}

// ---- kernel launchers shared between translation units ----------------------
struct Mat2 { amp_t m[4]; };
struct Mat4 { amp_t m[16]; };

// Launchers of the gate-by-gate kernels. `batch` registers of 2^nq amplitudes
// lie back to back from s; each gets the same operation.
void launch_single(amp_t* s, uint32_t nq, uint32_t t, const Mat2& m, uint32_t batch = 1);
void launch_diag_single(amp_t* s, uint32_t nq, uint32_t t, amp_t d0, amp_t d1,
                        uint32_t batch = 1);
void launch_controlled_single(amp_t* s, uint32_t nq, uint32_t c, uint32_t t, const Mat2& m,
                              uint32_t batch = 1);
void launch_swap(amp_t* s, uint32_t nq, uint32_t a, uint32_t b, uint32_t batch = 1);
void launch_two_qubit(amp_t* s, uint32_t nq, uint32_t q0, uint32_t q1, const Mat4& m,
                      uint32_t batch = 1);
// Dense controlled unitaries with at least this many targets (4..6) run on
// FP64 MFMA; larger values keep them on the vector kernels.
extern uint32_t g_dense_mfma_min_targets;
void launch_controlled_unitary(amp_t* s, uint32_t nq, const uint32_t* ctl, uint32_t nc,
                               const uint32_t* tgt, uint32_t nt, const amp_t* m_host,
                               uint32_t batch = 1);
void launch_controlled_permutation(amp_t* s, uint32_t nq, const uint32_t* ctl, uint32_t nc,
                                   const uint32_t* tgt, uint32_t nt, const uint64_t* perm_host,
                                   uint32_t batch = 1);
void launch_pauli_rotation(amp_t* s, uint32_t nq, uint64_t x, uint64_t zy, uint32_t n_y,
                           double theta, uint32_t batch = 1);
void launch_fill_basis(amp_t* s, uint32_t nq, uint64_t index, uint32_t batch = 1);

// Stream-ordered device scratch: bump allocation from blocks owned by the
// engine, recycled at the start of every C-ABI call (all device work runs on
// one stream, so later copies into recycled memory are ordered after the
// kernels that read it).
void* arena(size_t bytes);
void arena_reset();
// Tables of a captured launch sequence (lqcs_graph_create): while set, stage()
// lays its tables out in memory the graph owns instead of the per-call arena,
// and keeps their bytes on the host for one copy once the capture ends. With
// dev null (the sizing capture, whose graph is never launched) it only records
// sizes and hands out placeholder addresses.
struct GraphTables {
    char* dev = nullptr;
    std::vector<char> host;
};
extern thread_local GraphTables* g_graph_tables;

// Copies a host table into arena memory on the engine stream (into the
// graph's memory during a capture).
template <class T>
T* stage(const T* host, size_t count) {
    if (GraphTables* g = g_graph_tables) {
        const size_t off = (g->host.size() + 255) & ~size_t{255};
        g->host.resize(off + count * sizeof(T) + 1);
        if (count) std::memcpy(g->host.data() + off, host, count * sizeof(T));
        return reinterpret_cast<T*>(g->dev ? g->dev + off : reinterpret_cast<char*>(uintptr_t{1} << 40) + off);
    }
    T* dev = static_cast<T*>(arena(count * sizeof(T) + 1));
    if (count) LQCS_HIP(hipMemcpyAsync(dev, host, count * sizeof(T), hipMemcpyHostToDevice, stream()));
    return dev;
}

}  // namespace lqcs_dev
