// LDS-tiled multi-gate passes.
//
// A pass picks a set S of L qubits (always containing the lowest `low` qubits,
// so every tile is made of 2^low-amplitude contiguous runs). For each of the
// 2^(n-L) settings of the other qubits a workgroup loads the 2^L amplitudes
// of that tile into LDS, applies every queued gate whose non-diagonal support
// lies in S, and writes the tile back: one HBM read+write for many gates.
//
// Qubits outside S are constant inside a tile, so controls and diagonal
// factors on them are tile-uniform: a control outside S decides whether the
// whole tile takes the gate, and a phase depending on a qubit outside S reads
// that bit from the tile's base index.
#include "common.hpp"
#include "tile.hpp"
#include "tile_apply.hpp"

namespace lqcs_dev {
namespace {

__device__ __forceinline__ uint64_t pdep64(uint64_t v, uint64_t mask) {
    uint64_t r = 0;
    for (uint64_t bit = 1; mask; bit <<= 1) {
        const uint64_t low = mask & (~mask + 1);
        if (v & bit) r |= low;
        mask ^= low;
    }
    return r;
}

template <int L, int T>
__global__ void __launch_bounds__(T) k_tile_pass(amp_t* __restrict__ s, uint32_t nq,
                                                 uint64_t smask, uint64_t smask_lo,
                                                 uint64_t smask_hi,
                                                 const TileOp* __restrict__ ops,
                                                 uint32_t n_ops, uint64_t tile0) {
    constexpr int N = 1 << L;
    constexpr int PER = N / T;
    __shared__ amp_t tile[N];

    amp_t* b = s + (uint64_t(blockIdx.y) << nq);
    const uint64_t full = (nq == 64) ? ~uint64_t{0} : ((uint64_t{1} << nq) - 1);
    const uint64_t base = pdep64(tile0 + blockIdx.x, full & ~smask);
    const uint64_t lo = pdep64(threadIdx.x, smask_lo);

    {
        amp_t v[PER];
#pragma unroll
        for (int k = 0; k < PER; ++k) v[k] = ld(b + (base | lo | pdep64(k, smask_hi)));
#pragma unroll
        for (int k = 0; k < PER; ++k) tile[threadIdx.x + k * T] = v[k];
    }
    __syncthreads();

    for (uint32_t o = 0; o < n_ops; ++o) {
        const TileOp& op = ops[o];
        if ((base & op.gctrl) != op.gctrl) continue;  // tile-uniform control
        if (op.member >= 0 && (op.member == static_cast<int32_t>(blockIdx.y)) != (op.only != 0)) continue;
        tile_apply<L, T>(tile, op, base);
        __syncthreads();
    }

    {
        amp_t v[PER];
#pragma unroll
        for (int k = 0; k < PER; ++k) v[k] = tile[threadIdx.x + k * T];
#pragma unroll
        for (int k = 0; k < PER; ++k) st(b + (base | lo | pdep64(k, smask_hi)), v[k]);
    }
}

// Splits the tile mask: the lowest log2(T) selected bits are indexed by the
// thread id, the rest by the per-thread item counter.
void split_mask(uint64_t smask, int log2t, uint64_t& lo, uint64_t& hi) {
    lo = 0;
    uint64_t m = smask;
    for (int i = 0; i < log2t && m; ++i) {
        const uint64_t bit = m & (~m + 1);
        lo |= bit;
        m ^= bit;
    }
    hi = m;
}

template <int L, int T>
void launch_t(amp_t* s, uint32_t nq, uint64_t smask, const TileOp* ops, uint32_t n_ops,
              uint32_t batch) {
    uint64_t lo, hi;
    split_mask(smask, __builtin_ctz(T), lo, hi);
    const uint64_t tiles = uint64_t{1} << (nq - L);
    // One workgroup per tile, launched in chunks that keep every grid within
    // the device's 2^32 work-items.
    constexpr uint64_t kChunk = uint64_t{1} << 22;
    for (uint64_t t0 = 0; t0 < tiles; t0 += kChunk) {
        const uint64_t n = std::min(kChunk, tiles - t0);
        k_tile_pass<L, T><<<dim3(static_cast<unsigned>(n), batch), T, 0, stream()>>>(s, nq, smask, lo, hi, ops,
                                                                                      n_ops, t0);
        LQCS_HIP(hipGetLastError());
    }
}

}  // namespace

void launch_tile_pass(amp_t* s, uint32_t nq, uint32_t L, uint64_t smask, const TileOp* d_ops,
                      uint32_t n_ops, uint32_t batch) {
    if (static_cast<uint32_t>(__builtin_popcountll(smask)) != L || L > nq)
        throw Error(3, "tile pass: mask does not select L qubits");
    switch (L) {
        case 6: launch_t<6, 64>(s, nq, smask, d_ops, n_ops, batch); return;
        case 7: launch_t<7, 64>(s, nq, smask, d_ops, n_ops, batch); return;
        case 8: launch_t<8, 64>(s, nq, smask, d_ops, n_ops, batch); return;
        case 9: launch_t<9, 128>(s, nq, smask, d_ops, n_ops, batch); return;
        case 10: launch_t<10, 256>(s, nq, smask, d_ops, n_ops, batch); return;
        case 11: launch_t<11, 256>(s, nq, smask, d_ops, n_ops, batch); return;
        case 12: launch_t<12, 256>(s, nq, smask, d_ops, n_ops, batch); return;
        default: throw Error(3, "tile pass: unsupported tile width");
    }
}

}  // namespace lqcs_dev
