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

// Iterations for `items` work items over T threads.
template <int items, int T>
constexpr int kIters = (items + T - 1) / T;

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
        const uint32_t lctrl = op.lctrl;
        switch (op.kind) {
            case TK_DENSE1: {
                const uint32_t t0 = op.t0;
                const amp_t m0 = op.m[0], m1 = op.m[1], m2 = op.m[2], m3 = op.m[3];
#pragma unroll
                for (int k = 0; k < kIters<N / 2, T>; ++k) {
                    const uint32_t p = threadIdx.x + k * T;
                    if (N / 2 < T && p >= N / 2) break;
                    const uint32_t i0 = static_cast<uint32_t>(insert_zero_bit(p, t0));
                    if ((i0 & lctrl) != lctrl) continue;
                    const uint32_t i1 = i0 | (1u << t0);
                    const amp_t a0 = tile[i0], a1 = tile[i1];
                    tile[i0] = cmac2(m0, a0, m1, a1);
                    tile[i1] = cmac2(m2, a0, m3, a1);
                }
                break;
            }
            case TK_DENSE2: {
                const uint32_t t0 = op.t0, t1 = op.t1;
                const uint32_t lo2 = t0 < t1 ? t0 : t1, hi2 = t0 < t1 ? t1 : t0;
#pragma unroll
                for (int k = 0; k < kIters<N / 4, T>; ++k) {
                    const uint32_t p = threadIdx.x + k * T;
                    if (N / 4 < T && p >= N / 4) break;
                    const uint32_t i00 = static_cast<uint32_t>(insert_two_zero_bits(p, lo2, hi2));
                    if ((i00 & lctrl) != lctrl) continue;
                    const uint32_t idx[4] = {i00, i00 | (1u << t0), i00 | (1u << t1),
                                             i00 | (1u << t0) | (1u << t1)};
                    amp_t a[4];
#pragma unroll
                    for (int r = 0; r < 4; ++r) a[r] = tile[idx[r]];
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        amp_t acc = cmac2(op.m[4 * r], a[0], op.m[4 * r + 1], a[1]);
                        tile[idx[r]] = cadd(acc, cmac2(op.m[4 * r + 2], a[2], op.m[4 * r + 3], a[3]));
                    }
                }
                break;
            }
            case TK_DIAG: {
                const uint32_t g0 = (base & op.dg0) ? 1u : 0u;
                const uint32_t g1 = (base & op.dg1) ? 2u : 0u;
                const uint32_t dl0 = op.dl0, dl1 = op.dl1;
                const amp_t d[4] = {op.m[0], op.m[1], op.m[2], op.m[3]};
#pragma unroll
                for (int k = 0; k < PER; ++k) {
                    const uint32_t j = threadIdx.x + k * T;
                    if ((j & lctrl) != lctrl) continue;
                    const uint32_t sel = (g0 | ((j & dl0) ? 1u : 0u)) | (g1 | ((j & dl1) ? 2u : 0u));
                    tile[j] = cmul(tile[j], d[sel]);
                }
                break;
            }
            case TK_PAULI: {
                const uint32_t xl = op.xl, zl = op.zl;
                const uint32_t gpar = __popcll(base & op.zg) & 1u;
                const amp_t c = op.m[0], yp = op.m[1];
                if (xl == 0) {
                    const amp_t dp = cadd(c, yp), dm = cadd(c, cscale(yp, -1.0));
#pragma unroll
                    for (int k = 0; k < PER; ++k) {
                        const uint32_t j = threadIdx.x + k * T;
                        tile[j] = cmul(tile[j], ((__popc(j & zl) ^ gpar) & 1u) ? dm : dp);
                    }
                } else {
                    const uint32_t lsb = __builtin_ctz(xl);
#pragma unroll
                    for (int k = 0; k < kIters<N / 2, T>; ++k) {
                        const uint32_t p = threadIdx.x + k * T;
                        if (N / 2 < T && p >= N / 2) break;
                        const uint32_t i = static_cast<uint32_t>(insert_zero_bit(p, lsb));
                        const uint32_t j = i ^ xl;
                        const double si = ((__popc(i & zl) ^ gpar) & 1u) ? -1.0 : 1.0;
                        const double sj = ((__popc(j & zl) ^ gpar) & 1u) ? -1.0 : 1.0;
                        const amp_t ai = tile[i], aj = tile[j];
                        tile[i] = cmac2(c, ai, cscale(yp, sj), aj);
                        tile[j] = cmac2(cscale(yp, si), ai, c, aj);
                    }
                }
                break;
            }
            default: break;
        }
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
