// One tile-pass op applied to a 2^L-amplitude tile held in LDS by T
// threads, shared by the tile-pass kernel and the persistent trajectory
// kernel. The caller filters tile-uniform controls (gctrl) and batch-member
// restrictions, and synchronises the block after the op.
#pragma once

#include "tile.hpp"

namespace lqcs_dev {

// Iterations for `items` work items over T threads.
template <int items, int T>
constexpr int kIters = (items + T - 1) / T;

// `base` is the tile's global base index (bits outside the tile).
template <int L, int T>
__device__ __forceinline__ void tile_apply(amp_t* tile, const TileOp& op, uint64_t base) {
    constexpr int N = 1 << L;
    constexpr int PER = N / T;
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
            const amp_t d0 = op.m[0], d1 = op.m[1], d2 = op.m[2], d3 = op.m[3];
#pragma unroll
            for (int k = 0; k < PER; ++k) {
                const uint32_t j = threadIdx.x + k * T;
                if ((j & lctrl) != lctrl) continue;
                const uint32_t sel = (g0 | ((j & dl0) ? 1u : 0u)) | (g1 | ((j & dl1) ? 2u : 0u));
                tile[j] = cmul(tile[j], (sel & 2u) ? ((sel & 1u) ? d3 : d2) : ((sel & 1u) ? d1 : d0));
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
}

}  // namespace lqcs_dev
