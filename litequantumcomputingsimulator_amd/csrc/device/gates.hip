// Gate-by-gate kernels: one HBM pass per operation, one kernel per entry of
// the reference kernel seam (src/backend/kernels/kernels.hpp:13-62). They are
// the fallback for ops the tiled passes do not take and the baseline the
// planner is measured against.
//
// Every kernel streams complex128 amplitudes with 16-byte loads; a thread owns
// kItems independent groups so each wave keeps several loads in flight.
// blockIdx.y selects a register in a batch (stride = 2^n amplitudes).
#include "common.hpp"

#include <vector>

namespace lqcs_dev {
namespace {

constexpr int kItems = 2;

__device__ __forceinline__ uint64_t group_index(uint64_t blk, int k) {
    return (blk * kItems + k) * kBlock + threadIdx.x;
}
// Grid-stride over blocks of kItems * kBlock groups: launches are capped at
// kMaxGrid workgroups, so registers of any width stay within the launch limits.
#define LQCS_FOR_BLOCKS(blk, groups) \
    for (uint64_t blk = blockIdx.x; blk * (kItems * kBlock) < (groups); blk += gridDim.x)
constexpr uint64_t kMaxGrid = uint64_t{1} << 22;
inline unsigned capped(uint64_t blocks) { return static_cast<unsigned>(std::min<uint64_t>(blocks, kMaxGrid)); }
inline unsigned grid_capped(uint64_t items, int per_thread = 1) { return capped(grid_for(items, per_thread)); }
// Launch errors (a shape beyond the device limits, a missing code object)
// surface at the launch, not at the next synchronisation.
inline void launched() { LQCS_HIP(hipGetLastError()); }

__device__ __forceinline__ amp_t* batch_base(amp_t* s, uint32_t nq) {
    return s + (uint64_t(blockIdx.y) << nq);
}

// gate_kernels.cpp:26-39 -- pairs (i0, i0 | 2^t)
__global__ void __launch_bounds__(kBlock) k_single(amp_t* __restrict__ s, uint32_t nq, uint32_t t,
                                                   Mat2 m) {
    amp_t* b = batch_base(s, nq);
    const uint64_t groups = uint64_t{1} << (nq - 1);
    LQCS_FOR_BLOCKS(blk, groups) {
        amp_t a0[kItems], a1[kItems];
        uint64_t i0[kItems];
#pragma unroll
        for (int k = 0; k < kItems; ++k) {
            const uint64_t g = group_index(blk, k);
            i0[k] = insert_zero_bit(g < groups ? g : 0, t);
            a0[k] = ld(b + i0[k]);
            a1[k] = ld(b + (i0[k] | (uint64_t{1} << t)));
        }
#pragma unroll
        for (int k = 0; k < kItems; ++k) {
            if (group_index(blk, k) >= groups) continue;
            st(b + i0[k], cmac2(m.m[0], a0[k], m.m[1], a1[k]));
            st(b + (i0[k] | (uint64_t{1} << t)), cmac2(m.m[2], a0[k], m.m[3], a1[k]));
        }
    }
}

// :41-47 -- phase multiply, every amplitude
__global__ void __launch_bounds__(kBlock) k_diag_single(amp_t* __restrict__ s, uint32_t nq,
                                                        uint32_t t, amp_t d0, amp_t d1) {
    amp_t* b = batch_base(s, nq);
    const uint64_t n = uint64_t{1} << nq;
    LQCS_FOR_BLOCKS(blk, n) {
        amp_t a[kItems];
#pragma unroll
        for (int k = 0; k < kItems; ++k) {
            const uint64_t i = group_index(blk, k);
            a[k] = ld(b + (i < n ? i : 0));
        }
#pragma unroll
        for (int k = 0; k < kItems; ++k) {
            const uint64_t i = group_index(blk, k);
            if (i >= n) continue;
            st(b + i, cmul(a[k], ((i >> t) & 1) ? d1 : d0));
        }
    }
}

// :49-69 -- only the control=1 half is touched
__global__ void __launch_bounds__(kBlock) k_controlled_single(amp_t* __restrict__ s, uint32_t nq,
                                                              uint32_t c, uint32_t t, Mat2 m) {
    amp_t* b = batch_base(s, nq);
    const unsigned lo = c < t ? c : t, hi = c < t ? t : c;
    const uint64_t groups = uint64_t{1} << (nq - 2);
    LQCS_FOR_BLOCKS(blk, groups) {
        amp_t a0[kItems], a1[kItems];
        uint64_t i0[kItems];
#pragma unroll
        for (int k = 0; k < kItems; ++k) {
            const uint64_t g = group_index(blk, k);
            i0[k] = insert_two_zero_bits(g < groups ? g : 0, lo, hi) | (uint64_t{1} << c);
            a0[k] = ld(b + i0[k]);
            a1[k] = ld(b + (i0[k] | (uint64_t{1} << t)));
        }
#pragma unroll
        for (int k = 0; k < kItems; ++k) {
            if (group_index(blk, k) >= groups) continue;
            st(b + i0[k], cmac2(m.m[0], a0[k], m.m[1], a1[k]));
            st(b + (i0[k] | (uint64_t{1} << t)), cmac2(m.m[2], a0[k], m.m[3], a1[k]));
        }
    }
}

// :71-84 -- exchange |a=1,b=0> <-> |a=0,b=1>
__global__ void __launch_bounds__(kBlock) k_swap(amp_t* __restrict__ s, uint32_t nq, uint32_t qa,
                                                 uint32_t qb) {
    amp_t* b = batch_base(s, nq);
    const unsigned lo = qa < qb ? qa : qb, hi = qa < qb ? qb : qa;
    const uint64_t groups = uint64_t{1} << (nq - 2);
    LQCS_FOR_BLOCKS(blk, groups) {
        amp_t x[kItems], y[kItems];
        uint64_t base[kItems];
#pragma unroll
        for (int k = 0; k < kItems; ++k) {
            const uint64_t g = group_index(blk, k);
            base[k] = insert_two_zero_bits(g < groups ? g : 0, lo, hi);
            x[k] = ld(b + (base[k] | (uint64_t{1} << qa)));
            y[k] = ld(b + (base[k] | (uint64_t{1} << qb)));
        }
#pragma unroll
        for (int k = 0; k < kItems; ++k) {
            if (group_index(blk, k) >= groups) continue;
            st(b + (base[k] | (uint64_t{1} << qa)), y[k]);
            st(b + (base[k] | (uint64_t{1} << qb)), x[k]);
        }
    }
}

// :86-106 -- dense 4x4, local bit 0 <-> q0
__global__ void __launch_bounds__(kBlock) k_two_qubit(amp_t* __restrict__ s, uint32_t nq,
                                                      uint32_t q0, uint32_t q1, Mat4 m) {
    amp_t* b = batch_base(s, nq);
    const unsigned lo = q0 < q1 ? q0 : q1, hi = q0 < q1 ? q1 : q0;
    const uint64_t groups = uint64_t{1} << (nq - 2);
    for (uint64_t g = uint64_t(blockIdx.x) * kBlock + threadIdx.x; g < groups;
         g += uint64_t(gridDim.x) * kBlock) {
        const uint64_t base = insert_two_zero_bits(g, lo, hi);
        const uint64_t idx[4] = {base, base | (uint64_t{1} << q0), base | (uint64_t{1} << q1),
                                 base | (uint64_t{1} << q0) | (uint64_t{1} << q1)};
        amp_t a[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) a[r] = ld(b + idx[r]);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            amp_t acc = cmac2(m.m[4 * r], a[0], m.m[4 * r + 1], a[1]);
            acc = cadd(acc, cmac2(m.m[4 * r + 2], a[2], m.m[4 * r + 3], a[3]));
            st(b + idx[r], acc);
        }
    }
}

// Controlled k-target layout (:118-139): fixed positions sorted ascending.
struct Layout {
    uint32_t fixed[LQCS_LAYOUT_MAX];
    uint32_t n_fixed;
    uint32_t nt;
    uint64_t c_mask;
};

__device__ __forceinline__ uint64_t expand(uint64_t g, const Layout& L) {
    for (uint32_t i = 0; i < L.n_fixed; ++i) g = insert_zero_bit(g, L.fixed[i]);
    return g | L.c_mask;
}

// :143-167 for small k: one thread per group, amplitudes in registers,
// matrix staged in LDS by the block.
template <int K>
__global__ void __launch_bounds__(kBlock) k_cu_small(amp_t* __restrict__ s, uint32_t nq, Layout L,
                                                     const uint64_t* __restrict__ offs,
                                                     const amp_t* __restrict__ mat) {
    constexpr int D = 1 << K;
    __shared__ amp_t sm[D * D];
    for (int i = threadIdx.x; i < D * D; i += blockDim.x) sm[i] = mat[i];
    __syncthreads();
    amp_t* b = batch_base(s, nq);
    const uint64_t groups = uint64_t{1} << (nq - L.n_fixed);
    for (uint64_t g = uint64_t(blockIdx.x) * kBlock + threadIdx.x; g < groups;
         g += uint64_t(gridDim.x) * kBlock) {
        const uint64_t base = expand(g, L);
        amp_t a[D];
#pragma unroll
        for (int l = 0; l < D; ++l) a[l] = ld(b + (base | offs[l]));
#pragma unroll
        for (int r = 0; r < D; ++r) {
            amp_t acc = make_double2(0.0, 0.0);
#pragma unroll
            for (int c = 0; c < D; ++c) acc = cadd(acc, cmul(sm[r * D + c], a[c]));
            st(b + (base | offs[r]), acc);
        }
    }
}

// :143-167 for larger k: one block per group, the group staged in LDS.
__global__ void __launch_bounds__(kBlock) k_cu_large(amp_t* __restrict__ s, uint32_t nq, Layout L,
                                                     const uint64_t* __restrict__ offs,
                                                     const amp_t* __restrict__ mat) {
    extern __shared__ amp_t grp[];
    const uint32_t D = 1u << L.nt;
    amp_t* b = batch_base(s, nq);
    const uint64_t groups = uint64_t{1} << (nq - L.n_fixed);
    for (uint64_t g = blockIdx.x; g < groups; g += gridDim.x) {
        const uint64_t base = expand(g, L);
        for (uint32_t l = threadIdx.x; l < D; l += blockDim.x) grp[l] = ld(b + (base | offs[l]));
        __syncthreads();
        amp_t out[LQCS_CU_LARGE_ROWS];
        int k = 0;
        for (uint32_t r = threadIdx.x; r < D; r += blockDim.x, ++k) {
            amp_t acc = make_double2(0.0, 0.0);
            for (uint32_t c = 0; c < D; ++c) acc = cadd(acc, cmul(mat[uint64_t(r) * D + c], grp[c]));
            out[k] = acc;
        }
        __syncthreads();  // every read of grp is done before the next group lands
        k = 0;
        for (uint32_t r = threadIdx.x; r < D; r += blockDim.x, ++k) st(b + (base | offs[r]), out[k]);
    }
}

// :143-167 for dense blocks of k = 4..6 targets on FP64 MFMA.
//
// The op is a complex GEMM Y = U X: U is 2^k x 2^k and the columns of X are
// the groups (one column = the 2^k amplitudes of one setting of the other
// qubits, controls at 1). It runs as four real GEMMs, Yr = Ur Xr - Ui Xi and
// Yi = Ui Xr + Ur Xi, on v_mfma_f64_16x16x4 tiles: 16 output rows x 16
// columns, 4 rows of X per step. A wave owns S slabs of 16 columns; lane l
// loads amplitude (4b + l/16) of column l%16 as one 16-byte (re, im) pair,
// which is the B operand of the Xr and the Xi product of row block b. U sits
// in LDS in fragment order: the A operand of output tile t and row block b
// for lane l is U[16t + l%16][4b + l/16], so each lane's 16-byte read is
// contiguous with its neighbours'. A tile's accumulators hold rows
// l/16 + 4v, v = 0..3, of column l%16.
struct DenseOffsets {
    uint64_t blk[16];  // target offset of local index 4b, b < 2^k / 4
    uint64_t lo[4];    // target offset of local index 0..3
};

typedef double v4d __attribute__((ext_vector_type(4)));

__device__ __forceinline__ v4d mfma4(double a, double b, v4d c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

template <int K, int S>
__global__ void __launch_bounds__(kBlock) k_cu_mfma(amp_t* __restrict__ s, uint32_t nq, Layout L,
                                                    DenseOffsets o, const amp_t* __restrict__ mat,
                                                    uint64_t iters) {
    constexpr int D = 1 << K, NB = D / 4, NT = D / 16;
    __shared__ amp_t frag[D * D];
    for (int e = threadIdx.x; e < D * D; e += kBlock) {
        const int lane = e & 63, tb = e >> 6, t = tb / NB, b = tb % NB;
        frag[e] = mat[(16 * t + (lane & 15)) * D + 4 * b + (lane >> 4)];
    }
    __syncthreads();
    amp_t* base_ptr = batch_base(s, nq);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint64_t ld_lane = o.lo[lane >> 4];
    for (uint64_t it = uint64_t(blockIdx.x) * (kBlock / 64) + wave; it < iters;
         it += uint64_t(gridDim.x) * (kBlock / 64)) {
        uint64_t col[S];
        amp_t x[S][NB];
#pragma unroll
        for (int q = 0; q < S; ++q) {
            col[q] = expand((it * S + q) * 16 + (lane & 15), L);
#pragma unroll
            for (int b = 0; b < NB; ++b) x[q][b] = ld(base_ptr + (col[q] | o.blk[b] | ld_lane));
        }
        v4d yr[S][NT], yi[S][NT];
        // U fits in registers up to k = 5 (the compiler hoists its LDS reads
        // out of the loop); at k = 6 it is re-read from LDS every iteration.
        int zero = 0;
        if (K == 6) asm volatile("" : "+v"(zero));
#pragma unroll
        for (int q = 0; q < S; ++q)
#pragma unroll
            for (int t = 0; t < NT; ++t) yr[q][t] = yi[q][t] = v4d{0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int t = 0; t < NT; ++t)
#pragma unroll
            for (int b = 0; b < NB; ++b) {
                const amp_t u = frag[(t * NB + b) * 64 + lane + zero];
#pragma unroll
                for (int q = 0; q < S; ++q) {
                    yr[q][t] = mfma4(u.x, x[q][b].x, yr[q][t]);
                    yr[q][t] = mfma4(-u.y, x[q][b].y, yr[q][t]);
                    yi[q][t] = mfma4(u.y, x[q][b].x, yi[q][t]);
                    yi[q][t] = mfma4(u.x, x[q][b].y, yi[q][t]);
                }
            }
#pragma unroll
        for (int q = 0; q < S; ++q)
#pragma unroll
            for (int t = 0; t < NT; ++t)
#pragma unroll
                for (int v = 0; v < 4; ++v)
                    st(base_ptr + (col[q] | o.blk[4 * t] | ld_lane | o.blk[v]),
                       make_double2(yr[q][t][v], yi[q][t][v]));
    }
}

// :169-189 -- |l> -> |perm[l]> on the control subspace, staged through LDS.
// A round stages max(D, kPermSlots) amplitudes: kPermSlots / D whole groups
// when D is small, so that every thread has work; consecutive threads walk l
// within a group (contiguous when the targets are the low qubits).
constexpr uint32_t kPermSlots = 2048;

__global__ void __launch_bounds__(kBlock) k_cperm(amp_t* __restrict__ s, uint32_t nq, Layout L,
                                                  const uint64_t* __restrict__ offs,
                                                  const uint64_t* __restrict__ perm) {
    // Dynamic LDS: the round's amplitudes, then (several groups per round)
    // the first index of each group, ~0 past the last group.
    extern __shared__ amp_t grp[];
    const uint32_t log_d = L.nt, D = 1u << log_d;
    const uint32_t log_g = log_d < 11 ? 11 - log_d : 0, per_round = D << log_g;
    uint64_t* gbase = reinterpret_cast<uint64_t*>(grp + per_round);
    amp_t* b = batch_base(s, nq);
    const uint64_t groups = uint64_t{1} << (nq - L.n_fixed);
    const uint64_t rounds = (groups + (uint64_t{1} << log_g) - 1) >> log_g;
    for (uint64_t r = blockIdx.x; r < rounds; r += gridDim.x) {
        const uint64_t g0 = r << log_g;
        uint64_t only = 0;  // one group per round: its first index
        if (log_g == 0) {
            only = expand(g0, L);
        } else {
            for (uint32_t gi = threadIdx.x; gi < (1u << log_g); gi += kBlock)
                gbase[gi] = g0 + gi < groups ? expand(g0 + gi, L) : ~uint64_t{0};
            __syncthreads();
        }
        for (uint32_t e = threadIdx.x; e < per_round; e += kBlock) {
            const uint64_t base = log_g == 0 ? only : gbase[e >> log_d];
            if (base != ~uint64_t{0}) grp[e] = ld(b + (base | offs[e & (D - 1)]));
        }
        __syncthreads();
        for (uint32_t e = threadIdx.x; e < per_round; e += kBlock) {
            const uint64_t base = log_g == 0 ? only : gbase[e >> log_d];
            if (base != ~uint64_t{0}) st(b + (base | offs[perm[e & (D - 1)]]), grp[e]);
        }
        __syncthreads();  // grp and gbase are reused by the next round
    }
}

// The same by cycles, for target and control sets that leave the lowest
// qubits free: one thread per group, so consecutive threads touch consecutive
// amplitudes. Each thread rotates every cycle of the permutation in place,
// c_0 -> c_1 -> ... (the amplitude at c_i moves to c_{i+1} = perm[c_i]),
// reading 16 cycle entries ahead of its writes; fixed points are not touched.
// cyc holds the cycles back to back, starts[k]..starts[k+1] the k-th.
__global__ void __launch_bounds__(kBlock) k_cperm_cycles(amp_t* __restrict__ s, uint32_t nq, Layout L,
                                                         const uint64_t* __restrict__ offs,
                                                         const uint32_t* __restrict__ cyc,
                                                         const uint32_t* __restrict__ starts, uint32_t n_cycles) {
    constexpr int kAhead = 16;
    amp_t* b = batch_base(s, nq);
    const uint64_t groups = uint64_t{1} << (nq - L.n_fixed);
    for (uint64_t g = uint64_t(blockIdx.x) * kBlock + threadIdx.x; g < groups; g += uint64_t(gridDim.x) * kBlock) {
        amp_t* base = b + expand(g, L);
        for (uint32_t k = 0; k < n_cycles; ++k) {
            const uint32_t first = starts[k], len = starts[k + 1] - first;
            amp_t carry = ld(base + offs[cyc[first + len - 1]]);  // moves to c_0
            for (uint32_t i = 0; i < len; i += kAhead) {
                amp_t x[kAhead];
#pragma unroll
                for (int j = 0; j < kAhead; ++j)
                    if (i + j < len) x[j] = ld(base + offs[cyc[first + i + j]]);
#pragma unroll
                for (int j = 0; j < kAhead; ++j)
                    if (i + j < len) {
                        st(base + offs[cyc[first + i + j]], carry);
                        carry = x[j];
                    }
            }
        }
    }
}

// :191-225 -- exp(-i theta/2 P): pairs (i, i^x) with i's lowest x bit clear
__global__ void __launch_bounds__(kBlock) k_pauli_pairs(amp_t* __restrict__ s, uint32_t nq,
                                                        uint64_t x, uint64_t zy, uint32_t lsb,
                                                        amp_t c, amp_t yphase) {
    amp_t* b = batch_base(s, nq);
    const uint64_t groups = uint64_t{1} << (nq - 1);
    LQCS_FOR_BLOCKS(blk, groups) {
        amp_t ai[kItems], aj[kItems];
        uint64_t ii[kItems];
#pragma unroll
        for (int k = 0; k < kItems; ++k) {
            const uint64_t g = group_index(blk, k);
            ii[k] = insert_zero_bit(g < groups ? g : 0, lsb);
            ai[k] = ld(b + ii[k]);
            aj[k] = ld(b + (ii[k] ^ x));
        }
#pragma unroll
        for (int k = 0; k < kItems; ++k) {
            if (group_index(blk, k) >= groups) continue;
            const uint64_t i = ii[k], j = i ^ x;
            const double si = (__popcll(i & zy) & 1) ? -1.0 : 1.0;
            const double sj = (__popcll(j & zy) & 1) ? -1.0 : 1.0;
            st(b + i, cmac2(c, ai[k], cscale(yphase, sj), aj[k]));
            st(b + j, cmac2(cscale(yphase, si), ai[k], c, aj[k]));
        }
    }
}

__global__ void __launch_bounds__(kBlock) k_pauli_diag(amp_t* __restrict__ s, uint32_t nq,
                                                       uint64_t zy, amp_t c, amp_t yphase) {
    amp_t* b = batch_base(s, nq);
    const uint64_t n = uint64_t{1} << nq;
    const amp_t dp = cadd(c, yphase), dm = cadd(c, cscale(yphase, -1.0));
    LQCS_FOR_BLOCKS(blk, n) {
        amp_t a[kItems];
#pragma unroll
        for (int k = 0; k < kItems; ++k) {
            const uint64_t i = group_index(blk, k);
            a[k] = ld(b + (i < n ? i : 0));
        }
#pragma unroll
        for (int k = 0; k < kItems; ++k) {
            const uint64_t i = group_index(blk, k);
            if (i >= n) continue;
            st(b + i, cmul(a[k], (__popcll(i & zy) & 1) ? dm : dp));
        }
    }
}

__global__ void k_fill_basis(amp_t* __restrict__ s, uint32_t nq, uint64_t index) {
    amp_t* b = batch_base(s, nq);
    const uint64_t n = uint64_t{1} << nq;
    for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += uint64_t(gridDim.x) * blockDim.x)
        b[i] = make_double2(i == index ? 1.0 : 0.0, 0.0);
}

Layout make_layout(const uint32_t* ctl, uint32_t nc, const uint32_t* tgt, uint32_t nt) {
    Layout L{};
    if (nc + nt > LQCS_LAYOUT_MAX) throw Error(3, "operation touches too many qubits");
    L.n_fixed = 0;
    for (uint32_t i = 0; i < nc; ++i) L.fixed[L.n_fixed++] = ctl[i];
    for (uint32_t i = 0; i < nt; ++i) L.fixed[L.n_fixed++] = tgt[i];
    for (uint32_t i = 1; i < L.n_fixed; ++i)
        for (uint32_t j = i; j > 0 && L.fixed[j - 1] > L.fixed[j]; --j) {
            const uint32_t tmp = L.fixed[j];
            L.fixed[j] = L.fixed[j - 1];
            L.fixed[j - 1] = tmp;
        }
    L.c_mask = 0;
    for (uint32_t i = 0; i < nc; ++i) L.c_mask |= uint64_t{1} << ctl[i];
    L.nt = nt;
    return L;
}

std::vector<uint64_t> target_offsets(const uint32_t* tgt, uint32_t nt) {
    std::vector<uint64_t> off(size_t{1} << nt, 0);
    for (uint64_t l = 0; l < off.size(); ++l)
        for (uint32_t j = 0; j < nt; ++j)
            if ((l >> j) & 1) off[l] |= uint64_t{1} << tgt[j];
    return off;
}

}  // namespace

uint32_t g_dense_mfma_min_targets = 4;

void launch_single(amp_t* s, uint32_t nq, uint32_t t, const Mat2& m, uint32_t batch) {
    k_single<<<dim3(grid_capped(uint64_t{1} << (nq - 1), kItems), batch), kBlock, 0, stream()>>>(s, nq, t, m);
    launched();
}

void launch_diag_single(amp_t* s, uint32_t nq, uint32_t t, amp_t d0, amp_t d1, uint32_t batch) {
    k_diag_single<<<dim3(grid_capped(uint64_t{1} << nq, kItems), batch), kBlock, 0, stream()>>>(s, nq, t, d0, d1);
    launched();
}

void launch_controlled_single(amp_t* s, uint32_t nq, uint32_t c, uint32_t t, const Mat2& m,
                              uint32_t batch) {
    k_controlled_single<<<dim3(grid_capped(uint64_t{1} << (nq - 2), kItems), batch), kBlock, 0, stream()>>>(
        s, nq, c, t, m);
    launched();
}

void launch_swap(amp_t* s, uint32_t nq, uint32_t a, uint32_t b, uint32_t batch) {
    k_swap<<<dim3(grid_capped(uint64_t{1} << (nq - 2), kItems), batch), kBlock, 0, stream()>>>(s, nq, a, b);
    launched();
}

void launch_two_qubit(amp_t* s, uint32_t nq, uint32_t q0, uint32_t q1, const Mat4& m,
                      uint32_t batch) {
    k_two_qubit<<<dim3(grid_capped(uint64_t{1} << (nq - 2)), batch), kBlock, 0, stream()>>>(s, nq, q0, q1, m);
    launched();
}

void launch_controlled_unitary(amp_t* s, uint32_t nq, const uint32_t* ctl, uint32_t nc,
                               const uint32_t* tgt, uint32_t nt, const amp_t* m_host,
                               uint32_t batch) {
    const Layout L = make_layout(ctl, nc, tgt, nt);
    const std::vector<uint64_t> off = target_offsets(tgt, nt);
    const uint64_t D = uint64_t{1} << nt;
    const uint64_t groups = uint64_t{1} << (nq - L.n_fixed);
    const amp_t* d_mat = stage(m_host, D * D);
    if (nt >= g_dense_mfma_min_targets && nt >= 4 && nt <= 6) {
        const uint32_t S = nt == 4 ? 4 : nt == 5 ? 2 : 1;
        const uint64_t slabs = groups / 16;
        if (groups >= 16 * S) {
            DenseOffsets o{};
            for (uint64_t b = 0; b < D / 4; ++b) o.blk[b] = off[4 * b];
            for (int v = 0; v < 4; ++v) o.lo[v] = off[v];
            const uint64_t iters = slabs / S;
            const uint64_t want = (iters + kBlock / 64 - 1) / (kBlock / 64);
            const unsigned grid = static_cast<unsigned>(std::min<uint64_t>(want, uint64_t(num_cus()) * 8));
            const dim3 g(grid, batch);
            switch (nt) {
                case 4: k_cu_mfma<4, 4><<<g, kBlock, 0, stream()>>>(s, nq, L, o, d_mat, iters); launched(); return;
                case 5: k_cu_mfma<5, 2><<<g, kBlock, 0, stream()>>>(s, nq, L, o, d_mat, iters); launched(); return;
                default: k_cu_mfma<6, 1><<<g, kBlock, 0, stream()>>>(s, nq, L, o, d_mat, iters); launched(); return;
            }
        }
    }
    const uint64_t* d_off = stage(off.data(), off.size());
    const dim3 g(grid_capped(groups), batch);
    switch (nt) {
        case 1: k_cu_small<1><<<g, kBlock, 0, stream()>>>(s, nq, L, d_off, d_mat); launched(); return;
        case 2: k_cu_small<2><<<g, kBlock, 0, stream()>>>(s, nq, L, d_off, d_mat); launched(); return;
        case 3: k_cu_small<3><<<g, kBlock, 0, stream()>>>(s, nq, L, d_off, d_mat); launched(); return;
        case 4: k_cu_small<4><<<g, kBlock, 0, stream()>>>(s, nq, L, d_off, d_mat); launched(); return;
        default: break;
    }
    if (D > uint64_t(kBlock) * LQCS_CU_LARGE_ROWS)
        throw Error(3, "controlled unitary wider than the engine supports");
    k_cu_large<<<dim3(capped(groups), batch), kBlock, D * sizeof(amp_t), stream()>>>(
        s, nq, L, d_off, d_mat);
    launched();
}

void launch_controlled_permutation(amp_t* s, uint32_t nq, const uint32_t* ctl, uint32_t nc,
                                   const uint32_t* tgt, uint32_t nt, const uint64_t* perm_host,
                                   uint32_t batch) {
    const Layout L = make_layout(ctl, nc, tgt, nt);
    const std::vector<uint64_t> off = target_offsets(tgt, nt);
    const uint64_t D = uint64_t{1} << nt;
    if (D * sizeof(amp_t) > 64 * 1024) throw Error(3, "permutation wider than the engine supports");
    const uint64_t groups = uint64_t{1} << (nq - L.n_fixed);
    const uint64_t* d_off = stage(off.data(), off.size());
    // Lowest qubits free and enough groups for one thread each: rotate cycles.
    if (L.n_fixed > 0 && L.fixed[0] >= 4 && groups >= uint64_t(kBlock)) {
        std::vector<uint32_t> cyc, starts{0};
        std::vector<uint8_t> seen(D, 0);
        for (uint64_t l = 0; l < D; ++l) {
            if (seen[l] || perm_host[l] == l) continue;
            for (uint64_t c = l; !seen[c]; c = perm_host[c]) {
                seen[c] = 1;
                cyc.push_back(static_cast<uint32_t>(c));
            }
            starts.push_back(static_cast<uint32_t>(cyc.size()));
        }
        const uint32_t n_cycles = static_cast<uint32_t>(starts.size() - 1);
        if (n_cycles == 0) return;  // the identity
        const uint32_t* d_cyc = stage(cyc.data(), cyc.size());
        const uint32_t* d_starts = stage(starts.data(), starts.size());
        k_cperm_cycles<<<dim3(grid_capped(groups), batch), kBlock, 0, stream()>>>(s, nq, L, d_off, d_cyc, d_starts,
                                                                                 n_cycles);
        launched();
        return;
    }
    const uint64_t* d_perm = stage(perm_host, D);
    const uint64_t per_round = std::max<uint64_t>(D, kPermSlots);
    const uint64_t groups_per_round = per_round / D;
    const uint64_t rounds = (groups + groups_per_round - 1) / groups_per_round;
    const size_t lds = per_round * sizeof(amp_t) + (groups_per_round > 1 ? groups_per_round * sizeof(uint64_t) : 0);
    k_cperm<<<dim3(capped(rounds), batch), kBlock, lds, stream()>>>(s, nq, L, d_off, d_perm);
    launched();
}

void launch_pauli_rotation(amp_t* s, uint32_t nq, uint64_t x, uint64_t zy, uint32_t n_y,
                           double theta, uint32_t batch) {
    const double h = theta / 2;
    const amp_t c = make_double2(std::cos(h), 0.0);
    const amp_t mis = make_double2(0.0, -std::sin(h));
    static const amp_t kIPow[4] = {{1, 0}, {0, 1}, {-1, 0}, {0, -1}};
    const amp_t ip = kIPow[n_y % 4];
    const amp_t yphase = make_double2(mis.x * ip.x - mis.y * ip.y, mis.x * ip.y + mis.y * ip.x);
    if (x == 0) {
        k_pauli_diag<<<dim3(grid_capped(uint64_t{1} << nq, kItems), batch), kBlock, 0, stream()>>>(s, nq, zy, c,
                                                                                                  yphase);
        launched();
        return;
    }
    const uint32_t lsb = static_cast<uint32_t>(__builtin_ctzll(x));
    k_pauli_pairs<<<dim3(grid_capped(uint64_t{1} << (nq - 1), kItems), batch), kBlock, 0, stream()>>>(
        s, nq, x, zy, lsb, c, yphase);
    launched();
}

void launch_fill_basis(amp_t* s, uint32_t nq, uint64_t index, uint32_t batch) {
    const uint64_t n = uint64_t{1} << nq;
    const unsigned grid = static_cast<unsigned>(std::min<uint64_t>(grid_for(n), 65536));
    k_fill_basis<<<dim3(grid, batch), kBlock, 0, stream()>>>(s, nq, index);
    launched();
}

}  // namespace lqcs_dev
