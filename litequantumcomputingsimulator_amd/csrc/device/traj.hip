// Batched noisy trajectories (the per-shot path of run_per_shot,
// src/backend/statevector_simulator.cpp:122-190).
//
// B shots evolve side by side as B registers in one allocation (blockIdx.y =
// shot). Draw order is the reference's: per unitary instruction, channel-major
// then qubit order; one draw per Measure/Reset; shot s owns draws
// [s*D, (s+1)*D).
//
// The reference picks a Kraus branch per (gate, channel, qubit) with
// p_i = ||K_i psi||^2 of the normalised state and applies K_i / sqrt(p_i).
// Most of those choices are known before the state is:
//   * a channel whose K_i^dag K_i are all multiples c_i of the identity
//     (depolarizing, bit/phase flip) has p_i = c_i whatever psi is, so the
//     host decides the branch from the draw alone;
//   * for any other channel (amplitude damping), p_0 >= lambda_min(K_0^dag K_0),
//     so a draw below that bound picks branch 0 whatever psi is.
// Every decision compares ratios, so the registers are kept unnormalised
// between reductions: the default branch is applied as K_0 without 1/sqrt(p).
// A shot leaves this default path only at an "event": another branch of a
// state-independent channel (applied as the unitary K_i / sqrt(c_i)), or a
// draw the bound does not settle (one 2x2-statistics read of that shot's
// register, then the reference's rule on the normalised probabilities).
//
// Between state-dependent events the whole batch shares one op sequence (the
// circuit's gates, the default branches, and the other branches of
// state-independent channels as ops restricted to one batch register), which
// runs as tiled passes over the batch: a trajectory costs a few HBM passes
// per tile set instead of two per noise site. Trailing measurements are
// sampled without collapsing. Branch choices equal the reference's except for
// draws within
// rounding of a cumulative boundary.
#include "common.hpp"
#include "engine.hpp"
#include "reduce.hpp"
#include "tile_apply.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

namespace lqcs_dev {
namespace {

constexpr int kMaxKraus = 8;
struct KrausSet {
    amp_t k[kMaxKraus][4];
    uint32_t n;
};

// Event members at one noise site: decide among all branches from the 2x2
// statistics of the member's (unnormalised) register, then apply K/sqrt(p),
// which also renormalises it. Block row j handles member members[j].
__global__ void __launch_bounds__(kBlock) k_kraus_members(amp_t* __restrict__ s, uint32_t nq, uint32_t q,
                                                          const double* __restrict__ stats, KrausSet ks,
                                                          const double* __restrict__ draws, uint64_t draw_stride,
                                                          uint64_t draw_offset, const uint32_t* __restrict__ members) {
    const uint32_t r = members[blockIdx.y];
    const double* st = stats + 4 * blockIdx.y;
    const double norm2 = st[0] + st[1];
    const double u = draws[uint64_t(r) * draw_stride + draw_offset];
    // sample_kraus_branch on the normalised state: skip p <= 0, first with
    // u < cum, falling back to the last non-zero branch.
    int chosen = -1;
    double chosen_p = 0.0, cum = 0.0;
    for (uint32_t i = 0; i < ks.n; ++i) {
        const amp_t* K = ks.k[i];
        const double c00 = cnorm(K[0]) + cnorm(K[2]);
        const double c11 = cnorm(K[1]) + cnorm(K[3]);
        const double xr = K[0].x * K[1].x + K[0].y * K[1].y + K[2].x * K[3].x + K[2].y * K[3].y;
        const double xi = K[0].x * K[1].y - K[0].y * K[1].x + K[2].x * K[3].y - K[2].y * K[3].x;
        const double p = (c00 * st[0] + c11 * st[1] + 2.0 * (xr * st[2] - xi * st[3])) / norm2;
        if (p <= 0.0) continue;
        chosen = static_cast<int>(i);
        chosen_p = p;
        cum += p;
        if (u < cum) break;
    }
    if (chosen < 0) return;
    const double inv = 1.0 / sqrt(chosen_p * norm2);
    const amp_t m0 = cscale(ks.k[chosen][0], inv), m1 = cscale(ks.k[chosen][1], inv);
    const amp_t m2 = cscale(ks.k[chosen][2], inv), m3 = cscale(ks.k[chosen][3], inv);
    amp_t* b = s + (uint64_t(r) << nq);
    const uint64_t half = uint64_t{1} << (nq - 1);
    for (uint64_t g = uint64_t(blockIdx.x) * kBlock + threadIdx.x; g < half; g += uint64_t(gridDim.x) * kBlock) {
        const uint64_t i0 = insert_zero_bit(g, q), i1 = i0 | (uint64_t{1} << q);
        const amp_t a0 = b[i0], a1 = b[i1];
        b[i0] = cmac2(m0, a0, m1, a1);
        b[i1] = cmac2(m2, a0, m3, a1);
    }
}

// A 2x2 on qubit q of every batch member except those flagged in skip.
__global__ void __launch_bounds__(kBlock) k_single_skip(amp_t* __restrict__ s, uint32_t nq, uint32_t q, Mat2 m,
                                                        const uint8_t* __restrict__ skip) {
    if (skip[blockIdx.y]) return;
    amp_t* b = s + (uint64_t(blockIdx.y) << nq);
    const uint64_t half = uint64_t{1} << (nq - 1);
    for (uint64_t g = uint64_t(blockIdx.x) * kBlock + threadIdx.x; g < half; g += uint64_t(gridDim.x) * kBlock) {
        const uint64_t i0 = insert_zero_bit(g, q), i1 = i0 | (uint64_t{1} << q);
        const amp_t a0 = b[i0], a1 = b[i1];
        b[i0] = cmac2(m.m[0], a0, m.m[1], a1);
        b[i1] = cmac2(m.m[2], a0, m.m[3], a1);
    }
}

// Measure / reset of qubit q on every member (collapse_qubit / reset_qubit,
// gate_kernels.cpp:227-260) from unnormalised statistics: p1 = S11/(S00+S11),
// outcome u < p1, the kept branch scaled by 1/sqrt(S_kept) (renormalised).
__global__ void __launch_bounds__(kBlock) k_collapse_stats(amp_t* __restrict__ s, uint32_t nq, uint32_t q,
                                                           const double* __restrict__ stats,
                                                           const double* __restrict__ draws, uint64_t draw_stride,
                                                           uint64_t draw_offset, int reset, uint8_t* __restrict__ creg,
                                                           uint32_t creg_stride, int32_t clbit) {
    const uint32_t r = blockIdx.y;
    amp_t* b = s + (uint64_t(r) << nq);
    const double s00 = stats[4 * r], s11 = stats[4 * r + 1];
    const double p1 = s11 / (s00 + s11);
    const double u = draws[uint64_t(r) * draw_stride + draw_offset];
    const bool one = u < p1;
    const double scale = 1.0 / sqrt(one ? s11 : s00);
    const uint64_t half = uint64_t{1} << (nq - 1);
    const uint64_t g = uint64_t(blockIdx.x) * kBlock + threadIdx.x;
    if (g == 0 && clbit >= 0) creg[uint64_t(r) * creg_stride + clbit] = one ? 1 : 0;
    for (uint64_t k = g; k < half; k += uint64_t(gridDim.x) * kBlock) {
        const uint64_t i0 = insert_zero_bit(k, q), i1 = i0 | (uint64_t{1} << q);
        const amp_t a0 = b[i0], a1 = b[i1];
        const amp_t z = make_double2(0.0, 0.0);
        if (one) {
            if (reset) {
                b[i0] = cscale(a1, scale);
                b[i1] = z;
            } else {
                b[i0] = z;
                b[i1] = cscale(a1, scale);
            }
        } else {
            b[i0] = cscale(a0, scale);
            b[i1] = z;
        }
    }
}

// Terminal measurements (every step after them is a measurement too): the
// state is not needed afterwards, so instead of collapsing it, the sums for
// each outcome run over the amplitudes that agree with the member's outcomes
// so far (fixed qubits F, values V[member]), split by the measured qubit q.
// p1 = S1 / (S0 + S1) is the collapsed state's probability up to rounding:
// collapse and renormalisation only zero the other amplitudes and scale the
// kept ones by one factor. When F is the low qubits 0..k-1 (measure_all), the
// amplitudes read are (g << k) | V.
__device__ __forceinline__ double tsum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

__global__ void __launch_bounds__(kBlock) k_term_partial(const amp_t* __restrict__ s, uint32_t nq, uint64_t fixed,
                                                         const uint64_t* __restrict__ vals, uint32_t q,
                                                         double* __restrict__ part) {
    const uint32_t r = blockIdx.y;
    const amp_t* b = s + (uint64_t(r) << nq);
    const uint64_t v = vals[r], full = (uint64_t{1} << nq) - 1, free_bits = full & ~fixed;
    const uint32_t k = static_cast<uint32_t>(__popcll(fixed));
    const bool prefix = fixed == ((uint64_t{1} << k) - 1);
    const uint64_t n = uint64_t{1} << (nq - k), qbit = uint64_t{1} << q;
    double acc0 = 0.0, acc1 = 0.0;
    for (uint64_t g = uint64_t(blockIdx.x) * kBlock + threadIdx.x; g < n; g += uint64_t(gridDim.x) * kBlock) {
        uint64_t i;
        if (prefix) {
            i = (g << k) | v;
        } else {
            i = v;
            uint64_t m = free_bits;
            for (uint64_t bit = 1; m; bit <<= 1) {
                const uint64_t low = m & (~m + 1);
                if (g & bit) i |= low;
                m ^= low;
            }
        }
        const double w = cnorm(b[i]);
        if (i & qbit) acc1 += w; else acc0 += w;
    }
    __shared__ double red[2][kBlock / 64];
    acc0 = tsum(acc0);
    acc1 = tsum(acc1);
    if ((threadIdx.x & 63) == 0) {
        red[0][threadIdx.x >> 6] = acc0;
        red[1][threadIdx.x >> 6] = acc1;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double t0 = 0.0, t1 = 0.0;
        for (int w = 0; w < kBlock / 64; ++w) {
            t0 += red[0][w];
            t1 += red[1][w];
        }
        part[2 * (uint64_t(r) * gridDim.x + blockIdx.x)] = t0;
        part[2 * (uint64_t(r) * gridDim.x + blockIdx.x) + 1] = t1;
    }
}

// Sums each member's partials (fixed order), decides the outcome with the
// member's draw (u < p1) and records it.
__global__ void __launch_bounds__(kBlock) k_term_decide(const double* __restrict__ part, uint32_t blocks,
                                                        const double* __restrict__ draws, uint64_t draw_stride,
                                                        uint64_t draw_offset, uint32_t q, uint64_t* __restrict__ vals,
                                                        uint8_t* __restrict__ creg, uint32_t creg_stride,
                                                        uint32_t clbit) {
    const uint32_t r = blockIdx.x;
    double a0 = 0.0, a1 = 0.0;
    for (uint32_t j = threadIdx.x; j < blocks; j += kBlock) {
        a0 += part[2 * (uint64_t(r) * blocks + j)];
        a1 += part[2 * (uint64_t(r) * blocks + j) + 1];
    }
    __shared__ double red[2][kBlock / 64];
    a0 = tsum(a0);
    a1 = tsum(a1);
    if ((threadIdx.x & 63) == 0) {
        red[0][threadIdx.x >> 6] = a0;
        red[1][threadIdx.x >> 6] = a1;
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    double s0 = 0.0, s1 = 0.0;
    for (int w = 0; w < kBlock / 64; ++w) {
        s0 += red[0][w];
        s1 += red[1][w];
    }
    const double p1 = s1 / (s0 + s1);
    const bool one = draws[uint64_t(r) * draw_stride + draw_offset] < p1;
    creg[uint64_t(r) * creg_stride + clbit] = one ? 1 : 0;
    vals[r] = one ? (vals[r] | (uint64_t{1} << q)) : (vals[r] & ~(uint64_t{1} << q));
}

unsigned pass_blocks(uint32_t nq) {
    const uint64_t want = ((uint64_t{1} << (nq - 1)) + kBlock * 4 - 1) / (kBlock * 4);
    return static_cast<unsigned>(std::max<uint64_t>(1, std::min<uint64_t>(want, 1u << 16)));
}

// How a channel's branch is chosen.
struct ChannelPlan {
    bool mixture = false;            // every K_i^dag K_i = c_i * I
    std::vector<double> c;           // mixture: the c_i
    int def = -1;                    // branch taken without an event
    double bound = 0.0;              // general: draws below this pick `def`
    bool def_identity = false;       // default branch is the identity (nothing to apply)
    Mat2 def_op{};                   // default branch as applied (unitary for mixtures, K_0 otherwise)
    std::vector<Mat2> unit;          // mixture: K_i / sqrt(c_i)
};

ChannelPlan plan_channel(const KrausSet& ks) {
    ChannelPlan cp;
    cp.mixture = true;
    cp.c.resize(ks.n);
    double scale = 0.0;
    for (uint32_t i = 0; i < ks.n; ++i) {
        const amp_t* K = ks.k[i];
        const double g00 = cnorm(K[0]) + cnorm(K[2]), g11 = cnorm(K[1]) + cnorm(K[3]);
        const double gr = K[0].x * K[1].x + K[0].y * K[1].y + K[2].x * K[3].x + K[2].y * K[3].y;
        const double gi = K[0].x * K[1].y - K[0].y * K[1].x + K[2].x * K[3].y - K[2].y * K[3].x;
        scale = std::max(scale, std::max(g00, g11));
        if (std::abs(g00 - g11) > 1e-14 * std::max(1.0, g00) || std::hypot(gr, gi) > 1e-14 * std::max(1.0, g00))
            cp.mixture = false;
        cp.c[i] = 0.5 * (g00 + g11);
    }
    if (cp.mixture) {
        for (uint32_t i = 0; i < ks.n; ++i) {
            Mat2 u{};
            const double inv = cp.c[i] > 0.0 ? 1.0 / std::sqrt(cp.c[i]) : 0.0;
            for (int e = 0; e < 4; ++e) u.m[e] = cscale(ks.k[i][e], inv);
            cp.unit.push_back(u);
            if (cp.def < 0 && cp.c[i] > 0.0) cp.def = static_cast<int>(i);
        }
        if (cp.def >= 0) {
            cp.def_op = cp.unit[cp.def];
            const amp_t* u = cp.def_op.m;
            cp.def_identity = std::abs(u[0].x - 1.0) < 1e-15 && std::abs(u[0].y) < 1e-15 &&
                              std::abs(u[3].x - 1.0) < 1e-15 && std::abs(u[3].y) < 1e-15 && cnorm(u[1]) == 0.0 &&
                              cnorm(u[2]) == 0.0;
        }
        return cp;
    }
    // General channel: branch 0 (the first non-zero operator) is certain for
    // draws below lambda_min(K_0^dag K_0), with a margin for rounding.
    for (uint32_t i = 0; i < ks.n && cp.def < 0; ++i) {
        const amp_t* K = ks.k[i];
        if (cnorm(K[0]) + cnorm(K[1]) + cnorm(K[2]) + cnorm(K[3]) > 0.0) cp.def = static_cast<int>(i);
    }
    if (cp.def < 0) return cp;
    const amp_t* K = ks.k[cp.def];
    const double a = cnorm(K[0]) + cnorm(K[2]), d = cnorm(K[1]) + cnorm(K[3]);
    const double br = K[0].x * K[1].x + K[0].y * K[1].y + K[2].x * K[3].x + K[2].y * K[3].y;
    const double bi = K[0].x * K[1].y - K[0].y * K[1].x + K[2].x * K[3].y - K[2].y * K[3].x;
    const double lmin = 0.5 * (a + d) - std::sqrt(0.25 * (a - d) * (a - d) + br * br + bi * bi);
    cp.bound = std::max(0.0, lmin * (1.0 - 1e-12) - 1e-15);
    for (int e = 0; e < 4; ++e) cp.def_op.m[e] = K[e];
    return cp;
}

// Branch of a mixture channel for draw u (host restatement of the kernel's rule).
int mixture_branch(const ChannelPlan& cp, double u) {
    int chosen = -1;
    double cum = 0.0;
    for (size_t i = 0; i < cp.c.size(); ++i) {
        if (cp.c[i] <= 0.0) continue;
        chosen = static_cast<int>(i);
        cum += cp.c[i];
        if (u < cum) break;
    }
    return chosen;
}

// a * b^dagger of two 2x2 matrices (row-major).
Mat2 mul_adj(const Mat2& a, const Mat2& b) {
    Mat2 r{};
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j) {
            amp_t acc = make_double2(0.0, 0.0);
            for (int k = 0; k < 2; ++k) {
                const amp_t bc = make_double2(b.m[2 * j + k].x, -b.m[2 * j + k].y);  // conj(b[j][k])
                acc = cadd(acc, cmul(a.m[2 * i + k], bc));
            }
            r.m[2 * i + j] = acc;
        }
    return r;
}

DecodedOp single_op(uint32_t q, const Mat2& m) {
    DecodedOp d;
    d.targets = {q};
    d.touched = uint64_t{1} << q;
    const bool diag = cnorm(m.m[1]) == 0.0 && cnorm(m.m[2]) == 0.0;
    d.diagonal = diag;
    if (diag) {
        d.kind = LQCS_OP_DIAGONAL_SINGLE;
        d.m = {m.m[0], m.m[3]};
    } else {
        d.kind = LQCS_OP_SINGLE;
        d.m = {m.m[0], m.m[1], m.m[2], m.m[3]};
    }
    return d;
}

// Greedy single-qubit fusion of a shared run (the rule of the reference's
// gate_fusion.cpp:32-68, applied to the run's ops, whose noise branches are
// not unitary): uncontrolled 2x2 ops on a qubit multiply into one pending
// matrix, emitted before the next op that touches the qubit (or at the end).
// Ops restricted to one batch register touch their qubits like any other.
// The default branch of amplitude damping after every gate then folds into
// the gates around it instead of costing an op each.
void fuse_run(std::vector<DecodedOp>& ops) {
    std::vector<DecodedOp> out;
    out.reserve(ops.size());
    std::vector<Mat2> acc(64);
    uint64_t have = 0;
    auto as_mat = [](const DecodedOp& d) {
        Mat2 m{};
        if (d.kind == LQCS_OP_DIAGONAL_SINGLE) {
            m.m[0] = d.m[0];
            m.m[3] = d.m[1];
        } else {
            for (int e = 0; e < 4; ++e) m.m[e] = d.m[e];
        }
        return m;
    };
    auto emit = [&](uint32_t q) {
        out.push_back(single_op(q, acc[q]));
        have &= ~(uint64_t{1} << q);
    };
    for (DecodedOp& d : ops) {
        const bool one = (d.kind == LQCS_OP_SINGLE || d.kind == LQCS_OP_DIAGONAL_SINGLE) && d.controls.empty() &&
                         d.member < 0 && d.targets.size() == 1;
        if (one) {
            const uint32_t q = d.targets[0];
            const Mat2 m = as_mat(d);
            if (have & (uint64_t{1} << q)) {
                Mat2 r{};
                for (int i = 0; i < 2; ++i)
                    for (int j = 0; j < 2; ++j)
                        r.m[2 * i + j] = cadd(cmul(m.m[2 * i], acc[q].m[j]), cmul(m.m[2 * i + 1], acc[q].m[2 + j]));
                acc[q] = r;
            } else {
                acc[q] = m;
                have |= uint64_t{1} << q;
            }
            continue;
        }
        for (uint64_t t = d.touched & have; t; t &= t - 1) emit(static_cast<uint32_t>(__builtin_ctzll(t)));
        out.push_back(std::move(d));
    }
    for (uint64_t t = have; t; t &= t - 1) emit(static_cast<uint32_t>(__builtin_ctzll(t)));
    ops.swap(out);
}

// ---- persistent trajectories -------------------------------------------------
// A register of up to 12 qubits fits one workgroup's LDS (64 KiB), so a whole
// trajectory runs inside one workgroup: one launch covers every shot (one
// workgroup each), the register never leaves LDS, every branch decision is
// taken in the kernel (the draw against a mixture's weights or a general
// channel's default bound, the Kraus rule on the register's 2x2 statistics
// when the bound does not settle it, measurement and reset), and the only
// HBM traffic is the shot's draws and outcomes. Registers narrower than 6
// qubits run in a 6-qubit tile whose extra qubits stay |0>.
constexpr uint32_t kPersistMaxQubits = 12;

enum PStepKind : uint32_t { PS_GATE = 0, PS_NOISE = 1, PS_MEASURE = 2, PS_RESET = 3 };

struct PStep {
    uint32_t kind;
    uint32_t q;     // noise, measure and reset: the qubit
    uint32_t arg;   // gate: index in the gate table; noise: channel; measure: clbit
    uint32_t pad;
    uint64_t draw;  // noise, measure and reset: draw offset within the shot
};

struct ChannelDev {
    amp_t k[kMaxKraus][4];     // Kraus operators (row-major 2x2)
    amp_t unit[kMaxKraus][4];  // mixture: K_i / sqrt(c_i)
    double c[kMaxKraus];       // mixture: K_i^dag K_i = c_i I
    double bound;              // general: draws below it take branch `def`
    uint32_t n;
    uint32_t mixture;
    int32_t def;               // default branch, -1 when every operator is zero
    uint32_t def_identity;     // mixture whose default branch is the identity
};

// m (row-major 2x2) on qubit q of the LDS register.
template <int L, int T>
__device__ __forceinline__ void lds_2x2(amp_t* tile, uint32_t q, amp_t m0, amp_t m1, amp_t m2, amp_t m3) {
    constexpr int N = 1 << L;
#pragma unroll
    for (int k = 0; k < kIters<N / 2, T>; ++k) {
        const uint32_t p = threadIdx.x + k * T;
        if (N / 2 < T && p >= N / 2) break;
        const uint32_t i0 = static_cast<uint32_t>(insert_zero_bit(p, q)), i1 = i0 | (1u << q);
        const amp_t a0 = tile[i0], a1 = tile[i1];
        tile[i0] = cmac2(m0, a0, m1, a1);
        tile[i1] = cmac2(m2, a0, m3, a1);
    }
}

// 2x2 statistics of qubit q over the block (S00, S11, Re and Im of
// sum conj(a0) a1), summed in a fixed order; every thread receives them.
template <int L, int T>
__device__ __forceinline__ void lds_stats(const amp_t* tile, uint32_t q, double* red, double (&st)[4]) {
    constexpr int N = 1 << L;
    double a[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int k = 0; k < kIters<N / 2, T>; ++k) {
        const uint32_t p = threadIdx.x + k * T;
        if (N / 2 < T && p >= N / 2) break;
        const uint32_t i0 = static_cast<uint32_t>(insert_zero_bit(p, q)), i1 = i0 | (1u << q);
        const amp_t a0 = tile[i0], a1 = tile[i1];
        a[0] += cnorm(a0);
        a[1] += cnorm(a1);
        a[2] += a0.x * a1.x + a0.y * a1.y;
        a[3] += a0.x * a1.y - a0.y * a1.x;
    }
#pragma unroll
    for (int c = 0; c < 4; ++c) a[c] = tsum(a[c]);
    if ((threadIdx.x & 63) == 0) {
#pragma unroll
        for (int c = 0; c < 4; ++c) red[(threadIdx.x >> 6) * 4 + c] = a[c];
    }
    __syncthreads();
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        st[c] = 0.0;
        for (int w = 0; w < T / 64; ++w) st[c] += red[w * 4 + c];
    }
    __syncthreads();  // every read of red is done before it is written again
}

template <int L, int T>
__global__ void __launch_bounds__(T) k_traj_persist(const PStep* __restrict__ prog, uint32_t n_steps,
                                                    const TileOp* __restrict__ gates,
                                                    const ChannelDev* __restrict__ chans,
                                                    const double* __restrict__ draws, uint64_t D,
                                                    uint8_t* __restrict__ creg, uint32_t creg_w) {
    constexpr int N = 1 << L;
    constexpr int PER = N / T;
    __shared__ amp_t tile[N];
    __shared__ double red[4 * (T / 64)];
    const double* u_shot = draws + uint64_t(blockIdx.x) * D;
    uint8_t* c_shot = creg + uint64_t(blockIdx.x) * creg_w;
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        const uint32_t j = threadIdx.x + k * T;
        tile[j] = make_double2(j == 0 ? 1.0 : 0.0, 0.0);
    }
    for (uint32_t c = threadIdx.x; c < creg_w; c += T) c_shot[c] = 0;
    __syncthreads();
    for (uint32_t i = 0; i < n_steps; ++i) {
        const PStep st = prog[i];
        if (st.kind == PS_GATE) {
            tile_apply<L, T>(tile, gates[st.arg], 0);
            __syncthreads();
            continue;
        }
        const double u = u_shot[st.draw];
        if (st.kind == PS_NOISE) {
            const ChannelDev& ch = chans[st.arg];
            if (ch.def < 0) continue;
            if (ch.mixture) {
                // sample_kraus_branch with p_i = c_i whatever the state
                int k = -1;
                double cum = 0.0;
                for (uint32_t b = 0; b < ch.n; ++b) {
                    if (ch.c[b] <= 0.0) continue;
                    k = static_cast<int>(b);
                    cum += ch.c[b];
                    if (u < cum) break;
                }
                if (k == ch.def && ch.def_identity) continue;
                lds_2x2<L, T>(tile, st.q, ch.unit[k][0], ch.unit[k][1], ch.unit[k][2], ch.unit[k][3]);
                __syncthreads();
                continue;
            }
            if (u < ch.bound) {  // branch `def` whatever the state (kept unnormalised)
                const amp_t* K = ch.k[ch.def];
                lds_2x2<L, T>(tile, st.q, K[0], K[1], K[2], K[3]);
                __syncthreads();
                continue;
            }
            double sv[4];
            lds_stats<L, T>(tile, st.q, red, sv);
            const double norm2 = sv[0] + sv[1];
            int chosen = -1;
            double chosen_p = 0.0, cum = 0.0;
            for (uint32_t b = 0; b < ch.n; ++b) {
                const amp_t* K = ch.k[b];
                const double c00 = cnorm(K[0]) + cnorm(K[2]);
                const double c11 = cnorm(K[1]) + cnorm(K[3]);
                const double xr = K[0].x * K[1].x + K[0].y * K[1].y + K[2].x * K[3].x + K[2].y * K[3].y;
                const double xi = K[0].x * K[1].y - K[0].y * K[1].x + K[2].x * K[3].y - K[2].y * K[3].x;
                const double p = (c00 * sv[0] + c11 * sv[1] + 2.0 * (xr * sv[2] - xi * sv[3])) / norm2;
                if (p <= 0.0) continue;
                chosen = static_cast<int>(b);
                chosen_p = p;
                cum += p;
                if (u < cum) break;
            }
            if (chosen < 0) continue;
            const double inv = 1.0 / sqrt(chosen_p * norm2);
            const amp_t* K = ch.k[chosen];
            lds_2x2<L, T>(tile, st.q, cscale(K[0], inv), cscale(K[1], inv), cscale(K[2], inv), cscale(K[3], inv));
            __syncthreads();
            continue;
        }
        // measure / reset (collapse_qubit / reset_qubit): p1 = S11 / (S00 + S11)
        double sv[4];
        lds_stats<L, T>(tile, st.q, red, sv);
        const bool one = u < sv[1] / (sv[0] + sv[1]);
        const double scale = 1.0 / sqrt(one ? sv[1] : sv[0]);
        const bool reset = st.kind == PS_RESET;
        const amp_t z = make_double2(0.0, 0.0);
#pragma unroll
        for (int k = 0; k < kIters<N / 2, T>; ++k) {
            const uint32_t p = threadIdx.x + k * T;
            if (N / 2 < T && p >= N / 2) break;
            const uint32_t i0 = static_cast<uint32_t>(insert_zero_bit(p, st.q)), i1 = i0 | (1u << st.q);
            const amp_t a0 = tile[i0], a1 = tile[i1];
            if (one) {
                tile[i0] = reset ? cscale(a1, scale) : z;
                tile[i1] = reset ? z : cscale(a1, scale);
            } else {
                tile[i0] = cscale(a0, scale);
                tile[i1] = z;
            }
        }
        if (!reset && threadIdx.x == 0) c_shot[st.arg] = one ? 1 : 0;
        __syncthreads();
    }
}

// Whether every op of the program runs as a tile op of a whole-register tile.
bool persistent_fits(uint32_t nq, const std::vector<DecodedOp>& ops) {
    if (nq > kPersistMaxQubits) return false;
    for (const DecodedOp& d : ops)
        if (tile_requirement(d) == ~uint64_t{0}) return false;
    return true;
}

template <int L, int T>
void launch_persist(unsigned shots, const PStep* prog, uint32_t n_steps, const TileOp* gates,
                    const ChannelDev* chans, const double* draws, uint64_t D, uint8_t* creg, uint32_t creg_w) {
    k_traj_persist<L, T><<<shots, T, 0, stream()>>>(prog, n_steps, gates, chans, draws, D, creg, creg_w);
    LQCS_HIP(hipGetLastError());
}

// All shots through k_traj_persist, in launches of at most 2^16 shots.
void run_persistent(uint32_t nq, uint32_t n_clbits, const std::vector<DecodedOp>& ops, const lqcs_traj_step* steps,
                    uint64_t n_steps, const uint32_t* noise_qubits, const std::vector<KrausSet>& sets,
                    const std::vector<ChannelPlan>& cplan, const std::vector<uint64_t>& off, const double* draws,
                    uint64_t D, uint64_t n_shots, uint8_t* clbits_out) {
    const uint32_t L = std::max<uint32_t>(6, nq);
    const uint64_t smask = (uint64_t{1} << L) - 1;
    std::vector<TileOp> gates;
    std::vector<PStep> prog;
    const uint32_t n_channels = static_cast<uint32_t>(sets.size());
    for (uint64_t i = 0; i < n_steps; ++i) {
        const lqcs_traj_step& st = steps[i];
        PStep ps;
        std::memset(&ps, 0, sizeof ps);
        if (st.kind != LQCS_TOP_UNITARY) {
            ps.kind = st.kind == LQCS_TOP_MEASURE ? PS_MEASURE : PS_RESET;
            ps.q = st.qubit;
            ps.arg = st.kind == LQCS_TOP_MEASURE ? st.clbit : 0;
            ps.draw = off[i];
            prog.push_back(ps);
            continue;
        }
        ps.kind = PS_GATE;
        ps.arg = static_cast<uint32_t>(gates.size());
        gates.push_back(make_tile_op(ops[st.op_index], smask));
        prog.push_back(ps);
        uint64_t dd = off[i];
        for (uint32_t c = 0; c < n_channels; ++c)
            for (uint32_t j = 0; j < st.n_noise_qubits; ++j, ++dd) {
                PStep ns;
                std::memset(&ns, 0, sizeof ns);
                ns.kind = PS_NOISE;
                ns.q = noise_qubits[st.noise_qubit_offset + j];
                ns.arg = c;
                ns.draw = dd;
                prog.push_back(ns);
            }
    }
    if (prog.size() > UINT32_MAX) throw Error(LQCS_ERR_INVALID_ARGUMENT, "trajectory program too long");
    std::vector<ChannelDev> chans(std::max<uint32_t>(1, n_channels));
    for (uint32_t c = 0; c < n_channels; ++c) {
        ChannelDev& cd = chans[c];
        std::memset(&cd, 0, sizeof cd);
        const ChannelPlan& cp = cplan[c];
        cd.n = sets[c].n;
        cd.mixture = cp.mixture ? 1u : 0u;
        cd.def = cp.def;
        cd.def_identity = cp.def_identity ? 1u : 0u;
        cd.bound = cp.bound;
        for (uint32_t b = 0; b < sets[c].n; ++b) {
            for (int e = 0; e < 4; ++e) cd.k[b][e] = sets[c].k[b][e];
            if (cp.mixture) {
                cd.c[b] = cp.c[b];
                for (int e = 0; e < 4; ++e) cd.unit[b][e] = cp.unit[b].m[e];
            }
        }
    }
    const uint32_t creg_w = std::max<uint32_t>(1, n_clbits);
    const uint64_t chunk = std::min<uint64_t>(n_shots, uint64_t{1} << 16);
    const PStep* d_prog = stage(prog.data(), prog.size());
    const TileOp* d_gates = stage(gates.data(), std::max<size_t>(1, gates.size()));
    const ChannelDev* d_chans = stage(chans.data(), chans.size());
    double* d_draws = static_cast<double*>(arena(sizeof(double) * chunk * std::max<uint64_t>(D, 1)));
    uint8_t* d_creg = static_cast<uint8_t*>(arena(chunk * creg_w));
    std::vector<uint8_t> host_creg(chunk * creg_w);
    const uint32_t np = static_cast<uint32_t>(prog.size());
    for (uint64_t s0 = 0; s0 < n_shots; s0 += chunk) {
        const uint64_t nb = std::min(chunk, n_shots - s0);
        if (D)
            LQCS_HIP(hipMemcpyAsync(d_draws, draws + s0 * D, nb * D * sizeof(double), hipMemcpyHostToDevice,
                                    stream()));
        const unsigned g = static_cast<unsigned>(nb);
        switch (L) {
            case 6: launch_persist<6, 64>(g, d_prog, np, d_gates, d_chans, d_draws, D, d_creg, creg_w); break;
            case 7: launch_persist<7, 64>(g, d_prog, np, d_gates, d_chans, d_draws, D, d_creg, creg_w); break;
            case 8: launch_persist<8, 64>(g, d_prog, np, d_gates, d_chans, d_draws, D, d_creg, creg_w); break;
            case 9: launch_persist<9, 128>(g, d_prog, np, d_gates, d_chans, d_draws, D, d_creg, creg_w); break;
            case 10: launch_persist<10, 256>(g, d_prog, np, d_gates, d_chans, d_draws, D, d_creg, creg_w); break;
            case 11: launch_persist<11, 256>(g, d_prog, np, d_gates, d_chans, d_draws, D, d_creg, creg_w); break;
            default: launch_persist<12, 256>(g, d_prog, np, d_gates, d_chans, d_draws, D, d_creg, creg_w); break;
        }
        LQCS_HIP(hipMemcpyAsync(host_creg.data(), d_creg, nb * creg_w, hipMemcpyDeviceToHost, stream()));
        sync();
        for (uint64_t b = 0; b < nb; ++b)
            std::memcpy(clbits_out + (s0 + b) * n_clbits, host_creg.data() + b * creg_w, n_clbits);
    }
}

template <typename T>
T* dev_array(size_t n) {
    return static_cast<T*>(device_alloc(std::max<size_t>(n, 1) * sizeof(T)));
}

}  // namespace
}  // namespace lqcs_dev

using namespace lqcs_dev;

extern "C" lqcs_status lqcs_trajectories_run(uint32_t nq, uint32_t n_clbits, const lqcs_program* prog,
                                             const lqcs_traj_step* steps, uint64_t n_steps,
                                             const uint32_t* noise_qubits, const lqcs_channel* channels,
                                             uint32_t n_channels, const double* draws,
                                             uint64_t D, uint64_t n_shots, uint32_t max_batch,
                                             uint32_t kernel, uint8_t* clbits_out) {
    return run_guarded([&] {
        if (nq == 0 || nq > LQCS_DEVICE_MAX_QUBITS) throw Error(LQCS_ERR_INVALID_ARGUMENT, "bad register width");
        if (!prog || (n_steps && !steps)) throw Error(LQCS_ERR_INVALID_ARGUMENT, "null program or steps");
        if (n_shots == 0) return;
        if (!draws || !clbits_out) throw Error(LQCS_ERR_INVALID_ARGUMENT, "null draws or output");
        if (n_channels && !channels) throw Error(LQCS_ERR_INVALID_ARGUMENT, "null channels");
        ensure_ready();
        arena_reset();
        std::vector<DecodedOp> ops;
        for (uint64_t k = 0; k < prog->n_ops; ++k) ops.push_back(decode(*prog, k, nq));
        std::vector<KrausSet> sets(n_channels);
        std::vector<ChannelPlan> cplan(n_channels);
        for (uint32_t c = 0; c < n_channels; ++c) {
            if (channels[c].n_ops == 0 || channels[c].n_ops > kMaxKraus)
                throw Error(LQCS_ERR_INVALID_ARGUMENT, "channel needs 1..8 Kraus operators");
            std::memset(&sets[c], 0, sizeof(KrausSet));
            sets[c].n = channels[c].n_ops;
            for (uint32_t i = 0; i < channels[c].n_ops; ++i)
                for (int e = 0; e < 4; ++e)
                    sets[c].k[i][e] = make_double2(channels[c].ops[8 * i + 2 * e], channels[c].ops[8 * i + 2 * e + 1]);
            cplan[c] = plan_channel(sets[c]);
        }
        // Draw offsets per step and a consistency check against D.
        std::vector<uint64_t> off(n_steps);
        uint64_t d = 0, n_unitary = 0;
        double events_per_shot = 0.0;
        for (uint64_t i = 0; i < n_steps; ++i) {
            off[i] = d;
            const lqcs_traj_step& st = steps[i];
            if (st.kind == LQCS_TOP_UNITARY) {
                if (st.op_index >= ops.size()) throw Error(LQCS_ERR_INVALID_ARGUMENT, "step op index out of range");
                d += uint64_t(n_channels) * st.n_noise_qubits;
                ++n_unitary;
                for (uint32_t c = 0; c < n_channels; ++c) {
                    // Only state-dependent events split the shared run (mixture
                    // branches ride along as single-register ops).
                    const ChannelPlan& cp = cplan[c];
                    if (!cp.mixture) events_per_shot += st.n_noise_qubits * std::max(0.0, 1.0 - cp.bound);
                }
                for (uint32_t j = 0; j < st.n_noise_qubits; ++j)
                    if (noise_qubits[st.noise_qubit_offset + j] >= nq)
                        throw Error(LQCS_ERR_OUT_OF_RANGE, "noise qubit out of range");
            } else if (st.kind == LQCS_TOP_MEASURE || st.kind == LQCS_TOP_RESET) {
                if (st.qubit >= nq) throw Error(LQCS_ERR_OUT_OF_RANGE, "measured qubit out of range");
                if (st.kind == LQCS_TOP_MEASURE && st.clbit >= n_clbits)
                    throw Error(LQCS_ERR_OUT_OF_RANGE, "clbit out of range");
                d += 1;
            } else {
                throw Error(LQCS_ERR_INVALID_ARGUMENT, "unknown trajectory step kind");
            }
        }
        if (d != D) throw Error(LQCS_ERR_INVALID_ARGUMENT, "draws_per_shot does not match the steps");
        if (kernel > LQCS_TRAJ_PERSISTENT) throw Error(LQCS_ERR_INVALID_ARGUMENT, "unknown trajectory kernel");
        const bool fits = persistent_fits(nq, ops);
        if (kernel == LQCS_TRAJ_PERSISTENT && !fits)
            throw Error(LQCS_ERR_INVALID_ARGUMENT, "persistent trajectories need at most 12 qubits and ops of at "
                                                   "most two targets");
        if (fits && kernel != LQCS_TRAJ_SHARED_RUNS) {
            run_persistent(nq, n_clbits, ops, steps, n_steps, noise_qubits, sets, cplan, off, draws, D, n_shots,
                           clbits_out);
            return;
        }

        // Trailing measurements (nothing but measurements after them) sample
        // without collapsing (k_term_*).
        uint64_t term = n_steps;
        while (term > 0 && steps[term - 1].kind == LQCS_TOP_MEASURE) --term;

        // Batch size: at most max_batch and ~1/2 of free HBM. Each
        // state-dependent event of any register ends the batch's shared run,
        // one more pass over every register: a batch of B registers takes
        // about P0 + e B passes (P0 without events, e events per shot), which
        // per shot grows with B. So B is the fewest registers whose tiles fill
        // the GPU a few times over, or, where events are rarer, the most that
        // keep e B within a fifth of P0. (2048 shots of the 24-qubit noisy
        // workload: 115 trajectories/s at B = 79, 274 at B = 4.)
        size_t free_b = 0, total_b = 0;
        LQCS_HIP(hipMemGetInfo(&free_b, &total_b));
        const uint64_t reg_bytes = (uint64_t{1} << nq) * sizeof(amp_t);
        uint64_t B = std::max<uint64_t>(1, (free_b / 2) / reg_bytes);
        B = std::min<uint64_t>(B, max_batch ? max_batch : 65535);
        B = std::min<uint64_t>(B, n_shots);
        B = std::min<uint64_t>(B, 65535);
        if (!max_batch && events_per_shot > 0.0) {
            const TileChoice t = default_tile(nq);
            const double tiles = std::ldexp(1.0, static_cast<int>(nq) - static_cast<int>(t.L ? t.L : nq));
            const double fill = std::ceil(4.0 * 3.0 * num_cus() / tiles);
            const double p0 = std::max(1.0, double(n_unitary) / 30.0);
            const double target = std::max({1.0, fill, std::floor(0.2 * p0 / events_per_shot)});
            B = std::min<uint64_t>(B, static_cast<uint64_t>(target));
        }
        amp_t* states = nullptr;
        LQCS_HIP(hipMalloc(&states, B * reg_bytes));
        const uint32_t creg_w = std::max<uint32_t>(1, n_clbits);
        uint8_t* d_creg = dev_array<uint8_t>(B * creg_w);
        double* d_draws = dev_array<double>(B * std::max<uint64_t>(D, 1));
        double* d_stats = dev_array<double>(B * 4);
        uint32_t* d_members = dev_array<uint32_t>(B);
        uint8_t* d_skip = dev_array<uint8_t>(B);
        const unsigned term_blocks = static_cast<unsigned>(std::min<uint64_t>(pass_blocks(nq), 1024));
        uint64_t* d_vals = dev_array<uint64_t>(B);
        double* d_term = dev_array<double>(B * term_blocks * 2);
        struct Free {
            std::vector<void*> p;
            ~Free() {
                for (void* x : p) device_free(x);
            }
        } guard{{states, d_creg, d_draws, d_stats, d_members, d_skip, d_vals, d_term}};
        std::vector<uint8_t> host_creg(B * creg_w);
        const TileChoice tc = default_tile(nq);

        std::vector<DecodedOp> pending;
        uint32_t nb = 0;
        auto flush = [&] {
            if (pending.empty()) return;
            fuse_run(pending);
            const Plan plan = make_plan(tc.L ? LQCS_PLAN_GREEDY : LQCS_PLAN_GATE_BY_GATE, pending, nq, tc.L, tc.low);
            execute_plan(states, nq, tc.L, plan, pending, nb);
            pending.clear();
        };

        for (uint64_t s0 = 0; s0 < n_shots; s0 += B) {
            nb = static_cast<uint32_t>(std::min<uint64_t>(B, n_shots - s0));
            sync();
            arena_reset();
            launch_fill_basis(states, nq, 0, nb);
            LQCS_HIP(hipMemsetAsync(d_creg, 0, uint64_t(nb) * creg_w, stream()));
            if (D)
                LQCS_HIP(hipMemcpyAsync(d_draws, draws + s0 * D, uint64_t(nb) * D * sizeof(double),
                                        hipMemcpyHostToDevice, stream()));
            for (uint64_t i = 0; i < n_steps; ++i) {
                const lqcs_traj_step& st = steps[i];
                if (i >= term) {  // a trailing measurement
                    if (i == term) {
                        flush();
                        LQCS_HIP(hipMemsetAsync(d_vals, 0, uint64_t(nb) * sizeof(uint64_t), stream()));
                    }
                    uint64_t fixed = 0;
                    for (uint64_t j = term; j < i; ++j) fixed |= uint64_t{1} << steps[j].qubit;
                    k_term_partial<<<dim3(term_blocks, nb), kBlock, 0, stream()>>>(states, nq, fixed, d_vals, st.qubit,
                                                                                  d_term);
                    LQCS_HIP(hipGetLastError());
                    k_term_decide<<<nb, kBlock, 0, stream()>>>(d_term, term_blocks, d_draws, D, off[i], st.qubit,
                                                               d_vals, d_creg, creg_w, st.clbit);
                    LQCS_HIP(hipGetLastError());
                    continue;
                }
                if (st.kind != LQCS_TOP_UNITARY) {
                    flush();
                    reduce_qstats(states, nq, st.qubit, nb, d_stats);
                    const int reset = st.kind == LQCS_TOP_RESET;
                    k_collapse_stats<<<dim3(pass_blocks(nq), nb), kBlock, 0, stream()>>>(
                        states, nq, st.qubit, d_stats, d_draws, D, off[i], reset, d_creg, creg_w,
                        reset ? -1 : static_cast<int32_t>(st.clbit));
                    LQCS_HIP(hipGetLastError());
                    continue;
                }
                pending.push_back(ops[st.op_index]);
                uint64_t dd = off[i];
                for (uint32_t c = 0; c < n_channels; ++c) {
                    const ChannelPlan& cp = cplan[c];
                    for (uint32_t j = 0; j < st.n_noise_qubits; ++j, ++dd) {
                        const uint32_t q = noise_qubits[st.noise_qubit_offset + j];
                        if (cp.def < 0) continue;  // all operators zero: nothing happens
                        // Events of this site: (member, branch) for mixtures, or
                        // (member, -1) when the bound does not settle the draw.
                        std::vector<std::pair<uint32_t, int>> ev;
                        for (uint32_t b = 0; b < nb; ++b) {
                            const double u = draws[(s0 + b) * D + dd];
                            if (cp.mixture) {
                                const int k = mixture_branch(cp, u);
                                if (k != cp.def) ev.emplace_back(b, k);
                            } else if (!(u < cp.bound)) {
                                ev.emplace_back(b, -1);
                            }
                        }
                        if (ev.empty()) {
                            if (!cp.def_identity) pending.push_back(single_op(q, cp.def_op));
                            continue;
                        }
                        if (cp.mixture) {
                            // State-independent branches stay in the shared run: the
                            // default on every register, then on each event register
                            // its branch times the default's inverse (both unitary).
                            if (!cp.def_identity) pending.push_back(single_op(q, cp.def_op));
                            for (const auto& e : ev) {
                                DecodedOp op = single_op(q, cp.def_identity ? cp.unit[e.second]
                                                                            : mul_adj(cp.unit[e.second], cp.def_op));
                                op.member = static_cast<int32_t>(e.first);
                                op.only = true;
                                pending.push_back(std::move(op));
                            }
                            continue;
                        }
                        flush();
                        // Undecided members: statistics of their own registers first.
                        std::vector<uint32_t> undecided;
                        for (const auto& e : ev)
                            if (e.second < 0) undecided.push_back(e.first);
                        for (size_t k = 0; k < undecided.size(); ++k)
                            reduce_qstats(states + (uint64_t(undecided[k]) << nq), nq, q, 1, d_stats + 4 * k);
                        // Default branch on everyone else: with one event register, as
                        // the next shared run's first op; otherwise now, skipping them.
                        if (!cp.def_identity && ev.size() == 1) {
                            DecodedOp op = single_op(q, cp.def_op);
                            op.member = static_cast<int32_t>(ev[0].first);
                            op.only = false;
                            pending.push_back(std::move(op));
                        } else if (!cp.def_identity) {
                            std::vector<uint8_t> skip(nb, 0);
                            for (const auto& e : ev) skip[e.first] = 1;
                            LQCS_HIP(hipMemcpyAsync(d_skip, skip.data(), nb, hipMemcpyHostToDevice, stream()));
                            k_single_skip<<<dim3(pass_blocks(nq), nb), kBlock, 0, stream()>>>(states, nq, q,
                                                                                             cp.def_op, d_skip);
                            LQCS_HIP(hipGetLastError());
                        }
                        // Undecided: the reference's rule on their statistics.
                        if (!undecided.empty()) {
                            LQCS_HIP(hipMemcpyAsync(d_members, undecided.data(), undecided.size() * sizeof(uint32_t),
                                                    hipMemcpyHostToDevice, stream()));
                            k_kraus_members<<<dim3(pass_blocks(nq), static_cast<unsigned>(undecided.size())), kBlock,
                                              0, stream()>>>(states, nq, q, d_stats, sets[c], d_draws, D, dd,
                                                             d_members);
                            LQCS_HIP(hipGetLastError());
                        }
                        // The host vectors above are read by queued copies.
                        sync();
                    }
                }
            }
            flush();
            LQCS_HIP(hipMemcpyAsync(host_creg.data(), d_creg, uint64_t(nb) * creg_w, hipMemcpyDeviceToHost,
                                    stream()));
            sync();
            for (uint32_t b = 0; b < nb; ++b)
                std::memcpy(clbits_out + (s0 + b) * n_clbits, host_creg.data() + uint64_t(b) * creg_w, n_clbits);
        }
    });
}
