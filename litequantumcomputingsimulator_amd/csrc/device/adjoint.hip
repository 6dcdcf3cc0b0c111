// Adjoint differentiation of E(theta) = <psi|H|psi> for circuits whose
// parameters are Pauli-rotation angles, U_k = exp(-i theta_k/2 P_k).
//
//   dE/dtheta_k = 2 Re <lam_k| (-i/2) P_k |psi_k> = Im <lam_k|P_k|psi_k>,
//   psi_k = U_k ... U_1 |0>,  lam_k = U_{k+1}^dag ... U_G^dag H |psi_G>.
//
// One forward sweep, one pass building lam = H psi (all terms gathered per
// amplitude, energy = Re <psi|lam> from the same pass), then one backward
// sweep in which each Pauli-rotation step reads psi and lam once, accumulates
// its inner product and un-rotates both registers in the same kernel.
// Consecutive rotations run fused (k_fwd_group, k_bwd_group): up to D + 1
// of them whose X masks span at most D dimensions share one pass over the
// registers, each thread holding one coset of 2^D amplitudes (D = 4 forward,
// 3 backward); registers of fewer than three qubits take the one-step
// kernels (k_fwd, k_bwd).
//
// The batched form runs many circuits of different widths in one launch per
// rotation group (or per op step for the one-step kernels): a flattened
// block list maps blocks to circuits, each circuit taking at most
// kMaxBlocks blocks (grid-stride inside), so per-parameter partial sums
// have a fixed, deterministic layout.
#include "common.hpp"
#include "engine.hpp"

#include <cmath>
#include <cstring>
#include <memory>
#include <type_traits>
#include <utility>
#include <vector>

namespace lqcs_dev {
namespace {

constexpr uint32_t kMaxBlocks = 256;  // blocks per circuit per launch

struct PauliStep {
    uint64_t x, zy;
    amp_t c, yphase;   // forward rotation; the inverse uses -yphase
    int64_t param;     // global parameter index or -1
    uint32_t is_pauli; // 0: not a Pauli rotation (handled outside the kernel)
    uint32_t pad;
};

struct Term {
    uint64_t x, zy;
    amp_t w;  // coeff * i^ny
};

struct CircuitDesc {
    uint64_t reg_off;    // psi at reg_off, lam at reg_off + 2^nq
    uint32_t nq;
    uint32_t blocks;     // blocks this circuit uses per launch
    uint32_t blk_start;  // first flattened block
    uint32_t n_ops;
    uint64_t op_off;
    uint64_t term_off, n_terms;
    uint32_t n_groups;   // fused rotation groups of this sweep (widths >= 3)
    uint32_t group_off;
};

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

__device__ __forceinline__ void block_sum2(double& a, double& b) {
    __shared__ double red[2][kBlock / 64];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    a = wave_sum(a);
    b = wave_sum(b);
    if (lane == 0) {
        red[0][wid] = a;
        red[1][wid] = b;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        a = 0.0;
        b = 0.0;
        for (int w = 0; w < kBlock / 64; ++w) {
            a += red[0][w];
            b += red[1][w];
        }
    }
    __syncthreads();
}

// Flattened block -> circuit (binary search over blk_start).
__device__ __forceinline__ uint32_t find_circuit(const CircuitDesc* cd, uint32_t nc, uint32_t blk) {
    uint32_t lo = 0, hi = nc;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) / 2;
        if (cd[mid].blk_start <= blk) lo = mid; else hi = mid;
    }
    return lo;
}

__device__ __forceinline__ double zsign(uint64_t i, uint64_t zy) { return (__popcll(i & zy) & 1) ? -1.0 : 1.0; }

// Forward: psi <- exp(-i theta/2 P) psi for step `step` of every circuit.
__global__ void __launch_bounds__(kBlock) k_fwd(amp_t* __restrict__ base, const CircuitDesc* __restrict__ cd,
                                                uint32_t nc, const PauliStep* __restrict__ ops,
                                                uint32_t step) {
    const uint32_t c = find_circuit(cd, nc, blockIdx.x);
    const CircuitDesc d = cd[c];
    if (step >= d.n_ops) return;
    const PauliStep op = ops[d.op_off + step];
    if (!op.is_pauli) return;
    amp_t* s = base + d.reg_off;
    const uint32_t lb = blockIdx.x - d.blk_start;
    const uint64_t stride = uint64_t(d.blocks) * kBlock;
    if (op.x == 0) {
        const amp_t dp = cadd(op.c, op.yphase), dm = cadd(op.c, cscale(op.yphase, -1.0));
        const uint64_t n = uint64_t{1} << d.nq;
        for (uint64_t i = uint64_t(lb) * kBlock + threadIdx.x; i < n; i += stride)
            s[i] = cmul(s[i], (__popcll(i & op.zy) & 1) ? dm : dp);
        return;
    }
    const uint32_t lsb = __builtin_ctzll(op.x);
    const uint64_t half = uint64_t{1} << (d.nq - 1);
    for (uint64_t g = uint64_t(lb) * kBlock + threadIdx.x; g < half; g += stride) {
        const uint64_t i = insert_zero_bit(g, lsb), j = i ^ op.x;
        const amp_t ai = s[i], aj = s[j];
        s[i] = cmac2(op.c, ai, cscale(op.yphase, zsign(j, op.zy)), aj);
        s[j] = cmac2(cscale(op.yphase, zsign(i, op.zy)), ai, op.c, aj);
    }
}

// lam = H psi (terms gathered per amplitude); energy partial = sum conj(psi) lam.
__global__ void __launch_bounds__(kBlock) k_hamiltonian(amp_t* __restrict__ base,
                                                        const CircuitDesc* __restrict__ cd, uint32_t nc,
                                                        const Term* __restrict__ terms,
                                                        double* __restrict__ epart) {
    const uint32_t c = find_circuit(cd, nc, blockIdx.x);
    const CircuitDesc d = cd[c];
    const amp_t* psi = base + d.reg_off;
    amp_t* lam = base + d.reg_off + (uint64_t{1} << d.nq);
    const uint32_t lb = blockIdx.x - d.blk_start;
    const uint64_t stride = uint64_t(d.blocks) * kBlock;
    const uint64_t n = uint64_t{1} << d.nq;
    double er = 0.0, ei = 0.0;
    for (uint64_t i = uint64_t(lb) * kBlock + threadIdx.x; i < n; i += stride) {
        amp_t acc = make_double2(0.0, 0.0);
        for (uint64_t t = 0; t < d.n_terms; ++t) {
            const Term tm = terms[d.term_off + t];
            const uint64_t j = i ^ tm.x;
            acc = cadd(acc, cmul(cscale(tm.w, zsign(j, tm.zy)), psi[j]));
        }
        lam[i] = acc;
        const amp_t p = psi[i];
        er += p.x * acc.x + p.y * acc.y;
        ei += p.x * acc.y - p.y * acc.x;
    }
    block_sum2(er, ei);
    if (threadIdx.x == 0) {
        epart[(uint64_t(c) * kMaxBlocks + lb) * 2] = er;
        epart[(uint64_t(c) * kMaxBlocks + lb) * 2 + 1] = ei;
    }
}

// Backward step: S += sum_i conj(lam_i) sign(i^x) psi_{i^x} (before the step),
// then psi, lam <- U^dag psi, U^dag lam.
__global__ void __launch_bounds__(kBlock) k_bwd(amp_t* __restrict__ base, const CircuitDesc* __restrict__ cd,
                                                uint32_t nc, const PauliStep* __restrict__ ops,
                                                uint32_t step, double* __restrict__ gpart) {
    const uint32_t c = find_circuit(cd, nc, blockIdx.x);
    const CircuitDesc d = cd[c];
    if (step >= d.n_ops) return;
    const PauliStep op = ops[d.op_off + step];
    if (!op.is_pauli) return;
    amp_t* psi = base + d.reg_off;
    amp_t* lam = psi + (uint64_t{1} << d.nq);
    const uint32_t lb = blockIdx.x - d.blk_start;
    const uint64_t stride = uint64_t(d.blocks) * kBlock;
    const amp_t yinv = cscale(op.yphase, -1.0);
    double sr = 0.0, si = 0.0;
    if (op.x == 0) {
        const amp_t dp = cadd(op.c, yinv), dm = cadd(op.c, cscale(yinv, -1.0));
        const uint64_t n = uint64_t{1} << d.nq;
        for (uint64_t i = uint64_t(lb) * kBlock + threadIdx.x; i < n; i += stride) {
            const amp_t p = psi[i], l = lam[i];
            const double sg = zsign(i, op.zy);
            sr += sg * (l.x * p.x + l.y * p.y);
            si += sg * (l.x * p.y - l.y * p.x);
            const bool odd = (__popcll(i & op.zy) & 1) != 0;
            psi[i] = cmul(p, odd ? dm : dp);
            lam[i] = cmul(l, odd ? dm : dp);
        }
    } else {
        const uint32_t lsb = __builtin_ctzll(op.x);
        const uint64_t half = uint64_t{1} << (d.nq - 1);
        for (uint64_t g = uint64_t(lb) * kBlock + threadIdx.x; g < half; g += stride) {
            const uint64_t i = insert_zero_bit(g, lsb), j = i ^ op.x;
            const amp_t pi = psi[i], pj = psi[j], li = lam[i], lj = lam[j];
            const double sgi = zsign(i, op.zy), sgj = zsign(j, op.zy);
            // conj(li) sgj pj + conj(lj) sgi pi
            sr += sgj * (li.x * pj.x + li.y * pj.y) + sgi * (lj.x * pi.x + lj.y * pi.y);
            si += sgj * (li.x * pj.y - li.y * pj.x) + sgi * (lj.x * pi.y - lj.y * pi.x);
            psi[i] = cmac2(op.c, pi, cscale(yinv, sgj), pj);
            psi[j] = cmac2(cscale(yinv, sgi), pi, op.c, pj);
            lam[i] = cmac2(op.c, li, cscale(yinv, sgj), lj);
            lam[j] = cmac2(cscale(yinv, sgi), li, op.c, lj);
        }
    }
    if (op.param < 0) return;  // uniform per block
    block_sum2(sr, si);
    if (threadIdx.x == 0) {
        gpart[(uint64_t(op.param) * kMaxBlocks + lb) * 2] = sr;
        gpart[(uint64_t(op.param) * kMaxBlocks + lb) * 2 + 1] = si;
    }
}

// ---- fused Pauli-rotation steps ----------------------------------------------
// Up to D + 1 consecutive rotations whose X masks span at most D dimensions
// (D = 3 or 4) run as one pass: the register splits into cosets of a
// 2^D-element subgroup {off[a]} containing every X mask, each thread loads
// one coset (2^D amplitudes, index i0 ^ off[a], i0 with the pivot bits zero)
// and applies the rotations in registers. Rotation k pairs coset element a
// with a ^ coef, since X_k = off[coef]; the Z sign of index i0 ^ off[a] is
// parity(i0 & zy) ^ bit a of sgn.
constexpr uint32_t kMaxGroupDim = 4;

struct GroupStep {
    uint64_t zy;
    amp_t c, y;       // forward rotation; the inverse uses -y
    int64_t param;    // global parameter index or -1
    uint32_t coef;    // X mask in coset coordinates
    uint32_t sgn;     // bit a: parity(off[a] & zy)
};

struct Group {
    uint64_t off[1u << kMaxGroupDim];  // XOR combinations of the basis masks
    uint32_t piv[kMaxGroupDim];        // pivot bits, ascending: bit piv[t] of off[a] is bit t of a
    uint32_t n_steps;
    uint32_t step_off;
};

// new[e] = c old[e] + y sign(e ^ C) old[e ^ C], sign bits in sb.
template <uint32_t NE, uint32_t C>
__device__ __forceinline__ void rot_c(amp_t (&v)[NE], amp_t c, amp_t y, uint32_t sb) {
    amp_t t[NE];
#pragma unroll
    for (uint32_t e = 0; e < NE; ++e) {
        const amp_t ys = ((sb >> (e ^ C)) & 1u) ? cscale(y, -1.0) : y;
        t[e] = cmac2(c, v[e], ys, v[e ^ C]);
    }
#pragma unroll
    for (uint32_t e = 0; e < NE; ++e) v[e] = t[e];
}

// coef is uniform per launch: a chain of scalar compares selects the form
// with constant register indices.
template <uint32_t NE, uint32_t C = 0>
__device__ __forceinline__ void rot_any(amp_t (&v)[NE], uint32_t coef, amp_t c, amp_t y, uint32_t sb) {
    if constexpr (C + 1 < NE) {
        if (coef != C) {
            rot_any<NE, C + 1>(v, coef, c, y, sb);
            return;
        }
    }
    rot_c<NE, C>(v, c, y, sb);
}

// sum_e conj(l[e]) sign(e ^ C) p[e ^ C]
template <uint32_t NE, uint32_t C>
__device__ __forceinline__ void dot_c(const amp_t (&l)[NE], const amp_t (&p)[NE], uint32_t sb, double& sr,
                                      double& si) {
#pragma unroll
    for (uint32_t e = 0; e < NE; ++e) {
        const double sg = ((sb >> (e ^ C)) & 1u) ? -1.0 : 1.0;
        const amp_t a = l[e], b = p[e ^ C];
        sr += sg * (a.x * b.x + a.y * b.y);
        si += sg * (a.x * b.y - a.y * b.x);
    }
}

template <uint32_t NE, uint32_t C = 0>
__device__ __forceinline__ void dot_any(const amp_t (&l)[NE], const amp_t (&p)[NE], uint32_t coef, uint32_t sb,
                                        double& sr, double& si) {
    if constexpr (C + 1 < NE) {
        if (coef != C) {
            dot_any<NE, C + 1>(l, p, coef, sb, sr, si);
            return;
        }
    }
    dot_c<NE, C>(l, p, sb, sr, si);
}

template <uint32_t D>
__device__ __forceinline__ uint64_t coset_base(uint64_t g, const Group& G) {
#pragma unroll
    for (uint32_t t = 0; t < D; ++t) g = insert_zero_bit(g, G.piv[t]);
    return g;
}

__device__ __forceinline__ uint32_t sign_bits(uint64_t i0, const GroupStep& st, uint32_t all) {
    return (__popcll(i0 & st.zy) & 1) ? (st.sgn ^ all) : st.sgn;
}

// Forward: group `grp` of every circuit, cosets of 2^D.
template <uint32_t D>
__global__ void __launch_bounds__(kBlock) k_fwd_group(amp_t* __restrict__ base, const CircuitDesc* __restrict__ cd,
                                                      uint32_t nc, const Group* __restrict__ groups,
                                                      const GroupStep* __restrict__ steps, uint32_t grp) {
    constexpr uint32_t NE = 1u << D;
    const uint32_t c = find_circuit(cd, nc, blockIdx.x);
    const CircuitDesc d = cd[c];
    if (grp >= d.n_groups) return;
    const Group& G = groups[d.group_off + grp];
    amp_t* s = base + d.reg_off;
    const uint32_t lb = blockIdx.x - d.blk_start;
    const uint64_t stride = uint64_t(d.blocks) * kBlock;
    const uint64_t n_cosets = uint64_t{1} << (d.nq - D);
    const uint32_t n_steps = G.n_steps;
    for (uint64_t g = uint64_t(lb) * kBlock + threadIdx.x; g < n_cosets; g += stride) {
        const uint64_t i0 = coset_base<D>(g, G);
        amp_t v[NE];
#pragma unroll
        for (uint32_t a = 0; a < NE; ++a) v[a] = s[i0 ^ G.off[a]];
#pragma unroll
        for (uint32_t k = 0; k < D + 1; ++k) {
            if (k >= n_steps) break;
            const GroupStep& st = steps[G.step_off + k];
            rot_any<NE>(v, st.coef, st.c, st.y, sign_bits(i0, st, (1u << NE) - 1));
        }
#pragma unroll
        for (uint32_t a = 0; a < NE; ++a) s[i0 ^ G.off[a]] = v[a];
    }
}

// Backward: for the group's steps in reverse, S_k += sum conj(lam) sign
// psi-shifted (before step k is undone), then psi, lam <- U_k^dag.
template <uint32_t D>
__global__ void __launch_bounds__(kBlock) k_bwd_group(amp_t* __restrict__ base, const CircuitDesc* __restrict__ cd,
                                                      uint32_t nc, const Group* __restrict__ groups,
                                                      const GroupStep* __restrict__ steps, uint32_t grp,
                                                      double* __restrict__ gpart) {
    constexpr uint32_t NE = 1u << D;
    constexpr uint32_t kSteps = D + 1;
    const uint32_t c = find_circuit(cd, nc, blockIdx.x);
    const CircuitDesc d = cd[c];
    if (grp >= d.n_groups) return;
    const Group& G = groups[d.group_off + grp];
    amp_t* psi = base + d.reg_off;
    amp_t* lam = psi + (uint64_t{1} << d.nq);
    const uint32_t lb = blockIdx.x - d.blk_start;
    const uint64_t stride = uint64_t(d.blocks) * kBlock;
    const uint64_t n_cosets = uint64_t{1} << (d.nq - D);
    const uint32_t n_steps = G.n_steps;
    double sr[kSteps] = {}, si[kSteps] = {};
    for (uint64_t g = uint64_t(lb) * kBlock + threadIdx.x; g < n_cosets; g += stride) {
        const uint64_t i0 = coset_base<D>(g, G);
        amp_t p[NE], l[NE];
#pragma unroll
        for (uint32_t a = 0; a < NE; ++a) {
            p[a] = psi[i0 ^ G.off[a]];
            l[a] = lam[i0 ^ G.off[a]];
        }
#pragma unroll
        for (uint32_t r = 0; r < kSteps; ++r) {
            const uint32_t k = kSteps - 1 - r;
            if (k >= n_steps) continue;
            const GroupStep& st = steps[G.step_off + k];
            const uint32_t sb = sign_bits(i0, st, (1u << NE) - 1);
            dot_any<NE>(l, p, st.coef, sb, sr[k], si[k]);
            const amp_t yinv = cscale(st.y, -1.0);
            rot_any<NE>(p, st.coef, st.c, yinv, sb);
            rot_any<NE>(l, st.coef, st.c, yinv, sb);
        }
#pragma unroll
        for (uint32_t a = 0; a < NE; ++a) {
            psi[i0 ^ G.off[a]] = p[a];
            lam[i0 ^ G.off[a]] = l[a];
        }
    }
#pragma unroll
    for (uint32_t k = 0; k < kSteps; ++k) {
        if (k >= n_steps) break;
        const int64_t param = steps[G.step_off + k].param;
        if (param < 0) continue;  // uniform per block
        block_sum2(sr[k], si[k]);
        if (threadIdx.x == 0) {
            gpart[(uint64_t(param) * kMaxBlocks + lb) * 2] = sr[k];
            gpart[(uint64_t(param) * kMaxBlocks + lb) * 2 + 1] = si[k];
        }
    }
}

// Row r of `part` (kMaxBlocks pairs) summed over its first counts[r] entries.
__global__ void __launch_bounds__(kBlock) k_rows(const double* __restrict__ part, const uint32_t* __restrict__ counts,
                                                 double* __restrict__ out) {
    const uint32_t r = blockIdx.x;
    double a = 0.0, b = 0.0;
    for (uint32_t j = threadIdx.x; j < counts[r]; j += kBlock) {
        a += part[(uint64_t(r) * kMaxBlocks + j) * 2];
        b += part[(uint64_t(r) * kMaxBlocks + j) * 2 + 1];
    }
    block_sum2(a, b);
    if (threadIdx.x == 0) {
        out[2 * r] = a;
        out[2 * r + 1] = b;
    }
}

__global__ void k_zero2(amp_t* __restrict__ base, const CircuitDesc* __restrict__ cd, uint32_t nc) {
    const uint32_t c = find_circuit(cd, nc, blockIdx.x);
    const CircuitDesc d = cd[c];
    const uint32_t lb = blockIdx.x - d.blk_start;
    const uint64_t n = uint64_t{1} << d.nq;
    amp_t* psi = base + d.reg_off;
    for (uint64_t i = uint64_t(lb) * kBlock + threadIdx.x; i < 2 * n; i += uint64_t(d.blocks) * kBlock)
        psi[i] = make_double2(i == 0 ? 1.0 : 0.0, 0.0);
}

// Pair coefficients of exp(-i theta/2 P): c = cos(theta/2) and
// yphase = -i sin(theta/2) i^ny.
void rotation_coeffs(const DecodedOp& d, amp_t& c, amp_t& yphase) {
    const double h = d.theta / 2, sn = std::sin(h);
    static const amp_t kIPow[4] = {{1, 0}, {0, 1}, {-1, 0}, {0, -1}};
    const amp_t ip = kIPow[d.n_y % 4];
    c = make_double2(std::cos(h), 0.0);
    yphase = make_double2(sn * ip.y, -sn * ip.x);  // (0, -sin h) * i^ny
}

PauliStep make_step(const DecodedOp& d, int64_t param) {
    PauliStep s;
    std::memset(&s, 0, sizeof s);
    s.param = param;
    if (d.kind != LQCS_OP_PAULI_ROTATION) return s;
    s.is_pauli = 1;
    s.x = d.x;
    s.zy = d.zy;
    rotation_coeffs(d, s.c, s.yphase);
    return s;
}

DecodedOp inverse_op(const DecodedOp& d) {
    DecodedOp v = d;
    auto conj_t = [](std::vector<amp_t>& m, uint64_t D) {
        std::vector<amp_t> o(m.size());
        for (uint64_t r = 0; r < D; ++r)
            for (uint64_t c = 0; c < D; ++c) o[c * D + r] = make_double2(m[r * D + c].x, -m[r * D + c].y);
        m.swap(o);
    };
    switch (d.kind) {
        case LQCS_OP_SINGLE:
        case LQCS_OP_CONTROLLED_SINGLE: conj_t(v.m, 2); break;
        case LQCS_OP_TWO_QUBIT: conj_t(v.m, 4); break;
        case LQCS_OP_CONTROLLED_UNITARY: conj_t(v.m, uint64_t{1} << d.targets.size()); break;
        case LQCS_OP_DIAGONAL_SINGLE:
            for (auto& z : v.m) z.y = -z.y;
            break;
        case LQCS_OP_CONTROLLED_PERMUTATION:
            for (uint64_t l = 0; l < d.perm.size(); ++l) v.perm[d.perm[l]] = l;
            break;
        case LQCS_OP_PAULI_ROTATION: v.theta = -d.theta; break;
        default: break;
    }
    return v;
}

struct Job {
    uint32_t nq;
    std::vector<DecodedOp> ops;
    std::vector<int64_t> param_of_op;  // the circuit's own parameter index per op, -1 for none
    uint64_t n_params;
    std::vector<Term> terms;
    bool all_pauli;
};

// Slot of op k's parameter within a group whose slots for this job start at
// `base`; -1 when the op carries no parameter.
int64_t param_slot(const Job& j, size_t k, uint64_t base) {
    return j.param_of_op[k] >= 0 ? static_cast<int64_t>(base) + j.param_of_op[k] : -1;
}

// One entry of a circuit's fused schedule: a non-Pauli op (run gate by gate)
// or a group of consecutive Pauli rotations.
struct Unit {
    bool pauli;
    uint32_t index;  // op index, or group index within the circuit
};

// Splits the Pauli rotations ops[first, last) into groups of at most D + 1
// whose X masks span at most D dimensions (greedy, in order). The span's
// basis is kept in reduced row-echelon form, pivot = highest set bit, and
// padded to D vectors with the highest free qubits so that every group
// walks cosets of 2^D.
// src receives the op index of every step it appends.
void group_rotations(const Job& j, uint64_t param_base, uint32_t D, size_t first, size_t last,
                     std::vector<Group>& groups, std::vector<GroupStep>& gsteps, std::vector<uint32_t>& src,
                     std::vector<Unit>& units, uint32_t& n_groups) {
    const uint32_t NE = 1u << D;
    size_t k = first;
    while (k < last) {
        uint64_t basis[kMaxGroupDim] = {};
        uint32_t piv[kMaxGroupDim] = {};
        uint32_t dim = 0;
        auto reduce = [&](uint64_t v) {
            for (uint32_t t = 0; t < dim; ++t)
                if ((v >> piv[t]) & 1) v ^= basis[t];
            return v;
        };
        auto add = [&](uint64_t v) {  // v reduced, nonzero
            const uint32_t p = 63 - static_cast<uint32_t>(__builtin_clzll(v));
            for (uint32_t t = 0; t < dim; ++t)
                if ((basis[t] >> p) & 1) basis[t] ^= v;
            basis[dim] = v;
            piv[dim++] = p;
        };
        size_t end = k;
        while (end < last && end - k < D + 1) {
            const uint64_t v = reduce(j.ops[end].x);
            if (v != 0) {
                if (dim == D) break;
                add(v);
            }
            ++end;
        }
        for (int q = static_cast<int>(j.nq) - 1; q >= 0 && dim < D; --q) {
            bool used = false;
            for (uint32_t t = 0; t < dim; ++t) used = used || piv[t] == static_cast<uint32_t>(q);
            if (!used) add(uint64_t{1} << q);
        }
        uint32_t order[kMaxGroupDim] = {0, 1, 2, 3};
        std::sort(order, order + D, [&](uint32_t a, uint32_t b) { return piv[a] < piv[b]; });
        Group G;
        std::memset(&G, 0, sizeof G);
        for (uint32_t t = 0; t < D; ++t) G.piv[t] = piv[order[t]];
        for (uint32_t a = 0; a < NE; ++a)
            for (uint32_t t = 0; t < D; ++t)
                if ((a >> t) & 1) G.off[a] ^= basis[order[t]];
        G.n_steps = static_cast<uint32_t>(end - k);
        G.step_off = static_cast<uint32_t>(gsteps.size());
        for (size_t s = k; s < end; ++s) {
            const PauliStep ps = make_step(j.ops[s], param_slot(j, s, param_base));
            GroupStep st;
            std::memset(&st, 0, sizeof st);
            st.zy = ps.zy;
            st.c = ps.c;
            st.y = ps.yphase;
            st.param = ps.param;
            for (uint32_t t = 0; t < D; ++t) st.coef |= static_cast<uint32_t>((ps.x >> G.piv[t]) & 1) << t;
            if (G.off[st.coef] != ps.x) throw Error(LQCS_ERR_LOGIC, "rotation group basis does not span the X mask");
            for (uint32_t a = 0; a < NE; ++a)
                st.sgn |= static_cast<uint32_t>(__builtin_popcountll(G.off[a] & ps.zy) & 1) << a;
            gsteps.push_back(st);
            src.push_back(static_cast<uint32_t>(s));
        }
        units.push_back(Unit{true, n_groups++});
        groups.push_back(G);
        k = end;
    }
}

// The fused schedule of one circuit (width >= D): maximal runs of Pauli
// rotations become groups of dimension D, every other op stays a unit of
// its own.
void schedule_job(const Job& j, uint64_t param_base, uint32_t D, std::vector<Group>& groups,
                  std::vector<GroupStep>& gsteps, std::vector<uint32_t>& src, std::vector<Unit>& units,
                  uint32_t& n_groups) {
    n_groups = 0;
    size_t s = 0;
    while (s < j.ops.size()) {
        if (j.ops[s].kind != LQCS_OP_PAULI_ROTATION) {
            units.push_back(Unit{false, static_cast<uint32_t>(s)});
            ++s;
            continue;
        }
        size_t e = s;
        while (e < j.ops.size() && j.ops[e].kind == LQCS_OP_PAULI_ROTATION) ++e;
        group_rotations(j, param_base, D, s, e, groups, gsteps, src, units, n_groups);
        s = e;
    }
}

uint32_t blocks_for(uint32_t nq) {
    const uint64_t want = ((uint64_t{1} << nq) + kBlock * 16 - 1) / (kBlock * 16);
    return static_cast<uint32_t>(std::max<uint64_t>(1, std::min<uint64_t>(kMaxBlocks, want)));
}

// One sweep's fused schedule: each circuit's rotation groups and, for a
// general job, its units in program order. src[i] names the (circuit, op)
// step i was built from, so a new angle can be written into it.
struct Sched {
    std::vector<CircuitDesc> cd;
    std::vector<Group> groups;
    std::vector<GroupStep> steps;
    std::vector<std::pair<uint32_t, uint32_t>> src;
    std::vector<Unit> units;
    uint32_t max_groups = 0;
    const CircuitDesc* d_cd = nullptr;
    const Group* d_groups = nullptr;
    GroupStep* d_steps = nullptr;
};

// Everything one group of jobs that share launches needs. When `general` is
// set the group holds one job whose non-Pauli ops run through the
// gate-by-gate kernels. Parameter partials are indexed within the group: job
// c's parameter p sits at slot param_base[c] + p, so the partial buffers and
// the reduction launch scale with the group's parameters, not the whole
// batch's. Built on the host (build_work), placed in device memory
// (place_work), then launched (launch_work) with no allocation or copy in
// between, so a launch sequence can be captured in a graph.
struct Work {
    std::vector<const Job*> jobs;
    bool general = false;
    std::vector<uint64_t> param_base;
    uint64_t n_slots = 0;
    std::vector<CircuitDesc> cd;
    std::vector<PauliStep> steps;  // the one-step kernels' table (widths below 3)
    std::vector<Term> terms;
    uint64_t reg = 0;              // amplitudes: psi and lam of every circuit
    uint32_t blk = 0, max_ops = 0;
    bool fused = false;
    uint32_t Df = 3;
    Sched sf, sb;
    std::vector<uint32_t> ecount, gcount;
    const CircuitDesc* d_cd = nullptr;
    PauliStep* d_steps = nullptr;
    const Term* d_terms = nullptr;
    double *epart = nullptr, *gpart = nullptr, *d_e = nullptr, *d_g = nullptr;
    const uint32_t *d_ecount = nullptr, *d_gcount = nullptr;
};

Work build_work(const std::vector<const Job*>& jobs, bool general) {
    Work w;
    w.jobs = jobs;
    w.general = general;
    const uint32_t nc = static_cast<uint32_t>(jobs.size());
    w.param_base.assign(nc, 0);
    for (uint32_t c = 0; c < nc; ++c) {
        w.param_base[c] = w.n_slots;
        w.n_slots += jobs[c]->n_params;
    }
    w.cd.resize(nc);
    // Fused rotation groups walk cosets of 2^D, so every width must be >= D.
    // The forward holds psi alone and takes D = 4 where the widths allow it;
    // the backward holds psi and lam and takes D = 3 (at D = 4 it spills).
    uint32_t min_nq = 64;
    for (const Job* j : jobs) min_nq = std::min(min_nq, j->nq);
    w.fused = min_nq >= 3;
    w.Df = min_nq >= kMaxGroupDim ? kMaxGroupDim : 3;
    for (uint32_t c = 0; c < nc; ++c) {
        const Job& j = *jobs[c];
        CircuitDesc& d = w.cd[c];
        d.reg_off = w.reg;
        d.nq = j.nq;
        d.blocks = blocks_for(j.nq);
        d.blk_start = w.blk;
        d.n_ops = static_cast<uint32_t>(j.ops.size());
        d.op_off = w.steps.size();
        d.term_off = w.terms.size();
        d.n_terms = j.terms.size();
        for (size_t k = 0; k < j.ops.size(); ++k)
            w.steps.push_back(make_step(j.ops[k], param_slot(j, k, w.param_base[c])));
        d.group_off = 0;
        d.n_groups = 0;
        w.terms.insert(w.terms.end(), j.terms.begin(), j.terms.end());
        w.reg += uint64_t{2} << j.nq;
        w.blk += d.blocks;
        w.max_ops = std::max<uint32_t>(w.max_ops, d.n_ops);
    }
    auto build_sched = [&](Sched& S, uint32_t D) {
        S.cd = w.cd;
        for (uint32_t c = 0; c < nc; ++c) {
            S.cd[c].group_off = static_cast<uint32_t>(S.groups.size());
            std::vector<Unit> u;
            std::vector<uint32_t> src;
            schedule_job(*jobs[c], w.param_base[c], D, S.groups, S.steps, src, u, S.cd[c].n_groups);
            for (uint32_t k : src) S.src.emplace_back(c, k);
            if (c == 0) S.units.swap(u);
            S.max_groups = std::max(S.max_groups, S.cd[c].n_groups);
        }
    };
    if (w.fused) {
        build_sched(w.sf, w.Df);
        build_sched(w.sb, 3);
    }
    w.ecount.resize(nc);
    for (uint32_t c = 0; c < nc; ++c) w.ecount[c] = w.cd[c].blocks;
    w.gcount.assign(std::max<uint64_t>(1, w.n_slots), 0);
    for (uint32_t c = 0; c < nc; ++c)
        for (size_t k = 0; k < jobs[c]->ops.size(); ++k)
            if (jobs[c]->param_of_op[k] >= 0) w.gcount[w.param_base[c] + jobs[c]->param_of_op[k]] = w.cd[c].blocks;
    return w;
}

// Each put*(host, bytes) returns device memory of `bytes` holding a copy of
// `host` (uninitialised when host is null): put_dyn receives the tables that
// depend on the rotation angles, put_out the energies and parameter sums,
// put everything else.
template <class Put, class PutDyn, class PutOut>
void place_work(Work& w, Put&& put, PutDyn&& put_dyn, PutOut&& put_out) {
    auto tab = [](auto&& to, const auto& v) {
        using T = typename std::decay_t<decltype(v)>::value_type;
        return static_cast<T*>(to(static_cast<const void*>(v.data()), std::max<size_t>(1, v.size()) * sizeof(T)));
    };
    const uint32_t nc = static_cast<uint32_t>(w.jobs.size());
    w.d_cd = tab(put, w.cd);
    w.d_steps = tab(put_dyn, w.steps);
    w.d_terms = tab(put, w.terms);
    w.epart = static_cast<double*>(put(nullptr, sizeof(double) * 2 * kMaxBlocks * nc));
    w.gpart = static_cast<double*>(put(nullptr, sizeof(double) * 2 * kMaxBlocks * std::max<uint64_t>(1, w.n_slots)));
    if (w.fused) {
        for (Sched* S : {&w.sf, &w.sb}) {
            S->d_cd = tab(put, S->cd);
            S->d_groups = tab(put, S->groups);
            S->d_steps = tab(put_dyn, S->steps);
        }
    }
    w.d_ecount = tab(put, w.ecount);
    w.d_gcount = tab(put, w.gcount);
    w.d_e = static_cast<double*>(put_out(nullptr, sizeof(double) * 2 * nc));
    w.d_g = static_cast<double*>(put_out(nullptr, sizeof(double) * 2 * w.gcount.size()));
}

// The group's whole launch sequence on the engine stream: kernel launches
// only, registers at `base` (w.reg amplitudes).
void launch_work(const Work& w, amp_t* base) {
    const uint32_t nc = static_cast<uint32_t>(w.jobs.size()), blk = w.blk;
    const Job& j0 = *w.jobs[0];
    auto fwd_group = [&](uint32_t g) {
        const Sched& S = w.sf;
        if (w.Df == 4) k_fwd_group<4><<<blk, kBlock, 0, stream()>>>(base, S.d_cd, nc, S.d_groups, S.d_steps, g);
        else k_fwd_group<3><<<blk, kBlock, 0, stream()>>>(base, S.d_cd, nc, S.d_groups, S.d_steps, g);
        LQCS_HIP(hipGetLastError());
    };
    auto bwd_group = [&](uint32_t g) {
        const Sched& S = w.sb;
        k_bwd_group<3><<<blk, kBlock, 0, stream()>>>(base, S.d_cd, nc, S.d_groups, S.d_steps, g, w.gpart);
        LQCS_HIP(hipGetLastError());
    };

    k_zero2<<<blk, kBlock, 0, stream()>>>(base, w.d_cd, nc);
    // forward
    if (w.fused && w.general) {
        for (const Unit& u : w.sf.units) {
            if (u.pauli) fwd_group(u.index);
            else run_single_op(base, j0.nq, j0.ops[u.index], 1);
        }
    } else if (w.fused) {
        for (uint32_t g = 0; g < w.sf.max_groups; ++g) fwd_group(g);
    }
    for (uint32_t s = 0; !w.fused && s < w.max_ops; ++s) {
        if (w.general && j0.ops[s].kind != LQCS_OP_PAULI_ROTATION) {
            run_single_op(base, j0.nq, j0.ops[s], 1);
            continue;
        }
        k_fwd<<<blk, kBlock, 0, stream()>>>(base, w.d_cd, nc, w.d_steps, s);
    }
    k_hamiltonian<<<blk, kBlock, 0, stream()>>>(base, w.d_cd, nc, w.d_terms, w.epart);
    // backward
    if (w.fused && w.general) {
        for (size_t k = w.sb.units.size(); k-- > 0;) {
            const Unit& u = w.sb.units[k];
            if (u.pauli) bwd_group(u.index);
            else run_single_op(base, j0.nq, inverse_op(j0.ops[u.index]), 2);
        }
    } else if (w.fused) {
        for (uint32_t g = w.sb.max_groups; g-- > 0;) bwd_group(g);
    }
    for (uint32_t s = w.max_ops; !w.fused && s-- > 0;) {
        if (w.general && j0.ops[s].kind != LQCS_OP_PAULI_ROTATION) {
            run_single_op(base, j0.nq, inverse_op(j0.ops[s]), 2);
            continue;
        }
        k_bwd<<<blk, kBlock, 0, stream()>>>(base, w.d_cd, nc, w.d_steps, s, w.gpart);
    }
    // per-circuit energies and per-parameter sums
    k_rows<<<nc, kBlock, 0, stream()>>>(w.epart, w.d_ecount, w.d_e);
    k_rows<<<static_cast<unsigned>(w.gcount.size()), kBlock, 0, stream()>>>(w.gpart, w.d_gcount, w.d_g);
    LQCS_HIP(hipGetLastError());
}

// Runs one group now: arena-staged tables, a register buffer for this call.
// energies_out[c] = <H> of job c; gsum holds the summed (re, im) of every
// parameter slot; param_base[c] is job c's first slot.
void run_group(const std::vector<const Job*>& jobs, bool general, std::vector<double>& energies_out,
               std::vector<double>& gsum, std::vector<uint64_t>& param_base) {
    Work w = build_work(jobs, general);
    auto put = [](const void* host, size_t bytes) {
        void* d = arena(bytes);
        if (host) LQCS_HIP(hipMemcpyAsync(d, host, bytes, hipMemcpyHostToDevice, stream()));
        return d;
    };
    place_work(w, put, put, put);
    amp_t* base = static_cast<amp_t*>(device_alloc(w.reg * sizeof(amp_t)));
    struct Free {
        void* p;
        ~Free() { device_free(p); }
    } guard{base};
    launch_work(w, base);
    const uint32_t nc = static_cast<uint32_t>(jobs.size());
    std::vector<double> e(2 * nc);
    LQCS_HIP(hipMemcpyAsync(e.data(), w.d_e, sizeof(double) * 2 * nc, hipMemcpyDeviceToHost, stream()));
    gsum.assign(2 * w.gcount.size(), 0.0);
    LQCS_HIP(hipMemcpyAsync(gsum.data(), w.d_g, sizeof(double) * gsum.size(), hipMemcpyDeviceToHost, stream()));
    sync();
    energies_out.resize(nc);
    for (uint32_t c = 0; c < nc; ++c) energies_out[c] = e[2 * c];
    param_base = w.param_base;
}

Job build_job(uint32_t nq, const lqcs_program& prog, const uint64_t* param_op, uint64_t n_params,
              const uint64_t* xs, const uint64_t* zys, const uint32_t* nys, const double* coeffs,
              uint64_t n_terms) {
    if (nq == 0 || nq > LQCS_DEVICE_MAX_QUBITS) throw Error(LQCS_ERR_INVALID_ARGUMENT, "bad register width");
    Job j;
    j.nq = nq;
    for (uint64_t k = 0; k < prog.n_ops; ++k) j.ops.push_back(decode(prog, k, nq));
    j.param_of_op.assign(j.ops.size(), -1);
    for (uint64_t p = 0; p < n_params; ++p) {
        if (param_op[p] >= j.ops.size()) throw Error(LQCS_ERR_OUT_OF_RANGE, "parameter op index out of range");
        if (j.ops[param_op[p]].kind != LQCS_OP_PAULI_ROTATION)
            throw Error(LQCS_ERR_INVALID_ARGUMENT, "parameter op must be a Pauli rotation");
        if (j.param_of_op[param_op[p]] >= 0) throw Error(LQCS_ERR_INVALID_ARGUMENT, "op bound to two parameters");
        j.param_of_op[param_op[p]] = static_cast<int64_t>(p);
    }
    j.n_params = n_params;
    const uint64_t full = (uint64_t{1} << nq) - 1;
    static const amp_t kIPow[4] = {{1, 0}, {0, 1}, {-1, 0}, {0, -1}};
    for (uint64_t t = 0; t < n_terms; ++t) {
        if ((xs[t] | zys[t]) & ~full) throw Error(LQCS_ERR_OUT_OF_RANGE, "pauli mask out of range");
        const amp_t ip = kIPow[nys[t] % 4];
        j.terms.push_back(Term{xs[t], zys[t], make_double2(coeffs[t] * ip.x, coeffs[t] * ip.y)});
    }
    j.all_pauli = true;
    for (const auto& o : j.ops) j.all_pauli = j.all_pauli && o.kind == LQCS_OP_PAULI_ROTATION;
    return j;
}

// grad_p = Im(i^ny_p * S_p) from the accumulated S_p at its slot.
double grad_from(const Job& j, uint64_t slot, uint64_t op_index, const double* g) {
    const double sr = g[2 * slot], si = g[2 * slot + 1];
    switch (j.ops[op_index].n_y % 4) {
        case 0: return si;
        case 1: return sr;   // Im(i (sr + i si)) = sr
        case 2: return -si;
        default: return -sr;
    }
}

// Argument checks shared by the batched entry points.
void check_batch(uint32_t n_circuits, const uint32_t* n_qubits, const lqcs_program* progs, const uint64_t* param_op,
                 const uint64_t* param_off, const uint64_t* x_masks, const uint64_t* zy_masks, const uint32_t* n_ys,
                 const double* coeffs, const uint64_t* term_off, bool need_grads) {
    if (!n_qubits || !progs || !param_off || !term_off) throw Error(LQCS_ERR_INVALID_ARGUMENT, "null argument");
    for (uint32_t c = 0; c < n_circuits; ++c)
        if (param_off[c + 1] < param_off[c] || term_off[c + 1] < term_off[c])
            throw Error(LQCS_ERR_INVALID_ARGUMENT, "offsets must be non-decreasing");
    if (param_off[n_circuits] > param_off[0] && (!param_op || !need_grads))
        throw Error(LQCS_ERR_INVALID_ARGUMENT, "null argument");
    if (term_off[n_circuits] > term_off[0] && (!x_masks || !zy_masks || !n_ys || !coeffs))
        throw Error(LQCS_ERR_INVALID_ARGUMENT, "null argument");
}

std::vector<Job> build_jobs(uint32_t n_circuits, const uint32_t* n_qubits, const lqcs_program* progs,
                            const uint64_t* param_op, const uint64_t* param_off, const uint64_t* x_masks,
                            const uint64_t* zy_masks, const uint32_t* n_ys, const double* coeffs,
                            const uint64_t* term_off) {
    std::vector<Job> jobs;
    jobs.reserve(n_circuits);
    for (uint32_t c = 0; c < n_circuits; ++c) {
        const uint64_t p0 = param_off[c], p1 = param_off[c + 1];
        const uint64_t t0 = term_off[c], t1 = term_off[c + 1];
        jobs.push_back(build_job(n_qubits[c], progs[c], param_op + p0, p1 - p0, x_masks + t0, zy_masks + t0,
                                 n_ys + t0, coeffs + t0, t1 - t0));
    }
    return jobs;
}

// The batch's launch groups, in circuit order: Pauli-only circuits share
// launches in groups whose registers fit half the free HBM (at most 4096
// circuits each); any other circuit is a general group of its own.
struct LaunchGroup {
    std::vector<uint32_t> ids;
    bool general;
};

std::vector<LaunchGroup> launch_groups(const std::vector<Job>& jobs) {
    size_t free_b = 0, total_b = 0;
    LQCS_HIP(hipMemGetInfo(&free_b, &total_b));
    const uint64_t budget = free_b / 2;
    std::vector<LaunchGroup> out;
    LaunchGroup cur{{}, false};
    uint64_t bytes = 0;
    auto flush = [&] {
        if (cur.ids.empty()) return;
        out.push_back(std::move(cur));
        cur = LaunchGroup{{}, false};
        bytes = 0;
    };
    for (uint32_t c = 0; c < jobs.size(); ++c) {
        const Job& j = jobs[c];
        if (!j.all_pauli) {
            flush();
            out.push_back(LaunchGroup{{c}, true});
            continue;
        }
        const uint64_t need = (uint64_t{2} << j.nq) * sizeof(amp_t);
        if (!cur.ids.empty() && (bytes + need > budget || cur.ids.size() >= 4096)) flush();
        cur.ids.push_back(c);
        bytes += need;
    }
    flush();
    return out;
}

}  // namespace
}  // namespace lqcs_dev

using namespace lqcs_dev;

extern "C" lqcs_status lqcs_adjoint_gradient(uint32_t nq, const lqcs_program* prog, const uint64_t* param_op,
                                             uint64_t n_params, const uint64_t* x_masks,
                                             const uint64_t* zy_masks, const uint32_t* n_ys,
                                             const double* coeffs, uint64_t n_terms, double* energy,
                                             double* grad) {
    return run_guarded([&] {
        if (!prog || (n_params && (!param_op || !grad)) || (n_terms && (!x_masks || !zy_masks || !n_ys || !coeffs)))
            throw Error(LQCS_ERR_INVALID_ARGUMENT, "null argument");
        ensure_ready();
        arena_reset();
        const Job j = build_job(nq, *prog, param_op, n_params, x_masks, zy_masks, n_ys, coeffs, n_terms);
        std::vector<double> e, g;
        std::vector<uint64_t> pb;
        run_group({&j}, !j.all_pauli, e, g, pb);
        if (energy) *energy = e[0];
        for (uint64_t p = 0; p < n_params; ++p) grad[p] = grad_from(j, p, param_op[p], g.data());
    });
}

extern "C" lqcs_status lqcs_adjoint_gradient_batched(uint32_t n_circuits, const uint32_t* n_qubits,
                                                     const lqcs_program* progs, const uint64_t* param_op,
                                                     const uint64_t* param_off, const uint64_t* x_masks,
                                                     const uint64_t* zy_masks, const uint32_t* n_ys,
                                                     const double* coeffs, const uint64_t* term_off,
                                                     double* energies, double* grads) {
    return run_guarded([&] {
        if (n_circuits == 0) return;
        check_batch(n_circuits, n_qubits, progs, param_op, param_off, x_masks, zy_masks, n_ys, coeffs, term_off,
                    grads != nullptr);
        ensure_ready();
        arena_reset();
        const std::vector<Job> jobs =
            build_jobs(n_circuits, n_qubits, progs, param_op, param_off, x_masks, zy_masks, n_ys, coeffs, term_off);
        for (const LaunchGroup& lg : launch_groups(jobs)) {
            std::vector<const Job*> group;
            for (uint32_t c : lg.ids) group.push_back(&jobs[c]);
            std::vector<double> e, g;
            std::vector<uint64_t> pb;
            run_group(group, lg.general, e, g, pb);
            for (size_t i = 0; i < lg.ids.size(); ++i) {
                const uint32_t c = lg.ids[i];
                if (energies) energies[c] = e[i];
                for (uint64_t p = param_off[c]; p < param_off[c + 1]; ++p)
                    grads[p] = grad_from(jobs[c], pb[i] + (p - param_off[c]), param_op[p], g.data());
            }
            arena_reset();
        }
    });
}

// ---- captured plans ----------------------------------------------------------
// A batch whose circuits, parameter positions and Hamiltonians stay fixed
// while the angles change (an optimiser loop): the launch groups are built
// and placed in device memory once, their whole launch sequence is captured
// in one HIP graph, and each run writes the new angles into the rotation
// tables (one host-to-device copy), replays the graph and reads every
// group's energies and parameter sums back (one copy).
struct lqcs_adjoint_plan {
    std::vector<Job> jobs;
    std::vector<uint64_t> param_op, param_off;
    std::vector<LaunchGroup> groups;
    std::vector<Work> works;
    void* buf = nullptr;        // tables and partial sums of every group
    amp_t* regs = nullptr;      // registers, used by the groups in turn
    char* dyn_host = nullptr;   // pinned mirror of the angle-dependent tables
    char* dyn_dev = nullptr;
    size_t dyn_bytes = 0;
    double* out_host = nullptr; // pinned copy of every group's energies and sums
    double* out_dev = nullptr;
    size_t out_bytes = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;

    ~lqcs_adjoint_plan() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        if (dyn_host) (void)hipHostFree(dyn_host);
        if (out_host) (void)hipHostFree(out_host);
        device_free(buf);
        device_free(dyn_dev);
        device_free(out_dev);
        device_free(regs);
    }
};

namespace lqcs_dev {
namespace {

// Writes job angles into the mirror of every angle-dependent table.
void refresh_angles(lqcs_adjoint_plan& P) {
    for (Work& w : P.works) {
        for (uint32_t c = 0; c < w.jobs.size(); ++c) {
            const Job& j = *w.jobs[c];
            auto* st = reinterpret_cast<PauliStep*>(P.dyn_host + (reinterpret_cast<char*>(w.d_steps) - P.dyn_dev));
            for (size_t k = 0; k < j.ops.size(); ++k)
                if (j.param_of_op[k] >= 0) rotation_coeffs(j.ops[k], st[w.cd[c].op_off + k].c, st[w.cd[c].op_off + k].yphase);
        }
        if (!w.fused) continue;
        for (Sched* S : {&w.sf, &w.sb}) {
            auto* st = reinterpret_cast<GroupStep*>(P.dyn_host + (reinterpret_cast<char*>(S->d_steps) - P.dyn_dev));
            for (size_t i = 0; i < S->src.size(); ++i) {
                const Job& j = *w.jobs[S->src[i].first];
                const uint32_t k = S->src[i].second;
                if (j.param_of_op[k] >= 0) rotation_coeffs(j.ops[k], st[i].c, st[i].y);
            }
        }
    }
}

}  // namespace
}  // namespace lqcs_dev

extern "C" lqcs_status lqcs_adjoint_plan_create(uint32_t n_circuits, const uint32_t* n_qubits,
                                                const lqcs_program* progs, const uint64_t* param_op,
                                                const uint64_t* param_off, const uint64_t* x_masks,
                                                const uint64_t* zy_masks, const uint32_t* n_ys,
                                                const double* coeffs, const uint64_t* term_off,
                                                lqcs_adjoint_plan** out) {
    return run_guarded([&] {
        if (!out) throw Error(LQCS_ERR_INVALID_ARGUMENT, "out is null");
        *out = nullptr;
        if (n_circuits == 0) throw Error(LQCS_ERR_INVALID_ARGUMENT, "a plan needs at least one circuit");
        check_batch(n_circuits, n_qubits, progs, param_op, param_off, x_masks, zy_masks, n_ys, coeffs, term_off, true);
        ensure_ready();
        arena_reset();
        auto P = std::make_unique<lqcs_adjoint_plan>();
        P->jobs = build_jobs(n_circuits, n_qubits, progs, param_op, param_off, x_masks, zy_masks, n_ys, coeffs,
                             term_off);
        for (const Job& j : P->jobs)
            for (const DecodedOp& o : j.ops)
                if (o.kind == LQCS_OP_CONTROLLED_UNITARY || o.kind == LQCS_OP_CONTROLLED_PERMUTATION)
                    throw Error(LQCS_ERR_INVALID_ARGUMENT,
                                "adjoint plan: controlled-unitary and permutation ops stage tables per launch");
        if (param_off[n_circuits]) P->param_op.assign(param_op, param_op + param_off[n_circuits]);
        P->param_off.assign(param_off, param_off + n_circuits + 1);
        P->groups = launch_groups(P->jobs);
        uint64_t max_reg = 0;
        for (const LaunchGroup& lg : P->groups) {
            std::vector<const Job*> group;
            for (uint32_t c : lg.ids) group.push_back(&P->jobs[c]);
            P->works.push_back(build_work(group, lg.general));
            max_reg = std::max(max_reg, P->works.back().reg);
        }
        // Sizes first (offsets from null), then the real placement.
        auto align = [](size_t b) { return (b + 255) & ~size_t{255}; };
        size_t n_buf = 0, n_dyn = 0, n_out = 0;
        auto count = [&](size_t& n) {
            return [&n, &align](const void*, size_t bytes) {
                n += align(bytes);
                return static_cast<void*>(nullptr);
            };
        };
        for (Work& w : P->works) place_work(w, count(n_buf), count(n_dyn), count(n_out));
        P->buf = device_alloc(std::max<size_t>(n_buf, 256));
        P->dyn_dev = static_cast<char*>(device_alloc(std::max<size_t>(n_dyn, 256)));
        P->out_dev = static_cast<double*>(device_alloc(std::max<size_t>(n_out, 256)));
        P->regs = static_cast<amp_t*>(device_alloc(max_reg * sizeof(amp_t)));
        LQCS_HIP(hipHostMalloc(reinterpret_cast<void**>(&P->dyn_host), std::max<size_t>(n_dyn, 256)));
        LQCS_HIP(hipHostMalloc(reinterpret_cast<void**>(&P->out_host), std::max<size_t>(n_out, 256)));
        P->dyn_bytes = n_dyn;
        P->out_bytes = n_out;
        std::vector<char> tables(n_buf);
        size_t o_buf = 0, o_dyn = 0, o_out = 0;
        auto put = [&](const void* host, size_t bytes) {
            if (host) std::memcpy(tables.data() + o_buf, host, bytes);
            void* d = static_cast<char*>(P->buf) + o_buf;
            o_buf += align(bytes);
            return d;
        };
        auto put_dyn = [&](const void* host, size_t bytes) {
            if (host) std::memcpy(P->dyn_host + o_dyn, host, bytes);
            void* d = P->dyn_dev + o_dyn;
            o_dyn += align(bytes);
            return d;
        };
        auto put_out = [&](const void*, size_t bytes) {
            void* d = reinterpret_cast<char*>(P->out_dev) + o_out;
            o_out += align(bytes);
            return d;
        };
        for (Work& w : P->works) place_work(w, put, put_dyn, put_out);
        lqcs_dev::sync();
        if (n_buf) LQCS_HIP(hipMemcpy(P->buf, tables.data(), n_buf, hipMemcpyHostToDevice));
        if (n_dyn) LQCS_HIP(hipMemcpy(P->dyn_dev, P->dyn_host, n_dyn, hipMemcpyHostToDevice));
        LQCS_HIP(hipStreamBeginCapture(stream(), hipStreamCaptureModeThreadLocal));
        try {
            for (const Work& w : P->works) launch_work(w, P->regs);
        } catch (...) {
            hipGraph_t dead = nullptr;
            (void)hipStreamEndCapture(stream(), &dead);
            if (dead) (void)hipGraphDestroy(dead);
            throw;
        }
        LQCS_HIP(hipStreamEndCapture(stream(), &P->graph));
        LQCS_HIP(hipGraphInstantiate(&P->exec, P->graph, nullptr, nullptr, 0));
        *out = P.release();
    });
}

extern "C" lqcs_status lqcs_adjoint_plan_run(lqcs_adjoint_plan* plan, const double* theta, uint64_t n_theta,
                                             double* energies, double* grads) {
    return run_guarded([&] {
        if (!plan) throw Error(LQCS_ERR_INVALID_ARGUMENT, "plan is null");
        lqcs_adjoint_plan& P = *plan;
        const uint32_t nc = static_cast<uint32_t>(P.jobs.size());
        const uint64_t n_p = P.param_off[nc];
        if (n_theta != n_p)
            throw Error(LQCS_ERR_INVALID_ARGUMENT, "adjoint plan: expected " + std::to_string(n_p) + " angles");
        if (n_p > P.param_off[0] && (!theta || !grads)) throw Error(LQCS_ERR_INVALID_ARGUMENT, "null argument");
        ensure_ready();
        for (uint32_t c = 0; c < nc; ++c)
            for (uint64_t p = P.param_off[c]; p < P.param_off[c + 1]; ++p)
                P.jobs[c].ops[P.param_op[p]].theta = theta[p];
        refresh_angles(P);
        if (P.dyn_bytes)
            LQCS_HIP(hipMemcpyAsync(P.dyn_dev, P.dyn_host, P.dyn_bytes, hipMemcpyHostToDevice, stream()));
        LQCS_HIP(hipGraphLaunch(P.exec, stream()));
        LQCS_HIP(hipMemcpyAsync(P.out_host, P.out_dev, P.out_bytes, hipMemcpyDeviceToHost, stream()));
        lqcs_dev::sync();
        for (size_t gi = 0; gi < P.works.size(); ++gi) {
            const Work& w = P.works[gi];
            const auto off = [&](const double* d) { return P.out_host + (d - P.out_dev); };
            const double* e = off(w.d_e);
            const double* g = off(w.d_g);
            const std::vector<uint32_t>& ids = P.groups[gi].ids;
            for (size_t i = 0; i < ids.size(); ++i) {
                const uint32_t c = ids[i];
                if (energies) energies[c] = e[2 * i];
                for (uint64_t p = P.param_off[c]; p < P.param_off[c + 1]; ++p)
                    grads[p] = grad_from(P.jobs[c], w.param_base[i] + (p - P.param_off[c]), P.param_op[p], g);
            }
        }
    });
}

extern "C" lqcs_status lqcs_adjoint_plan_destroy(lqcs_adjoint_plan* plan) {
    return run_guarded([&] {
        if (!plan) return;
        lqcs_dev::sync();
        delete plan;
    });
}
