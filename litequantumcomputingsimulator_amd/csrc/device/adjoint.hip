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

PauliStep make_step(const DecodedOp& d, int64_t param) {
    PauliStep s;
    std::memset(&s, 0, sizeof s);
    s.param = param;
    if (d.kind != LQCS_OP_PAULI_ROTATION) return s;
    s.is_pauli = 1;
    s.x = d.x;
    s.zy = d.zy;
    const double h = d.theta / 2, sn = std::sin(h);
    static const amp_t kIPow[4] = {{1, 0}, {0, 1}, {-1, 0}, {0, -1}};
    const amp_t ip = kIPow[d.n_y % 4];
    s.c = make_double2(std::cos(h), 0.0);
    s.yphase = make_double2(sn * ip.y, -sn * ip.x);  // (0, -sin h) * i^ny
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
void group_rotations(const Job& j, uint64_t param_base, uint32_t D, size_t first, size_t last,
                     std::vector<Group>& groups, std::vector<GroupStep>& gsteps, std::vector<Unit>& units,
                     uint32_t& n_groups) {
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
                  std::vector<GroupStep>& gsteps, std::vector<Unit>& units, uint32_t& n_groups) {
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
        group_rotations(j, param_base, D, s, e, groups, gsteps, units, n_groups);
        s = e;
    }
}

uint32_t blocks_for(uint32_t nq) {
    const uint64_t want = ((uint64_t{1} << nq) + kBlock * 16 - 1) / (kBlock * 16);
    return static_cast<uint32_t>(std::max<uint64_t>(1, std::min<uint64_t>(kMaxBlocks, want)));
}

// Runs a group of jobs that share launches. When `general` is set the group
// holds one job whose non-Pauli ops run through the gate-by-gate kernels.
// Parameter partials are indexed within the group: job c's parameter p sits
// at slot param_base[c] + p, so the partial buffers and the reduction launch
// scale with the group's parameters, not the whole batch's. `gsum` receives
// the summed (re, im) of every slot.
void run_group(const std::vector<const Job*>& jobs, bool general, std::vector<double>& energies_out,
               std::vector<double>& gsum, std::vector<uint64_t>& param_base) {
    const uint32_t nc = static_cast<uint32_t>(jobs.size());
    param_base.assign(nc, 0);
    uint64_t n_slots = 0;
    for (uint32_t c = 0; c < nc; ++c) {
        param_base[c] = n_slots;
        n_slots += jobs[c]->n_params;
    }
    std::vector<CircuitDesc> cd(nc);
    std::vector<PauliStep> steps;
    std::vector<Term> terms;
    uint64_t reg = 0;
    uint32_t blk = 0, max_ops = 0;
    // Fused rotation groups walk cosets of 2^D, so every width must be >= D.
    // The forward holds psi alone and takes D = 4 where the widths allow it;
    // the backward holds psi and lam and takes D = 3 (at D = 4 it spills).
    uint32_t min_nq = 64;
    for (const Job* j : jobs) min_nq = std::min(min_nq, j->nq);
    const bool fused = min_nq >= 3;
    const uint32_t Df = min_nq >= kMaxGroupDim ? kMaxGroupDim : 3, Db = 3;
    for (uint32_t c = 0; c < nc; ++c) {
        const Job& j = *jobs[c];
        cd[c].reg_off = reg;
        cd[c].nq = j.nq;
        cd[c].blocks = blocks_for(j.nq);
        cd[c].blk_start = blk;
        cd[c].n_ops = static_cast<uint32_t>(j.ops.size());
        cd[c].op_off = steps.size();
        cd[c].term_off = terms.size();
        cd[c].n_terms = j.terms.size();
        for (size_t k = 0; k < j.ops.size(); ++k) steps.push_back(make_step(j.ops[k], param_slot(j, k, param_base[c])));
        cd[c].group_off = 0;
        cd[c].n_groups = 0;
        terms.insert(terms.end(), j.terms.begin(), j.terms.end());
        reg += uint64_t{2} << j.nq;
        blk += cd[c].blocks;
        max_ops = std::max<uint32_t>(max_ops, cd[c].n_ops);
    }
    amp_t* base = static_cast<amp_t*>(device_alloc(reg * sizeof(amp_t)));
    struct Free {
        void* p;
        ~Free() { device_free(p); }
    } guard{base};
    const CircuitDesc* d_cd = stage(cd.data(), cd.size());
    const PauliStep* d_steps = stage(steps.data(), std::max<size_t>(1, steps.size()));
    const Term* d_terms = stage(terms.data(), std::max<size_t>(1, terms.size()));
    double* epart = static_cast<double*>(arena(sizeof(double) * 2 * kMaxBlocks * nc));
    double* gpart = static_cast<double*>(arena(sizeof(double) * 2 * kMaxBlocks * std::max<uint64_t>(1, n_slots)));

    // One fused schedule per sweep: each circuit's groups and, for jobs[0]
    // when `general`, its units in program order.
    struct Sched {
        std::vector<CircuitDesc> cd;
        std::vector<Group> groups;
        std::vector<GroupStep> steps;
        std::vector<Unit> units;
        uint32_t max_groups = 0;
        const CircuitDesc* d_cd = nullptr;
        const Group* d_groups = nullptr;
        const GroupStep* d_steps = nullptr;
    } sf, sb;
    auto build_sched = [&](Sched& S, uint32_t D) {
        S.cd = cd;
        for (uint32_t c = 0; c < nc; ++c) {
            S.cd[c].group_off = static_cast<uint32_t>(S.groups.size());
            std::vector<Unit> u;
            schedule_job(*jobs[c], param_base[c], D, S.groups, S.steps, u, S.cd[c].n_groups);
            if (c == 0) S.units.swap(u);
            S.max_groups = std::max(S.max_groups, S.cd[c].n_groups);
        }
        S.d_cd = stage(S.cd.data(), S.cd.size());
        S.d_groups = stage(S.groups.data(), std::max<size_t>(1, S.groups.size()));
        S.d_steps = stage(S.steps.data(), std::max<size_t>(1, S.steps.size()));
    };
    if (fused) {
        build_sched(sf, Df);
        build_sched(sb, Db);
    }
    auto fwd_group = [&](uint32_t g) {
        if (Df == 4) k_fwd_group<4><<<blk, kBlock, 0, stream()>>>(base, sf.d_cd, nc, sf.d_groups, sf.d_steps, g);
        else k_fwd_group<3><<<blk, kBlock, 0, stream()>>>(base, sf.d_cd, nc, sf.d_groups, sf.d_steps, g);
    };
    auto bwd_group = [&](uint32_t g) {
        k_bwd_group<3><<<blk, kBlock, 0, stream()>>>(base, sb.d_cd, nc, sb.d_groups, sb.d_steps, g, gpart);
    };

    k_zero2<<<blk, kBlock, 0, stream()>>>(base, d_cd, nc);
    // forward
    if (fused && general) {
        for (const Unit& u : sf.units) {
            if (u.pauli) fwd_group(u.index);
            else run_single_op(base, jobs[0]->nq, jobs[0]->ops[u.index], 1);
        }
    } else if (fused) {
        for (uint32_t g = 0; g < sf.max_groups; ++g) fwd_group(g);
    }
    for (uint32_t s = 0; !fused && s < max_ops; ++s) {
        if (general) {
            const DecodedOp& op = jobs[0]->ops[s];
            if (op.kind != LQCS_OP_PAULI_ROTATION) {
                run_single_op(base, jobs[0]->nq, op, 1);
                continue;
            }
        }
        k_fwd<<<blk, kBlock, 0, stream()>>>(base, d_cd, nc, d_steps, s);
    }
    k_hamiltonian<<<blk, kBlock, 0, stream()>>>(base, d_cd, nc, d_terms, epart);
    // backward
    if (fused && general) {
        for (size_t k = sb.units.size(); k-- > 0;) {
            const Unit& u = sb.units[k];
            if (u.pauli) bwd_group(u.index);
            else run_single_op(base, jobs[0]->nq, inverse_op(jobs[0]->ops[u.index]), 2);
        }
    } else if (fused) {
        for (uint32_t g = sb.max_groups; g-- > 0;) bwd_group(g);
    }
    for (uint32_t s = max_ops; !fused && s-- > 0;) {
        if (general) {
            const DecodedOp& op = jobs[0]->ops[s];
            if (op.kind != LQCS_OP_PAULI_ROTATION) {
                run_single_op(base, jobs[0]->nq, inverse_op(op), 2);
                continue;
            }
        }
        k_bwd<<<blk, kBlock, 0, stream()>>>(base, d_cd, nc, d_steps, s, gpart);
    }
    // per-circuit energies and per-parameter sums
    std::vector<uint32_t> ecount(nc);
    for (uint32_t c = 0; c < nc; ++c) ecount[c] = cd[c].blocks;
    double* d_e = static_cast<double*>(arena(sizeof(double) * 2 * nc));
    k_rows<<<nc, kBlock, 0, stream()>>>(epart, stage(ecount.data(), nc), d_e);
    std::vector<uint32_t> gcount(std::max<uint64_t>(1, n_slots), 0);
    for (uint32_t c = 0; c < nc; ++c)
        for (size_t k = 0; k < jobs[c]->ops.size(); ++k)
            if (jobs[c]->param_of_op[k] >= 0) gcount[param_base[c] + jobs[c]->param_of_op[k]] = cd[c].blocks;
    double* d_g = static_cast<double*>(arena(sizeof(double) * 2 * gcount.size()));
    k_rows<<<static_cast<unsigned>(gcount.size()), kBlock, 0, stream()>>>(gpart, stage(gcount.data(), gcount.size()), d_g);
    std::vector<double> e(2 * nc);
    LQCS_HIP(hipMemcpyAsync(e.data(), d_e, sizeof(double) * 2 * nc, hipMemcpyDeviceToHost, stream()));
    gsum.assign(2 * gcount.size(), 0.0);
    LQCS_HIP(hipMemcpyAsync(gsum.data(), d_g, sizeof(double) * gsum.size(), hipMemcpyDeviceToHost, stream()));
    sync();
    energies_out.resize(nc);
    for (uint32_t c = 0; c < nc; ++c) energies_out[c] = e[2 * c];
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
double grad_from(const Job& j, uint64_t slot, uint64_t op_index, const std::vector<double>& g) {
    const double sr = g[2 * slot], si = g[2 * slot + 1];
    switch (j.ops[op_index].n_y % 4) {
        case 0: return si;
        case 1: return sr;   // Im(i (sr + i si)) = sr
        case 2: return -si;
        default: return -sr;
    }
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
        for (uint64_t p = 0; p < n_params; ++p) grad[p] = grad_from(j, p, param_op[p], g);
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
        if (!n_qubits || !progs || !param_off || !term_off) throw Error(LQCS_ERR_INVALID_ARGUMENT, "null argument");
        for (uint32_t c = 0; c < n_circuits; ++c)
            if (param_off[c + 1] < param_off[c] || term_off[c + 1] < term_off[c])
                throw Error(LQCS_ERR_INVALID_ARGUMENT, "offsets must be non-decreasing");
        if (param_off[n_circuits] > param_off[0] && (!param_op || !grads))
            throw Error(LQCS_ERR_INVALID_ARGUMENT, "null argument");
        if (term_off[n_circuits] > term_off[0] && (!x_masks || !zy_masks || !n_ys || !coeffs))
            throw Error(LQCS_ERR_INVALID_ARGUMENT, "null argument");
        ensure_ready();
        arena_reset();
        std::vector<Job> jobs;
        jobs.reserve(n_circuits);
        for (uint32_t c = 0; c < n_circuits; ++c) {
            const uint64_t p0 = param_off[c], p1 = param_off[c + 1];
            const uint64_t t0 = term_off[c], t1 = term_off[c + 1];
            jobs.push_back(build_job(n_qubits[c], progs[c], param_op + p0, p1 - p0, x_masks + t0, zy_masks + t0,
                                     n_ys + t0, coeffs + t0, t1 - t0));
        }
        // Pauli-only circuits share launches in groups bounded by HBM use.
        size_t free_b = 0, total_b = 0;
        LQCS_HIP(hipMemGetInfo(&free_b, &total_b));
        const uint64_t budget = free_b / 2;
        std::vector<const Job*> group;
        std::vector<uint32_t> group_ids;
        uint64_t bytes = 0;
        auto flush = [&] {
            if (group.empty()) return;
            std::vector<double> e, g;
            std::vector<uint64_t> pb;
            run_group(group, false, e, g, pb);
            for (size_t i = 0; i < group.size(); ++i) {
                const uint32_t c = group_ids[i];
                if (energies) energies[c] = e[i];
                for (uint64_t p = param_off[c]; p < param_off[c + 1]; ++p)
                    grads[p] = grad_from(jobs[c], pb[i] + (p - param_off[c]), param_op[p], g);
            }
            group.clear();
            group_ids.clear();
            bytes = 0;
            arena_reset();
        };
        for (uint32_t c = 0; c < n_circuits; ++c) {
            const Job& j = jobs[c];
            if (!j.all_pauli) {
                flush();
                std::vector<double> e, g;
                std::vector<uint64_t> pb;
                run_group({&j}, true, e, g, pb);
                if (energies) energies[c] = e[0];
                for (uint64_t p = param_off[c]; p < param_off[c + 1]; ++p)
                    grads[p] = grad_from(j, p - param_off[c], param_op[p], g);
                arena_reset();
                continue;
            }
            const uint64_t need = (uint64_t{2} << j.nq) * sizeof(amp_t);
            if (!group.empty() && (bytes + need > budget || group.size() >= 4096)) flush();
            group.push_back(&j);
            group_ids.push_back(c);
            bytes += need;
        }
        flush();
    });
}
