// Dense-block plans: the op stream grouped into blocks of at most k qubits,
// each block multiplied on the host into one 2^k x 2^k unitary and applied as
// a controlled unitary without controls -- a complex GEMM on FP64 MFMA for
// k = 4..6 (k_cu_mfma, gates.hip).
//
// Grouping is first-fit in program order, as for tiled passes, but on every
// qubit an op touches (a block cannot take a diagonal on an outside qubit the
// way a tile takes a phase from its tile number): an op joins the current
// block while the union of qubits stays within k and no op left out before it
// shares a qubit with it; ops left out wait for a later block. Ops wider than
// k run on their own.
//
// A block replaces one pass per gate (or one tiled pass per few dozen gates)
// with one pass whose cost does not depend on how many gates it holds, so it
// pays on deep circuits over few qubits; the autotuner times it beside the
// tiled plans (CompiledCircuit::autotune).
#include "common.hpp"
#include "engine.hpp"

#include <cmath>
#include <vector>

#include "../../../include/lqcs_device.h"

namespace lqcs_dev {
namespace {

// M <- d M for the 2^k x 2^k block matrix M (row-major; row = local index);
// loc[q] is the local bit of global qubit q. Every column at once, so that
// the op's index tables are built once.
void apply_small(const DecodedOp& d, const std::vector<int>& loc, uint32_t k, std::vector<amp_t>& M) {
    const uint64_t n = uint64_t{1} << k;
    auto lbit = [&](uint32_t q) { return uint64_t{1} << loc[q]; };
    auto at = [&](uint64_t row, uint64_t col) -> amp_t& { return M[row * n + col]; };
    uint64_t cm = 0;
    for (uint32_t c : d.controls) cm |= lbit(c);
    const uint32_t m = static_cast<uint32_t>(d.targets.size());
    uint64_t tm = 0;
    for (uint32_t t : d.targets) tm |= lbit(t);
    std::vector<uint64_t> off(uint64_t{1} << m, 0);  // local offset of target-subspace index a (target j <-> bit j)
    for (uint64_t a = 0; a < off.size(); ++a)
        for (uint32_t j = 0; j < m; ++j)
            if (a & (uint64_t{1} << j)) off[a] |= lbit(d.targets[j]);
    switch (d.kind) {
        case LQCS_OP_SINGLE:
        case LQCS_OP_CONTROLLED_SINGLE:
        case LQCS_OP_TWO_QUBIT:
        case LQCS_OP_CONTROLLED_UNITARY: {
            const uint64_t dim = off.size();
            std::vector<amp_t> v(dim);
            for (uint64_t i = 0; i < n; ++i) {
                if ((i & tm) || (i & cm) != cm) continue;
                for (uint64_t col = 0; col < n; ++col) {
                    for (uint64_t a = 0; a < dim; ++a) v[a] = at(i | off[a], col);
                    for (uint64_t r = 0; r < dim; ++r) {
                        amp_t acc = make_double2(0.0, 0.0);
                        for (uint64_t a = 0; a < dim; ++a) acc = cadd(acc, cmul(d.m[r * dim + a], v[a]));
                        at(i | off[r], col) = acc;
                    }
                }
            }
            return;
        }
        case LQCS_OP_DIAGONAL_SINGLE: {
            const uint64_t b = lbit(d.targets[0]);
            for (uint64_t i = 0; i < n; ++i)
                if ((i & cm) == cm)
                    for (uint64_t col = 0; col < n; ++col) at(i, col) = cmul(at(i, col), (i & b) ? d.m[1] : d.m[0]);
            return;
        }
        case LQCS_OP_SWAP: {
            const uint64_t a = lbit(d.targets[0]), b = lbit(d.targets[1]);
            for (uint64_t i = 0; i < n; ++i)
                if ((i & a) && !(i & b) && (i & cm) == cm)
                    for (uint64_t col = 0; col < n; ++col) std::swap(at(i, col), at((i & ~a) | b, col));
            return;
        }
        case LQCS_OP_CONTROLLED_PERMUTATION: {
            std::vector<amp_t> v(off.size());
            for (uint64_t i = 0; i < n; ++i) {
                if ((i & tm) || (i & cm) != cm) continue;
                for (uint64_t col = 0; col < n; ++col) {
                    for (uint64_t a = 0; a < off.size(); ++a) v[a] = at(i | off[a], col);
                    for (uint64_t a = 0; a < off.size(); ++a) at(i | off[d.perm[a]], col) = v[a];
                }
            }
            return;
        }
        case LQCS_OP_PAULI_ROTATION: {
            // exp(-i theta/2 P): c = cos(theta/2), y = -i sin(theta/2) i^ny;
            // a_i' = c a_i + y s_j a_j over pairs (i, j = i ^ x), s = the Z
            // sign of the index (Z and Y letters); x = 0: a_i *= c + y s_i.
            uint64_t xl = 0, zl = 0;
            for (uint32_t q = 0; q < 64; ++q) {
                if (d.x & (uint64_t{1} << q)) xl |= lbit(q);
                if (d.zy & (uint64_t{1} << q)) zl |= lbit(q);
            }
            const double h = d.theta / 2, c = std::cos(h), sn = std::sin(h);
            static const amp_t kIPow[4] = {{1, 0}, {0, 1}, {-1, 0}, {0, -1}};
            const amp_t ip = kIPow[d.n_y % 4];
            const amp_t y = cmul(make_double2(0.0, -sn), ip);
            auto sgn = [&](uint64_t i) { return (__builtin_popcountll(i & zl) & 1) ? -1.0 : 1.0; };
            if (xl == 0) {
                for (uint64_t i = 0; i < n; ++i) {
                    const amp_t f = cadd(make_double2(c, 0.0), cscale(y, sgn(i)));
                    for (uint64_t col = 0; col < n; ++col) at(i, col) = cmul(at(i, col), f);
                }
                return;
            }
            const uint64_t lowx = xl & (~xl + 1);
            for (uint64_t i = 0; i < n; ++i) {
                if (i & lowx) continue;
                const uint64_t j = i ^ xl;
                const amp_t yi = cscale(y, sgn(i)), yj = cscale(y, sgn(j));
                for (uint64_t col = 0; col < n; ++col) {
                    const amp_t ai = at(i, col), aj = at(j, col);
                    at(i, col) = cadd(cscale(ai, c), cmul(yj, aj));
                    at(j, col) = cadd(cmul(yi, ai), cscale(aj, c));
                }
            }
            return;
        }
        default:
            throw Error(LQCS_ERR_INVALID_ARGUMENT, "dense fusion: unknown op kind");
    }
}

// The block's unitary on qubits S (ascending; local bit j <-> S[j]), row-major.
std::vector<amp_t> block_matrix(const std::vector<DecodedOp>& ops, const std::vector<uint32_t>& members,
                                const std::vector<uint32_t>& S) {
    const uint32_t k = static_cast<uint32_t>(S.size());
    const uint64_t n = uint64_t{1} << k;
    std::vector<int> loc(64, -1);
    for (uint32_t j = 0; j < k; ++j) loc[S[j]] = static_cast<int>(j);
    std::vector<amp_t> M(n * n, make_double2(0.0, 0.0));
    for (uint64_t r = 0; r < n; ++r) M[r * n + r] = make_double2(1.0, 0.0);
    for (uint32_t idx : members) apply_small(ops[idx], loc, k, M);
    return M;
}

}  // namespace

Plan plan_dense(const std::vector<DecodedOp>& ops, uint32_t nq, uint32_t k) {
    if (k < 1 || k > 6 || k > nq) throw Error(LQCS_ERR_INVALID_ARGUMENT, "dense blocks must have 1..6 qubits and <= n");
    Plan plan;
    std::vector<uint32_t> pending(ops.size());
    for (uint32_t i = 0; i < ops.size(); ++i) pending[i] = i;
    auto bits = [](uint64_t m) { return static_cast<uint32_t>(__builtin_popcountll(m)); };
    auto stand_alone = [&](uint32_t idx) {
        return bits(ops[idx].touched) > k || ops[idx].member >= 0;
    };
    while (!pending.empty()) {
        if (stand_alone(pending[0])) {
            PlanStep st;
            st.op = pending[0];
            plan.steps.push_back(st);
            pending.erase(pending.begin());
            continue;
        }
        uint64_t S = 0, blocked = 0;
        std::vector<uint32_t> taken, rest;
        for (uint32_t idx : pending) {
            const DecodedOp& d = ops[idx];
            if ((d.touched & blocked) || stand_alone(idx) || bits(S | d.touched) > k) {
                blocked |= d.touched;
                rest.push_back(idx);
                continue;
            }
            S |= d.touched;
            taken.push_back(idx);
        }
        // Pad to k qubits (lowest unused) so that k = 4..6 blocks take the MFMA kernel.
        for (uint32_t q = 0; q < nq && bits(S) < k; ++q) S |= uint64_t{1} << q;
        std::vector<uint32_t> Sv;
        for (uint32_t q = 0; q < nq; ++q)
            if (S & (uint64_t{1} << q)) Sv.push_back(q);
        DecodedOp blk;
        blk.kind = LQCS_OP_CONTROLLED_UNITARY;
        blk.targets = Sv;
        blk.touched = S;
        blk.m = block_matrix(ops, taken, Sv);
        PlanStep st;
        st.op = static_cast<uint32_t>(ops.size() + plan.fused.size());
        st.fused_ops = static_cast<uint32_t>(taken.size());
        plan.fused.push_back(std::move(blk));
        plan.steps.push_back(std::move(st));
        pending.swap(rest);
    }
    return plan;
}

}  // namespace lqcs_dev
