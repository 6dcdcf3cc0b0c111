// Pass planners beyond first-fit grouping.
//
// Gain planning picks each pass's tile set S qubit by qubit: starting from
// the forced low qubits it adds the qubit whose inclusion lets the most
// pending ops run in this pass (ops in order; an op runs when its
// non-diagonal support lies in S and no earlier op sharing a qubit was left
// behind). It looks across layer boundaries the way first-fit grouping does,
// but spends tile slots on the qubits that unblock the most work.
#include "engine.hpp"

#include <algorithm>

namespace lqcs_dev {
namespace {

// Ops of `pending` that run with tile set S; fills taken/rest when given.
size_t runnable(const std::vector<DecodedOp>& ops, const std::vector<uint64_t>& req,
                const std::vector<uint32_t>& pending, uint64_t S, std::vector<uint32_t>* taken,
                std::vector<uint32_t>* rest) {
    uint64_t blocked = 0;
    size_t count = 0;
    for (uint32_t idx : pending) {
        const uint64_t R = req[idx];
        if ((ops[idx].touched & blocked) || R == ~uint64_t{0} || (R & ~S)) {
            blocked |= ops[idx].touched;
            if (rest) rest->push_back(idx);
            continue;
        }
        ++count;
        if (taken) taken->push_back(idx);
    }
    return count;
}

}  // namespace

Plan plan_gain(const std::vector<DecodedOp>& ops, uint32_t nq, uint32_t L, uint32_t low) {
    Plan plan;
    std::vector<uint64_t> req(ops.size());
    for (size_t i = 0; i < ops.size(); ++i) req[i] = tile_requirement(ops[i]);
    std::vector<uint32_t> pending(ops.size());
    for (uint32_t i = 0; i < ops.size(); ++i) pending[i] = i;
    const uint64_t low_mask = (uint64_t{1} << low) - 1;
    auto bits = [](uint64_t m) { return static_cast<uint32_t>(__builtin_popcountll(m)); };
    while (!pending.empty()) {
        if (req[pending[0]] == ~uint64_t{0}) {
            PlanStep st;
            st.tiled = false;
            st.op = pending[0];
            plan.steps.push_back(st);
            pending.erase(pending.begin());
            continue;
        }
        uint64_t S = low_mask;
        size_t have = runnable(ops, req, pending, S, nullptr, nullptr);
        while (bits(S) < L) {
            int best_q = -1;
            size_t best = have;
            for (uint32_t q = 0; q < nq; ++q) {
                if (S & (uint64_t{1} << q)) continue;
                const size_t c = runnable(ops, req, pending, S | (uint64_t{1} << q), nullptr, nullptr);
                if (c > best) {
                    best = c;
                    best_q = static_cast<int>(q);
                }
            }
            if (best_q >= 0) {
                S |= uint64_t{1} << best_q;
                have = best;
                continue;
            }
            // No single qubit helps: admit the support of the first op left out.
            uint64_t blocked = 0, add = 0;
            for (uint32_t idx : pending) {
                const uint64_t R = req[idx];
                if ((ops[idx].touched & blocked) == 0 && R != ~uint64_t{0} && (R & ~S)) {
                    add = R & ~S;
                    break;
                }
                if ((ops[idx].touched & blocked) || R == ~uint64_t{0} || (R & ~S)) blocked |= ops[idx].touched;
            }
            if (add == 0 || bits(S | add) > L) break;
            S |= add;
            have = runnable(ops, req, pending, S, nullptr, nullptr);
        }
        for (uint32_t q = 0; q < nq && bits(S) < L; ++q) S |= uint64_t{1} << q;
        std::vector<uint32_t> taken, rest;
        runnable(ops, req, pending, S, &taken, &rest);
        PlanStep st;
        st.tiled = true;
        st.smask = S;
        for (uint32_t idx : taken) st.tile_ops.push_back(make_tile_op(ops[idx], S));
        plan.steps.push_back(std::move(st));
        pending.swap(rest);
    }
    return plan;
}

void apply_learned(amp_t* s, uint32_t nq, const std::vector<DecodedOp>& ops, uint32_t L, uint32_t low,
                   uint32_t batch) {
    const TileChoice def = default_tile(nq);
    if (L == 0) L = def.L;
    if (low == 0) low = def.low;
    if (L == 0) {
        apply_ops(s, nq, ops, LQCS_PLAN_GATE_BY_GATE, 0, 0, batch);
        return;
    }
    const Plan plan = plan_gain(ops, nq, L, low);
    execute_plan(s, nq, L, plan, ops, batch);
}

}  // namespace lqcs_dev
