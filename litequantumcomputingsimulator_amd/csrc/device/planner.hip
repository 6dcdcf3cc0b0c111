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
#include <cmath>
#include <cstring>

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

// ---- learned plan choice ---------------------------------------------------------

namespace {

// Relative predicted gain a learned pick needs over greedy grouping.
constexpr double kLearnedMargin = 0.02;

struct Choice {
    int plan = LQCS_PLAN_GREEDY;
    int low = 0;
    double predicted_ms = -1.0;
    uint32_t L = 0;  // tile width
};
thread_local Choice g_last_choice;

// The learned choice is a function of the ops, the width, the batch, the
// options, the engine switches the plans depend on, and the installed model.
// Repeated applies of one program (a compiled circuit) reuse it instead of
// planning six candidates and running the network again: tens of
// milliseconds of host time per apply at 32 qubits.
struct CachedChoice {
    uint64_t key;
    Choice choice;
};
thread_local std::vector<CachedChoice> g_choice_cache;

struct Fnv {
    uint64_t h = 1469598103934665603ull;
    void bytes(const void* p, size_t n) {
        const unsigned char* b = static_cast<const unsigned char*>(p);
        for (size_t i = 0; i < n; ++i) h = (h ^ b[i]) * 1099511628211ull;
    }
    template <typename T>
    void pod(const T& v) { bytes(&v, sizeof v); }
};

uint64_t choice_key(const std::vector<DecodedOp>& ops, uint32_t nq, uint32_t batch, uint32_t L, uint32_t low_req) {
    Fnv f;
    f.pod(nq);
    f.pod(batch);
    f.pod(L);
    f.pod(low_req);
    f.pod(cost_model_version());
    f.pod(rtile_enabled());
    f.pod(g_rtile_spec);
    f.pod(g_dense_mfma_min_targets);
    for (const DecodedOp& d : ops) {
        f.pod(d.kind);
        f.pod(d.member);
        f.pod(d.only);
        f.pod(d.x);
        f.pod(d.zy);
        f.pod(d.n_y);
        f.pod(d.theta);
        f.bytes(d.controls.data(), d.controls.size() * sizeof(uint32_t));
        f.pod(uint32_t(0xffffffffu));
        f.bytes(d.targets.data(), d.targets.size() * sizeof(uint32_t));
        f.pod(uint32_t(0xffffffffu));
        f.bytes(d.m.data(), d.m.size() * sizeof(amp_t));
        f.bytes(d.perm.data(), d.perm.size() * sizeof(uint64_t));
    }
    return f.h;
}

uint32_t lowest_bits_qubits_at_least(uint64_t smask, uint32_t tmask, uint32_t lanes, uint32_t qmin) {
    // Of the `lanes` lowest tile positions in tmask, how many map to qubits >= qmin.
    uint32_t count = 0, seen = 0;
    for (uint32_t p = 0; p < 12 && seen < lanes; ++p) {
        if (!(tmask & (1u << p))) continue;
        ++seen;
        uint64_t m = smask;
        for (uint32_t k = 0; k < p; ++k) m &= m - 1;
        if (static_cast<uint32_t>(__builtin_ctzll(m)) >= qmin) ++count;
    }
    return count;
}

}  // namespace

void step_features(const Plan& plan, const PlanStep& st, const std::vector<DecodedOp>& ops, uint32_t nq, uint32_t L, uint32_t batch,
                   double* f) {
    std::fill(f, f + LQCS_COST_FEATURES, 0.0);
    f[1] = nq;
    f[2] = std::log2(double(std::max<uint32_t>(batch, 1)));
    if (st.tiled) {
        f[0] = 1.0;
        f[18] = double(L) - 12.0;  // 13-qubit tiles (specialized passes only)
        f[15] = __builtin_popcountll(st.smask >> 20);
        f[16] = __builtin_ctzll(~st.smask);
        f[17] = double(st.tile_ops.size());
        RTileProgram rp;
        if (rtile_enabled() && build_rtile(st.tile_ops, L, st.smask, nq, rp)) {
            f[3] = double(rp.phases.size());
            for (const RPhase& ph : rp.phases) f[4] += ph.transpose;
            f[5] = std::ceil(double(rp.words.size()) / 512.0);  // op words, in 8 KiB units
            for (size_t i = 0; i + 1 < rp.words.size();) {
                RegOpHdr h;
                std::memcpy(&h, &rp.words[i], sizeof h);
                const uint32_t kind = h.code & 15u, rctl = (h.code >> 12) & 15u, nm = (h.code >> 16) & 0xFFFu;
                const bool free = rctl == 0 && h.tctl == 0;
                switch (kind) {
                    case RK_DENSE1: f[free ? 6 : 7] += 1.0; break;
                    case RK_X: f[free ? 8 : 9] += 1.0; break;
                    case RK_SEL1: f[h.la ? 7 : 6] += 1.0; break;  // a 2x2 op; per-lane select like a control
                    case RK_DENSE2:
                    case RK_DENSE2R: f[10] += 1.0; break;
                    case RK_DIAG:
                    case RK_DIAG1: f[11] += 1.0; break;
                    default: f[12] += 1.0; break;
                }
                if (h.gctrl) f[13] += 1.0;
                i += 2 + nm;
            }
            f[14] = lowest_bits_qubits_at_least(st.smask, rp.phases.front().tmask, 6, 8);
        }
    } else {
        const DecodedOp& d = step_op(plan, ops, st);
        f[18] = d.kind;
        f[19] = double(d.targets.size());
        f[20] = double(d.controls.size());
        f[21] = d.diagonal ? 1.0 : 0.0;
        uint64_t tm = 0;
        for (uint32_t t : d.targets) tm |= uint64_t{1} << t;
        if (d.kind == LQCS_OP_PAULI_ROTATION) tm = d.x | d.zy;
        if (tm) {
            f[22] = __builtin_ctzll(tm);
            f[23] = 63 - __builtin_clzll(tm);
        }
    }
}

Plan choose_plan(const lqcs_exec_options* opts, const std::vector<DecodedOp>& ops, uint32_t nq, uint32_t batch,
                 uint32_t& L) {
    int kind = opts ? opts->plan : LQCS_PLAN_AUTO;
    if (kind == LQCS_PLAN_AUTO) kind = LQCS_PLAN_GREEDY;
    if (kind < LQCS_PLAN_GATE_BY_GATE || kind > LQCS_PLAN_DENSE)
        throw Error(LQCS_ERR_INVALID_ARGUMENT, "unknown plan kind");
    const TileChoice def = default_tile(nq);
    const uint32_t L_req = opts && opts->tile_qubits > 0 ? static_cast<uint32_t>(opts->tile_qubits) : 0;
    if (kind == LQCS_PLAN_DENSE) {  // tile_qubits is the block width
        L = L_req ? L_req : std::min<uint32_t>(5, nq);
        g_last_choice = Choice{kind, 0, -1.0, L};
        return plan_dense(ops, nq, L);
    }
    L = L_req ? L_req : def.L;
    const uint32_t low_req = opts && opts->low_qubits > 0 ? static_cast<uint32_t>(opts->low_qubits) : 0;
    const uint32_t low = low_req ? low_req : def.low;
    if (kind != LQCS_PLAN_LEARNED || L == 0 || !cost_model_installed()) {
        // Without a cost model the learned plan falls back to greedy grouping.
        const int k = kind == LQCS_PLAN_LEARNED ? LQCS_PLAN_GREEDY : kind;
        g_last_choice = Choice{k, static_cast<int>(low), -1.0, L};
        return make_plan(k, ops, nq, L, low);
    }
    ensure_ready();
    const uint64_t key = choice_key(ops, nq, batch, L_req, low_req);
    for (const CachedChoice& c : g_choice_cache) {
        if (c.key != key) continue;
        g_last_choice = c.choice;
        L = c.choice.L;
        return make_plan(c.choice.plan, ops, nq, L, static_cast<uint32_t>(c.choice.low));
    }
    // Candidates: greedy and gain grouping at 2..4 low qubits; for specialized
    // passes at 3..5 low qubits on 12- and 13-qubit tiles (unless the call
    // fixes the width), from registers wide enough for 13-qubit tiles, and
    // dense blocks (below).
    std::vector<uint32_t> widths{L};
    std::vector<uint32_t> lows{2u, 3u, 4u};
    if (g_rtile_spec) {
        lows = {3u, 4u, 5u};
        // 13-qubit tiles run register programs only: no Pauli rotation may
        // act densely on more than four qubits (the LDS tile kernel, which
        // runs those on narrower tiles, stops at 12).
        bool narrow = true;
        for (const DecodedOp& d : ops)
            if (d.kind == LQCS_OP_PAULI_ROTATION && __builtin_popcountll(d.x) > 4) narrow = false;
        if (!L_req && L == 12 && nq >= 17 && narrow) widths.push_back(13);
    }
    std::vector<Plan> cands;
    std::vector<Choice> meta;
    for (uint32_t w : widths) {
        for (int k : {LQCS_PLAN_GREEDY, LQCS_PLAN_GAIN}) {
            for (uint32_t lw : lows) {
                if (low_req && lw != low_req) continue;
                if (lw + 2 > w) continue;
                cands.push_back(make_plan(k, ops, nq, w, lw));
                meta.push_back(Choice{k, static_cast<int>(lw), 0.0, w});
            }
        }
    }
    // Both models are also trained on dense 4..6-qubit blocks (FP64 MFMA; 12
    // qubits and up, so that the blocks take the MFMA kernel), which win on
    // deep circuits over few qubits: candidates too, unless the call fixes the
    // tile shape.
    if (!L_req && !low_req && nq >= 12)
        for (uint32_t k : {4u, 5u, 6u}) {
            cands.push_back(plan_dense(ops, nq, k));
            meta.push_back(Choice{LQCS_PLAN_DENSE, 0, 0.0, k});
        }
    if (cands.empty()) {
        g_last_choice = Choice{LQCS_PLAN_GREEDY, static_cast<int>(low), -1.0, L};
        return make_plan(LQCS_PLAN_GREEDY, ops, nq, L, low);
    }
    const uint32_t F = LQCS_COST_FEATURES;
    std::vector<double> rows;
    std::vector<size_t> first(cands.size() + 1, 0);
    for (size_t c = 0; c < cands.size(); ++c) {
        first[c] = rows.size() / F;
        for (const PlanStep& st : cands[c].steps) {
            rows.resize(rows.size() + F);
            step_features(cands[c], st, ops, nq, meta[c].L, batch, rows.data() + rows.size() - F);
        }
    }
    first[cands.size()] = rows.size() / F;
    const std::vector<double> ms = cost_model_predict_ms(rows, rows.size() / F);
    size_t best = 0, base = cands.size();
    for (size_t c = 0; c < cands.size(); ++c) {
        double t = 0.0;
        for (size_t r = first[c]; r < first[c + 1]; ++r) t += ms[r];
        meta[c].predicted_ms = t;
        if (t < meta[best].predicted_ms || c == 0) best = c;
        if (meta[c].plan == LQCS_PLAN_GREEDY && meta[c].low == static_cast<int>(low) && meta[c].L == L) base = c;
    }
    // The plan greedy grouping would run stays unless another is predicted
    // to be faster by kLearnedMargin: a near tie is within the model's error.
    if (base < cands.size() && meta[best].predicted_ms > (1.0 - kLearnedMargin) * meta[base].predicted_ms) best = base;
    g_last_choice = meta[best];
    L = meta[best].L;
    if (g_choice_cache.size() >= 16) g_choice_cache.erase(g_choice_cache.begin());
    g_choice_cache.push_back(CachedChoice{key, meta[best]});
    return std::move(cands[best]);
}

void apply_learned(amp_t* s, uint32_t nq, const std::vector<DecodedOp>& ops, uint32_t L, uint32_t low,
                   uint32_t batch) {
    lqcs_exec_options o{LQCS_PLAN_LEARNED, static_cast<int32_t>(L), static_cast<int32_t>(low), 0};
    uint32_t width = 0;
    const Plan plan = choose_plan(&o, ops, nq, batch, width);
    execute_plan(s, nq, width, plan, ops, batch);
}

void last_plan_choice(int* plan, int* low, double* predicted_ms, int* tile_qubits) {
    if (plan) *plan = g_last_choice.plan;
    if (low) *low = g_last_choice.low;
    if (predicted_ms) *predicted_ms = g_last_choice.predicted_ms;
    if (tile_qubits) *tile_qubits = static_cast<int>(g_last_choice.L);
}

}  // namespace lqcs_dev
