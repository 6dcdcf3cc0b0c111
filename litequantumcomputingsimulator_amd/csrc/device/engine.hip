// Engine context, the pass planner/executor and the C-ABI of lqcs_device.h.
//
// The planner turns the op stream (the reference kernel seam) into HBM
// passes. Greedy grouping walks the pending ops in order with a growing tile
// qubit set S: an op joins the pass when every op before it that shares a
// qubit has joined too, and its non-diagonal support fits in S (|S| <= L).
// Ops that do not fit block their qubits for the rest of the walk, so every
// op taken after a deferred one acts on disjoint qubits and commutes with it.
#include "common.hpp"
#include "engine.hpp"
#include "reduce.hpp"
#include "tile.hpp"

#include <algorithm>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../../include/lqcs_device.h"

namespace lqcs_dev {

// ---- context ------------------------------------------------------------------
namespace {

struct Context {
    bool ready = false;
    int device = 0;
    hipStream_t stream = nullptr;
    int cus = 0;
    std::vector<std::pair<char*, size_t>> blocks;
    size_t used = 0;
};

Context& ctx() {
    static Context c;
    return c;
}

thread_local std::string g_last_error;
}  // namespace

extern bool g_rtile_stream;  // rtile.hip
extern bool g_rtile_land;    // rtile.hip

namespace {
// Register tiles on (default) or off (LDS tiles only), for comparisons.
bool g_use_rtile = true;

// Launch timing state (lqcs_profile_*).
struct Profile {
    struct Launch {
        bool tiled;
        hipEvent_t a, b;
        uint64_t amps;  // amplitudes the launch passes over (batch * 2^n)
    };
    bool on = false;
    std::vector<Launch> launches;  // in launch order
    std::vector<hipEvent_t> pool;
};
Profile& prof() {
    static Profile p;
    return p;
}
hipEvent_t take_event() {
    Profile& p = prof();
    if (!p.pool.empty()) {
        hipEvent_t e = p.pool.back();
        p.pool.pop_back();
        return e;
    }
    hipEvent_t e;
    LQCS_HIP(hipEventCreate(&e));
    return e;
}
thread_local uint64_t g_last_passes = 0, g_last_launches = 0;

// Elapsed milliseconds of every recorded launch, in order; recycles events.
struct Drained {
    bool tiled;
    double ms;
    uint64_t amps;
};

std::vector<Drained> drain_launches() {
    Profile& p = prof();
    std::vector<Drained> out;
    out.reserve(p.launches.size());
    for (const auto& l : p.launches) {
        float t = 0.0f;
        LQCS_HIP(hipEventElapsedTime(&t, l.a, l.b));
        out.push_back({l.tiled, double(t), l.amps});
        p.pool.push_back(l.a);
        p.pool.push_back(l.b);
    }
    p.launches.clear();
    return out;
}

std::mutex& api_mutex() {
    static std::mutex m;
    return m;
}

}  // namespace

void check_hip(hipError_t e, const char* what) {
    if (e == hipSuccess) return;
    const int code = (e == hipErrorOutOfMemory) ? LQCS_ERR_OUT_OF_MEMORY : LQCS_ERR_DEVICE;
    throw Error(code, std::string(what) + ": " + hipGetErrorString(e));
}

hipStream_t stream() { return ctx().stream; }
int num_cus() { return ctx().cus; }

void init_device(int device) {
    Context& c = ctx();
    if (c.ready) {
        if (device != c.device) throw Error(LQCS_ERR_INVALID_ARGUMENT, "engine already bound to another device");
        return;
    }
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0)
        throw Error(LQCS_ERR_NO_DEVICE, "no HIP device visible");
    if (device < 0 || device >= count) throw Error(LQCS_ERR_INVALID_ARGUMENT, "device index out of range");
    LQCS_HIP(hipSetDevice(device));
    hipDeviceProp_t prop;
    LQCS_HIP(hipGetDeviceProperties(&prop, device));
    if (std::string(prop.gcnArchName).rfind("gfx950", 0) != 0)
        throw Error(LQCS_ERR_NO_DEVICE, std::string("engine is built for gfx950, found ") + prop.gcnArchName);
    LQCS_HIP(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking));
    c.cus = prop.multiProcessorCount;
    c.device = device;
    c.ready = true;
}

void ensure_ready() {
    if (!ctx().ready) init_device(0);
    LQCS_HIP(hipSetDevice(ctx().device));
}

thread_local GraphTables* g_graph_tables = nullptr;

void* arena(size_t bytes) {
    Context& c = ctx();
    bytes = (bytes + 255) & ~size_t{255};
    if (c.blocks.empty() || c.used + bytes > c.blocks.back().second) {
        size_t sz = std::max<size_t>(bytes, c.blocks.empty() ? (size_t{64} << 20) : c.blocks.back().second * 2);
        char* p = nullptr;
        LQCS_HIP(hipMalloc(&p, sz));
        c.blocks.emplace_back(p, sz);
        c.used = 0;
    }
    void* p = c.blocks.back().first + c.used;
    c.used += bytes;
    return p;
}

void arena_reset() {
    Context& c = ctx();
    if (c.blocks.size() > 1) {  // fold into one block big enough for next time
        LQCS_HIP(hipStreamSynchronize(c.stream));
        size_t total = 0;
        for (auto& b : c.blocks) {
            total += b.second;
            LQCS_HIP(hipFree(b.first));
        }
        c.blocks.clear();
        char* p = nullptr;
        LQCS_HIP(hipMalloc(&p, total));
        c.blocks.emplace_back(p, total);
    }
    c.used = 0;
}

void sync() { LQCS_HIP(hipStreamSynchronize(stream())); }

// ---- op decoding ------------------------------------------------------------
namespace {

amp_t cplx_at(const lqcs_program& p, uint64_t i) {
    if (i >= p.n_cplx) throw Error(LQCS_ERR_INVALID_ARGUMENT, "op data offset beyond complex pool");
    return make_double2(p.cplx[2 * i], p.cplx[2 * i + 1]);
}

bool is_zero(amp_t a) { return a.x == 0.0 && a.y == 0.0; }

}  // namespace

DecodedOp decode(const lqcs_program& p, uint64_t k, uint32_t nq) {
    const lqcs_op& o = p.ops[k];
    DecodedOp d;
    d.kind = o.kind;
    if (uint64_t(o.qubit_offset) + o.n_controls + o.n_targets > p.n_qubits)
        throw Error(LQCS_ERR_INVALID_ARGUMENT, "op qubits beyond qubit pool");
    d.controls.assign(p.qubits + o.qubit_offset, p.qubits + o.qubit_offset + o.n_controls);
    d.targets.assign(p.qubits + o.qubit_offset + o.n_controls,
                     p.qubits + o.qubit_offset + o.n_controls + o.n_targets);
    uint64_t seen = 0;
    for (uint32_t q : d.controls) {
        if (q >= nq) throw Error(LQCS_ERR_OUT_OF_RANGE, "control qubit out of range");
        if (seen & (uint64_t{1} << q)) throw Error(LQCS_ERR_INVALID_ARGUMENT, "duplicate qubit");
        seen |= uint64_t{1} << q;
        d.cmask |= uint64_t{1} << q;
    }
    for (uint32_t q : d.targets) {
        if (q >= nq) throw Error(LQCS_ERR_OUT_OF_RANGE, "target qubit out of range");
        if (seen & (uint64_t{1} << q)) throw Error(LQCS_ERR_INVALID_ARGUMENT, "duplicate qubit");
        seen |= uint64_t{1} << q;
    }
    d.touched = seen;
    const uint64_t full = (nq >= 64) ? ~uint64_t{0} : ((uint64_t{1} << nq) - 1);
    auto need_targets = [&](uint32_t n) {
        if (d.targets.size() != n) throw Error(LQCS_ERR_INVALID_ARGUMENT, "wrong target count for op kind");
    };
    switch (o.kind) {
        case LQCS_OP_SINGLE:
        case LQCS_OP_CONTROLLED_SINGLE:
            need_targets(1);
            if (o.kind == LQCS_OP_CONTROLLED_SINGLE && d.controls.size() != 1)
                throw Error(LQCS_ERR_INVALID_ARGUMENT, "controlled single op needs one control");
            for (int i = 0; i < 4; ++i) d.m.push_back(cplx_at(p, o.data_offset + i));
            break;
        case LQCS_OP_DIAGONAL_SINGLE:
            need_targets(1);
            for (int i = 0; i < 2; ++i) d.m.push_back(cplx_at(p, o.data_offset + i));
            break;
        case LQCS_OP_SWAP:
            need_targets(2);
            break;
        case LQCS_OP_TWO_QUBIT:
            need_targets(2);
            for (int i = 0; i < 16; ++i) d.m.push_back(cplx_at(p, o.data_offset + i));
            break;
        case LQCS_OP_CONTROLLED_UNITARY: {
            if (d.targets.empty()) throw Error(LQCS_ERR_INVALID_ARGUMENT, "unitary op without targets");
            if (d.targets.size() > 12) throw Error(LQCS_ERR_LOGIC, "dense op wider than 12 qubits");
            const uint64_t D = uint64_t{1} << d.targets.size();
            for (uint64_t i = 0; i < D * D; ++i) d.m.push_back(cplx_at(p, o.data_offset + i));
            break;
        }
        case LQCS_OP_CONTROLLED_PERMUTATION: {
            if (d.targets.empty()) throw Error(LQCS_ERR_INVALID_ARGUMENT, "permutation op without targets");
            if (d.targets.size() > 12) throw Error(LQCS_ERR_LOGIC, "permutation wider than 12 qubits");
            const uint64_t D = uint64_t{1} << d.targets.size();
            if (o.data_offset + D > p.n_perm) throw Error(LQCS_ERR_INVALID_ARGUMENT, "perm beyond pool");
            std::vector<uint8_t> hit(D, 0);
            for (uint64_t i = 0; i < D; ++i) {
                const uint64_t v = p.perm[o.data_offset + i];
                if (v >= D || hit[v]) throw Error(LQCS_ERR_INVALID_ARGUMENT, "permutation is not a bijection");
                hit[v] = 1;
                d.perm.push_back(v);
            }
            break;
        }
        case LQCS_OP_PAULI_ROTATION:
            if (!d.controls.empty() || !d.targets.empty())
                throw Error(LQCS_ERR_INVALID_ARGUMENT, "pauli rotation takes masks, not qubits");
            if ((o.x_mask | o.zy_mask) & ~full) throw Error(LQCS_ERR_OUT_OF_RANGE, "pauli mask out of range");
            d.x = o.x_mask;
            d.zy = o.zy_mask;
            d.n_y = o.n_y;
            d.theta = o.theta;
            d.touched = o.x_mask | o.zy_mask;
            break;
        default:
            throw Error(LQCS_ERR_INVALID_ARGUMENT, "unknown op kind");
    }
    // Diagonal recognition (exact zeros off the diagonal).
    if (o.kind == LQCS_OP_DIAGONAL_SINGLE) d.diagonal = true;
    if ((o.kind == LQCS_OP_SINGLE || o.kind == LQCS_OP_CONTROLLED_SINGLE) && is_zero(d.m[1]) && is_zero(d.m[2]))
        d.diagonal = true;
    if (o.kind == LQCS_OP_TWO_QUBIT || (o.kind == LQCS_OP_CONTROLLED_UNITARY && d.targets.size() <= 2)) {
        const uint64_t D = uint64_t{1} << d.targets.size();
        bool diag = true;
        for (uint64_t r = 0; r < D && diag; ++r)
            for (uint64_t c = 0; c < D; ++c)
                if (r != c && !is_zero(d.m[r * D + c])) { diag = false; break; }
        d.diagonal = diag;
    }
    if (o.kind == LQCS_OP_PAULI_ROTATION && d.x == 0) d.diagonal = true;
    return d;
}

// ---- gate-by-gate execution -------------------------------------------------
void run_single_op(amp_t* s, uint32_t nq, const DecodedOp& d, uint32_t batch) {
    if (d.member >= 0) {  // one batch register, or all but one: contiguous ranges
        DecodedOp all = d;
        all.member = -1;
        const uint32_t m = static_cast<uint32_t>(d.member);
        if (m >= batch) throw Error(LQCS_ERR_INVALID_ARGUMENT, "batch member out of range");
        if (d.only) {
            run_single_op(s + (uint64_t(m) << nq), nq, all, 1);
        } else {
            if (m > 0) run_single_op(s, nq, all, m);
            if (m + 1 < batch) run_single_op(s + (uint64_t(m + 1) << nq), nq, all, batch - m - 1);
        }
        return;
    }
    const uint32_t* t = d.targets.data();
    switch (d.kind) {
        case LQCS_OP_SINGLE: {
            Mat2 m{{d.m[0], d.m[1], d.m[2], d.m[3]}};
            launch_single(s, nq, t[0], m, batch);
            return;
        }
        case LQCS_OP_DIAGONAL_SINGLE:
            launch_diag_single(s, nq, t[0], d.m[0], d.m[1], batch);
            return;
        case LQCS_OP_CONTROLLED_SINGLE: {
            Mat2 m{{d.m[0], d.m[1], d.m[2], d.m[3]}};
            launch_controlled_single(s, nq, d.controls[0], t[0], m, batch);
            return;
        }
        case LQCS_OP_SWAP:
            launch_swap(s, nq, t[0], t[1], batch);
            return;
        case LQCS_OP_TWO_QUBIT: {
            Mat4 m;
            for (int i = 0; i < 16; ++i) m.m[i] = d.m[i];
            launch_two_qubit(s, nq, t[0], t[1], m, batch);
            return;
        }
        case LQCS_OP_CONTROLLED_UNITARY:
            launch_controlled_unitary(s, nq, d.controls.data(), d.controls.size(), t,
                                      d.targets.size(), d.m.data(), batch);
            return;
        case LQCS_OP_CONTROLLED_PERMUTATION:
            launch_controlled_permutation(s, nq, d.controls.data(), d.controls.size(), t,
                                          d.targets.size(), d.perm.data(), batch);
            return;
        case LQCS_OP_PAULI_ROTATION:
            launch_pauli_rotation(s, nq, d.x, d.zy, d.n_y, d.theta, batch);
            return;
        default:
            throw Error(LQCS_ERR_INVALID_ARGUMENT, "unknown op kind");
    }
}

// ---- tiled planning ------------------------------------------------------------
// ---- tile-pass encoding (shared with planner.hip)

// Qubits whose values must vary inside the tile for this op; ~0 when the op
// cannot run inside a tile at all.
uint64_t tile_requirement(const DecodedOp& d) {
    if (d.diagonal && (d.kind != LQCS_OP_CONTROLLED_UNITARY || d.targets.size() <= 2)) return 0;
    switch (d.kind) {
        case LQCS_OP_SINGLE:
        case LQCS_OP_CONTROLLED_SINGLE:
            return uint64_t{1} << d.targets[0];
        case LQCS_OP_SWAP:
        case LQCS_OP_TWO_QUBIT:
            return (uint64_t{1} << d.targets[0]) | (uint64_t{1} << d.targets[1]);
        case LQCS_OP_CONTROLLED_UNITARY:
        case LQCS_OP_CONTROLLED_PERMUTATION:
            if (d.targets.size() > 2) return ~uint64_t{0};
            return d.targets.size() == 1 ? (uint64_t{1} << d.targets[0])
                                         : ((uint64_t{1} << d.targets[0]) | (uint64_t{1} << d.targets[1]));
        case LQCS_OP_PAULI_ROTATION:
            return d.x;
        default:
            return ~uint64_t{0};
    }
}

static uint32_t local_pos(uint64_t smask, uint32_t q) {
    return static_cast<uint32_t>(__builtin_popcountll(smask & ((uint64_t{1} << q) - 1)));
}

static uint32_t to_local_mask(uint64_t smask, uint64_t global) {
    uint32_t m = 0;
    for (uint64_t g = global & smask; g; g &= g - 1) m |= 1u << local_pos(smask, __builtin_ctzll(g));
    return m;
}

TileOp make_tile_op(const DecodedOp& d, uint64_t smask) {
    TileOp t;
    std::memset(&t, 0, sizeof t);
    t.member = d.member;
    t.only = d.only ? 1u : 0u;
    t.lctrl = to_local_mask(smask, d.cmask);
    t.gctrl = d.cmask & ~smask;
    auto set_diag_qubit = [&](int slot, uint32_t q) {
        const uint64_t bit = uint64_t{1} << q;
        if (smask & bit) (slot ? t.dl1 : t.dl0) = 1u << local_pos(smask, q);
        else (slot ? t.dg1 : t.dg0) = bit;
    };
    if (d.kind == LQCS_OP_PAULI_ROTATION) {
        t.kind = TK_PAULI;
        t.xl = to_local_mask(smask, d.x);
        t.zl = to_local_mask(smask, d.zy);
        t.zg = d.zy & ~smask;
        const double h = d.theta / 2;
        static const amp_t kIPow[4] = {{1, 0}, {0, 1}, {-1, 0}, {0, -1}};
        const amp_t ip = kIPow[d.n_y % 4];
        const double ms = -std::sin(h);
        t.m[0] = make_double2(std::cos(h), 0.0);
        t.m[1] = make_double2(-ms * ip.y, ms * ip.x);  // (0, -sin) * i^ny
        return t;
    }
    if (d.diagonal) {
        t.kind = TK_DIAG;
        if (d.kind == LQCS_OP_DIAGONAL_SINGLE) {
            set_diag_qubit(0, d.targets[0]);
            t.m[0] = d.m[0];
            t.m[1] = d.m[1];
        } else if (d.targets.size() == 1) {
            set_diag_qubit(0, d.targets[0]);
            t.m[0] = d.m[0];
            t.m[1] = d.m[3];
        } else {
            set_diag_qubit(0, d.targets[0]);
            set_diag_qubit(1, d.targets[1]);
            for (int i = 0; i < 4; ++i) t.m[i] = d.m[i * 4 + i];
        }
        return t;
    }
    if (d.targets.size() == 1) {
        t.kind = TK_DENSE1;
        t.t0 = local_pos(smask, d.targets[0]);
        if (d.kind == LQCS_OP_CONTROLLED_PERMUTATION) {
            for (int c = 0; c < 2; ++c) t.m[d.perm[c] * 2 + c] = make_double2(1.0, 0.0);
        } else {
            for (int i = 0; i < 4; ++i) t.m[i] = d.m[i];
        }
        return t;
    }
    t.kind = TK_DENSE2;
    t.t0 = local_pos(smask, d.targets[0]);
    t.t1 = local_pos(smask, d.targets[1]);
    if (d.kind == LQCS_OP_SWAP) {
        t.m[0] = t.m[6] = t.m[9] = t.m[15] = make_double2(1.0, 0.0);
    } else if (d.kind == LQCS_OP_CONTROLLED_PERMUTATION) {
        for (int c = 0; c < 4; ++c) t.m[d.perm[c] * 4 + c] = make_double2(1.0, 0.0);
    } else {
        for (int i = 0; i < 16; ++i) t.m[i] = d.m[i];
    }
    return t;
}



Plan plan_greedy(const std::vector<DecodedOp>& ops, uint32_t nq, uint32_t L, uint32_t low) {
    Plan plan;
    std::vector<uint32_t> pending(ops.size());
    for (uint32_t i = 0; i < ops.size(); ++i) pending[i] = i;
    const uint64_t low_mask = (uint64_t{1} << low) - 1;
    while (!pending.empty()) {
        const DecodedOp& head = ops[pending[0]];
        if (tile_requirement(head) == ~uint64_t{0}) {
            PlanStep st;
            st.tiled = false;
            st.op = pending[0];
            plan.steps.push_back(st);
            pending.erase(pending.begin());
            continue;
        }
        uint64_t S = low_mask, blocked = 0;
        std::vector<uint32_t> taken, rest;
        for (uint32_t idx : pending) {
            const DecodedOp& d = ops[idx];
            const uint64_t R = tile_requirement(d);
            if ((d.touched & blocked) || R == ~uint64_t{0}) {
                blocked |= d.touched;
                rest.push_back(idx);
                continue;
            }
            const uint64_t grown = S | R;
            if (static_cast<uint32_t>(__builtin_popcountll(grown)) <= L) {
                S = grown;
                taken.push_back(idx);
            } else {
                blocked |= d.touched;
                rest.push_back(idx);
            }
        }
        // Pad S to exactly L qubits with the lowest unused ones.
        for (uint32_t q = 0; q < nq && static_cast<uint32_t>(__builtin_popcountll(S)) < L; ++q)
            S |= uint64_t{1} << q;
        PlanStep st;
        st.tiled = true;
        st.smask = S;
        for (uint32_t idx : taken) st.tile_ops.push_back(make_tile_op(ops[idx], S));
        plan.steps.push_back(std::move(st));
        pending.swap(rest);
    }
    return plan;
}

StagedPlan stage_plan(const Plan& plan, uint32_t nq, uint32_t L, uint32_t batch) {
    // Register-tile programs where possible, LDS tile programs otherwise;
    // specialized kernels for the register-tile programs when selected.
    StagedPlan sp;
    std::vector<std::string> spec_src;
    const bool nt = (uint64_t(batch) << nq) * sizeof(amp_t) >= (uint64_t{1} << 30);
    for (const auto& st : plan.steps) {
        spec_src.emplace_back();
        if (!st.tiled) {
            sp.where.push_back({0, 0, 0, 0});
            continue;
        }
        RTileProgram rp;
        if (g_use_rtile && build_rtile(st.tile_ops, L, st.smask, nq, rp)) {
            if (g_rtile_spec && !g_rtile_stream && !g_rtile_land && L >= 12)
                spec_src.back() = spec_source(rp, nq, L, st.smask, nt);
            if (L > 12 && spec_src.back().empty())
                throw Error(LQCS_ERR_LOGIC, "13-qubit tile pass without a specialized kernel");
            sp.where.push_back({2, sp.rph.size(), rp.phases.size(), sp.rwords.size()});
            sp.rph.insert(sp.rph.end(), rp.phases.begin(), rp.phases.end());
            sp.rwords.insert(sp.rwords.end(), rp.words.begin(), rp.words.end());
        } else {
            if (L > 12)
                throw Error(LQCS_ERR_INVALID_ARGUMENT,
                            "13-qubit tiles: an op acts densely on more than 4 tile qubits (a wide Pauli rotation); "
                            "use tiles of at most 12 qubits for this circuit");
            sp.where.push_back({1, sp.flat.size(), st.tile_ops.size(), 0});
            sp.flat.insert(sp.flat.end(), st.tile_ops.begin(), st.tile_ops.end());
        }
    }
    if (g_rtile_spec) sp.spec = spec_kernels(spec_src);
    return sp;
}

void launch_staged(amp_t* s, uint32_t nq, uint32_t L, const Plan& plan, const std::vector<DecodedOp>& ops,
                   uint32_t batch, const StagedPlan& sp, const TileOp* d_ops, const RPhase* d_ph,
                   const amp_t* d_rwords, bool timed) {
    for (size_t i = 0; i < plan.steps.size(); ++i) {
        const PlanStep& st = plan.steps[i];
        hipEvent_t e0 = nullptr, e1 = nullptr;
        if (timed) {
            e0 = take_event();
            e1 = take_event();
            LQCS_HIP(hipEventRecord(e0, stream()));
        }
        const StagedPlan::Where& w = sp.where[i];
        if (w.kind == 2 && i < sp.spec.size() && sp.spec[i]) {
            launch_spec(sp.spec[i], s, nq, L, d_rwords + w.words, batch);
        } else if (w.kind == 2) {
            launch_rtile(s, nq, L, st.smask, d_ph + w.first, static_cast<uint32_t>(w.count),
                         d_rwords ? d_rwords + w.words : nullptr, batch);
        } else if (w.kind == 1) {
            launch_tile_pass(s, nq, L, st.smask, d_ops + w.first, static_cast<uint32_t>(w.count), batch);
        } else {
            run_single_op(s, nq, step_op(plan, ops, st), batch);
        }
        if (timed) {
            LQCS_HIP(hipEventRecord(e1, stream()));
            prof().launches.push_back({st.tiled, e0, e1, uint64_t(batch) << nq});
        }
    }
}

void execute_plan(amp_t* s, uint32_t nq, uint32_t L, const Plan& plan,
                  const std::vector<DecodedOp>& ops, uint32_t batch) {
    const StagedPlan sp = stage_plan(plan, nq, L, batch);
    const TileOp* d_ops = sp.flat.empty() ? nullptr : stage(sp.flat.data(), sp.flat.size());
    const RPhase* d_ph = sp.rph.empty() ? nullptr : stage(sp.rph.data(), sp.rph.size());
    const amp_t* d_rwords = sp.rwords.empty() ? nullptr : stage(sp.rwords.data(), sp.rwords.size());
    launch_staged(s, nq, L, plan, ops, batch, sp, d_ops, d_ph, d_rwords, prof().on);
    g_last_passes = plan.steps.size();
    g_last_launches = plan.steps.size();
}

TileChoice default_tile(uint32_t nq) {
    TileChoice c;
    if (nq < 6) return c;  // tiles unused: gate-by-gate
    c.L = std::min<uint32_t>(12, nq <= 12 ? nq : std::max<uint32_t>(6, nq - 4));
    c.low = std::min<uint32_t>(4, c.L - 2);
    return c;
}

void apply_ops(amp_t* s, uint32_t nq, const std::vector<DecodedOp>& ops, int plan_kind,
               uint32_t L, uint32_t low, uint32_t batch) {
    if (plan_kind == LQCS_PLAN_AUTO) plan_kind = LQCS_PLAN_GREEDY;
    const TileChoice def = default_tile(nq);
    if (L == 0) L = def.L;
    if (low == 0) low = def.low;
    if (plan_kind == LQCS_PLAN_GATE_BY_GATE || L == 0) {
        for (const auto& d : ops) {
            const bool timed = prof().on;
            hipEvent_t e0 = nullptr, e1 = nullptr;
            if (timed) {
                e0 = take_event();
                e1 = take_event();
                LQCS_HIP(hipEventRecord(e0, stream()));
            }
            run_single_op(s, nq, d, batch);
            if (timed) {
                LQCS_HIP(hipEventRecord(e1, stream()));
                prof().launches.push_back({false, e0, e1, uint64_t(batch) << nq});
            }
        }
        g_last_passes = ops.size();
        g_last_launches = ops.size();
        return;
    }
    const Plan plan = make_plan(plan_kind, ops, nq, L, low);
    execute_plan(s, nq, L, plan, ops, batch);
}

bool rtile_enabled() { return g_use_rtile; }

// The tile kernel of one call (lqcs_exec_options.tile_kernel), restored on exit.
class KernelScope {
public:
    explicit KernelScope(const lqcs_exec_options* o)
        : use_(g_use_rtile), stream_(g_rtile_stream), land_(g_rtile_land), spec_(g_rtile_spec) {
        const int k = o ? o->tile_kernel : LQCS_TILE_DEFAULT;
        if (k == LQCS_TILE_DEFAULT) return;
        if (k < LQCS_TILE_LDS || k > LQCS_TILE_SPECIALIZED)
            throw Error(LQCS_ERR_INVALID_ARGUMENT, "unknown tile kernel");
        g_use_rtile = k != LQCS_TILE_LDS;
        g_rtile_stream = k == LQCS_TILE_STREAM;
        g_rtile_land = k == LQCS_TILE_LAND;
        g_rtile_spec = k == LQCS_TILE_SPECIALIZED;
    }
    ~KernelScope() {
        g_use_rtile = use_;
        g_rtile_stream = stream_;
        g_rtile_land = land_;
        g_rtile_spec = spec_;
    }
    KernelScope(const KernelScope&) = delete;
    KernelScope& operator=(const KernelScope&) = delete;

private:
    bool use_, stream_, land_, spec_;
};

Plan make_plan(int kind, const std::vector<DecodedOp>& ops, uint32_t nq, uint32_t L, uint32_t low) {
    if (kind == LQCS_PLAN_DENSE) return plan_dense(ops, nq, L);  // L: the block width
    Plan plan;
    if (kind == LQCS_PLAN_GATE_BY_GATE || L == 0) {
        for (uint32_t i = 0; i < ops.size(); ++i) {
            PlanStep st;
            st.op = i;
            plan.steps.push_back(st);
        }
        return plan;
    }
    const uint32_t max_l = g_rtile_spec ? 13u : 12u;  // 13-qubit tiles run on specialized passes only
    if (L < 6 || L > max_l || L > nq)
        throw Error(LQCS_ERR_INVALID_ARGUMENT, g_rtile_spec ? "tile width must be 6..13 and <= n"
                                                            : "tile width must be 6..12 and <= n (13 with specialized passes)");
    if (low + 2 > L) throw Error(LQCS_ERR_INVALID_ARGUMENT, "low qubits must leave two free tile slots");
    return kind == LQCS_PLAN_GAIN ? plan_gain(ops, nq, L, low) : plan_greedy(ops, nq, L, low);
}

void set_last_stats(uint64_t passes, uint64_t launches) {
    g_last_passes = passes;
    g_last_launches = launches;
}

}  // namespace lqcs_dev

// ================================================================================
// C-ABI
// ================================================================================
using namespace lqcs_dev;

namespace {

template <class F>
lqcs_status guarded(F&& f) {
    std::lock_guard<std::mutex> lock(api_mutex());
    try {
        f();
        return LQCS_OK;
    } catch (const Error& e) {
        g_last_error = e.what();
        return e.code;
    } catch (const std::bad_alloc&) {
        g_last_error = "host allocation failed";
        return LQCS_ERR_OUT_OF_MEMORY;
    } catch (const std::exception& e) {
        g_last_error = e.what();
        return LQCS_ERR_LOGIC;
    }
}

}  // namespace

namespace lqcs_dev {
lqcs_status run_guarded(const std::function<void()>& f) { return guarded(f); }
void* device_alloc(size_t bytes) {
    void* p = nullptr;
    if (hipMalloc(&p, bytes) != hipSuccess)
        throw Error(LQCS_ERR_OUT_OF_MEMORY, "device allocation of " + std::to_string(bytes) + " bytes failed");
    return p;
}
void device_free(void* p) {
    if (p) (void)hipFree(p);
}
}  // namespace lqcs_dev

namespace {

void need(const void* p, const char* what) {
    if (!p) throw Error(LQCS_ERR_INVALID_ARGUMENT, std::string(what) + " is null");
}

uint64_t amps_of(uint32_t nq) { return uint64_t{1} << nq; }

}  // namespace

extern "C" {

const char* lqcs_dev_last_error(void) { return g_last_error.c_str(); }

lqcs_status lqcs_dev_init(int device) {
    return guarded([&] { init_device(device); });
}

lqcs_status lqcs_dev_info(char* name, size_t name_len, int* cus, uint64_t* hbm_bytes) {
    return guarded([&] {
        ensure_ready();
        hipDeviceProp_t prop;
        LQCS_HIP(hipGetDeviceProperties(&prop, ctx().device));
        if (name && name_len) {
            std::snprintf(name, name_len, "%s (%s)", prop.name, prop.gcnArchName);
        }
        if (cus) *cus = prop.multiProcessorCount;
        if (hbm_bytes) *hbm_bytes = prop.totalGlobalMem;
    });
}

lqcs_status lqcs_dev_synchronize(void) {
    return guarded([&] {
        ensure_ready();
        sync();
        LQCS_HIP(hipDeviceSynchronize());
    });
}

lqcs_status lqcs_set_tile_kernel(int kernel) {
    return guarded([&] {
        if (kernel < 0 || kernel > 4)
            throw Error(LQCS_ERR_INVALID_ARGUMENT,
                        "tile kernel must be 0 (LDS), 1 (register), 2 (register, streaming), 3 (register, landing) "
                        "or 4 (register, specialized per pass)");
        g_use_rtile = kernel >= 1;
        g_rtile_stream = kernel == 2;
        g_rtile_land = kernel == 3;
        g_rtile_spec = kernel == 4;
    });
}

lqcs_status lqcs_set_dense_mfma_min_targets(uint32_t k) {
    return guarded([&] {
        if (k < 4) throw Error(LQCS_ERR_INVALID_ARGUMENT, "dense MFMA kernel needs at least 4 targets");
        g_dense_mfma_min_targets = k;
    });
}

lqcs_status lqcs_profile_enable(int enable) {
    return guarded([&] {
        ensure_ready();
        prof().on = enable != 0;
    });
}

lqcs_status lqcs_profile_read(uint64_t* tile_launches, double* tile_ms, uint64_t* other_launches,
                              double* other_ms, uint64_t* tile_amps) {
    return guarded([&] {
        ensure_ready();
        sync();
        uint64_t n[2] = {0, 0}, amps = 0;
        double ms[2] = {0.0, 0.0};
        for (const Drained& d : drain_launches()) {
            ++n[d.tiled ? 1 : 0];
            ms[d.tiled ? 1 : 0] += d.ms;
            if (d.tiled) amps += d.amps;
        }
        if (tile_launches) *tile_launches = n[1];
        if (tile_ms) *tile_ms = ms[1];
        if (other_launches) *other_launches = n[0];
        if (other_ms) *other_ms = ms[0];
        if (tile_amps) *tile_amps = amps;
    });
}

lqcs_status lqcs_profile_read_launches(double* ms, uint8_t* tiled, uint64_t max_launches, uint64_t* n_launches) {
    return guarded([&] {
        need(n_launches, "n_launches");
        ensure_ready();
        sync();
        const auto all = drain_launches();
        *n_launches = all.size();
        for (uint64_t i = 0; i < all.size() && i < max_launches; ++i) {
            if (ms) ms[i] = all[i].ms;
            if (tiled) tiled[i] = all[i].tiled ? 1 : 0;
        }
    });
}

lqcs_status lqcs_state_create(uint32_t n_qubits, lqcs_dev_state** out) {
    return guarded([&] {
        need(out, "out");
        *out = nullptr;
        if (n_qubits == 0) throw Error(LQCS_ERR_INVALID_ARGUMENT, "Statevector: num_qubits must be >= 1");
        if (n_qubits > LQCS_DEVICE_MAX_QUBITS)
            throw Error(LQCS_ERR_INVALID_ARGUMENT, "Statevector: " + std::to_string(n_qubits) +
                                                       " qubits exceed the device engine limit");
        ensure_ready();
        arena_reset();
        auto* st = new lqcs_dev_state;
        st->nq = n_qubits;
        const hipError_t e = hipMalloc(&st->d, amps_of(n_qubits) * sizeof(amp_t));
        if (e != hipSuccess) {
            delete st;
            throw Error(LQCS_ERR_OUT_OF_MEMORY, "Statevector: device allocation of " +
                                                    std::to_string(amps_of(n_qubits) * 16) + " bytes failed");
        }
        launch_fill_basis(st->d, n_qubits, 0);
        *out = st;
    });
}

lqcs_status lqcs_state_destroy(lqcs_dev_state* s) {
    return guarded([&] {
        if (!s) return;
        sync();
        LQCS_HIP(hipFree(s->d));
        delete s;
    });
}

lqcs_status lqcs_state_num_qubits(const lqcs_dev_state* s, uint32_t* n_qubits) {
    return guarded([&] {
        need(s, "state");
        need(n_qubits, "n_qubits");
        *n_qubits = s->nq;
    });
}

lqcs_status lqcs_state_set_basis(lqcs_dev_state* s, uint64_t index) {
    return guarded([&] {
        need(s, "state");
        if (index >= amps_of(s->nq)) throw Error(LQCS_ERR_OUT_OF_RANGE, "basis index out of range");
        ensure_ready();
        launch_fill_basis(s->d, s->nq, index);
    });
}

lqcs_status lqcs_state_upload(lqcs_dev_state* s, const double* amps) {
    return guarded([&] {
        need(s, "state");
        need(amps, "amps");
        ensure_ready();
        LQCS_HIP(hipMemcpyAsync(s->d, amps, amps_of(s->nq) * sizeof(amp_t), hipMemcpyHostToDevice, stream()));
        sync();
    });
}

lqcs_status lqcs_state_download(const lqcs_dev_state* s, double* amps) {
    return guarded([&] {
        need(s, "state");
        need(amps, "amps");
        ensure_ready();
        LQCS_HIP(hipMemcpyAsync(amps, s->d, amps_of(s->nq) * sizeof(amp_t), hipMemcpyDeviceToHost, stream()));
        sync();
    });
}

lqcs_status lqcs_state_copy(lqcs_dev_state* dst, const lqcs_dev_state* src) {
    return guarded([&] {
        need(dst, "dst");
        need(src, "src");
        if (dst->nq != src->nq) throw Error(LQCS_ERR_INVALID_ARGUMENT, "state widths differ");
        ensure_ready();
        LQCS_HIP(hipMemcpyAsync(dst->d, src->d, amps_of(src->nq) * sizeof(amp_t), hipMemcpyDeviceToDevice,
                                stream()));
    });
}

lqcs_status lqcs_state_device_ptr(lqcs_dev_state* s, void** ptr) {
    return guarded([&] {
        need(s, "state");
        need(ptr, "ptr");
        *ptr = s->d;
    });
}

lqcs_status lqcs_state_apply(lqcs_dev_state* s, const lqcs_program* prog, const lqcs_exec_options* opts) {
    return guarded([&] {
        need(s, "state");
        need(prog, "program");
        if (prog->n_ops && !prog->ops) throw Error(LQCS_ERR_INVALID_ARGUMENT, "ops is null");
        ensure_ready();
        arena_reset();
        std::vector<DecodedOp> ops;
        ops.reserve(prog->n_ops);
        for (uint64_t k = 0; k < prog->n_ops; ++k) ops.push_back(decode(*prog, k, s->nq));
        const KernelScope ks(opts);
        uint32_t L = 0;
        const Plan plan = choose_plan(opts, ops, s->nq, 1, L);
        execute_plan(s->d, s->nq, L, plan, ops, 1);
    });
}

struct lqcs_graph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    void* buf = nullptr;     // the tile programs the captured kernels read
    void* tables = nullptr;  // the stand-alone steps' tables (matrices, offsets, permutations)
    const lqcs_dev::amp_t* d = nullptr;  // the register buffer the kernels address
    uint32_t nq = 0;
};

lqcs_status lqcs_graph_create(lqcs_dev_state* s, const lqcs_program* prog, const lqcs_exec_options* opts,
                              lqcs_graph** out) {
    return guarded([&] {
        need(s, "state");
        need(prog, "program");
        need(out, "out");
        *out = nullptr;
        if (prog->n_ops && !prog->ops) throw Error(LQCS_ERR_INVALID_ARGUMENT, "ops is null");
        ensure_ready();
        std::vector<DecodedOp> ops;
        for (uint64_t k = 0; k < prog->n_ops; ++k) ops.push_back(decode(*prog, k, s->nq));
        const KernelScope ks(opts);
        uint32_t L = 0;
        const Plan plan = choose_plan(opts, ops, s->nq, 1, L);
        bool has_tables = false;  // stand-alone steps that stage tables per call
        for (const PlanStep& st : plan.steps)
            has_tables |= !st.tiled && (step_op(plan, ops, st).kind == LQCS_OP_CONTROLLED_UNITARY ||
                                        step_op(plan, ops, st).kind == LQCS_OP_CONTROLLED_PERMUTATION);
        const StagedPlan sp = stage_plan(plan, s->nq, L);
        auto pad = [](size_t b) { return (b + 255) & ~size_t{255}; };
        const size_t b_ops = pad(sp.flat.size() * sizeof(TileOp)), b_ph = pad(sp.rph.size() * sizeof(RPhase));
        const size_t b_w = pad(sp.rwords.size() * sizeof(amp_t));
        auto* g = new lqcs_graph;
        try {
            g->d = s->d;
            g->nq = s->nq;
            g->buf = device_alloc(std::max<size_t>(b_ops + b_ph + b_w, 256));
            char* base = static_cast<char*>(g->buf);
            auto* d_ops = reinterpret_cast<TileOp*>(base);
            auto* d_ph = reinterpret_cast<RPhase*>(base + b_ops);
            auto* d_w = reinterpret_cast<amp_t*>(base + b_ops + b_ph);
            sync();
            if (!sp.flat.empty()) LQCS_HIP(hipMemcpy(d_ops, sp.flat.data(), sp.flat.size() * sizeof(TileOp), hipMemcpyHostToDevice));
            if (!sp.rph.empty()) LQCS_HIP(hipMemcpy(d_ph, sp.rph.data(), sp.rph.size() * sizeof(RPhase), hipMemcpyHostToDevice));
            if (!sp.rwords.empty()) LQCS_HIP(hipMemcpy(d_w, sp.rwords.data(), sp.rwords.size() * sizeof(amp_t), hipMemcpyHostToDevice));
            // One capture of the launch sequence; with `tables`, stage() lays the
            // stand-alone steps' tables out in graph-owned memory.
            auto capture = [&](GraphTables* tables, hipGraph_t* out_graph) {
                struct Scope {
                    explicit Scope(GraphTables* t) { g_graph_tables = t; }
                    ~Scope() { g_graph_tables = nullptr; }
                } scope(tables);
                LQCS_HIP(hipStreamBeginCapture(stream(), hipStreamCaptureModeThreadLocal));
                try {
                    launch_staged(s->d, s->nq, L, plan, ops, 1, sp, sp.flat.empty() ? nullptr : d_ops,
                                  sp.rph.empty() ? nullptr : d_ph, sp.rwords.empty() ? nullptr : d_w, false);
                } catch (...) {
                    hipGraph_t dead = nullptr;
                    (void)hipStreamEndCapture(stream(), &dead);
                    if (dead) (void)hipGraphDestroy(dead);
                    throw;
                }
                LQCS_HIP(hipStreamEndCapture(stream(), out_graph));
            };
            if (!has_tables) {
                capture(nullptr, &g->graph);
            } else {
                // A sizing capture (never launched) measures the tables; the
                // second capture addresses them in the graph's memory, filled once
                // the capture has ended.
                GraphTables tables;
                hipGraph_t sizing = nullptr;
                capture(&tables, &sizing);
                if (sizing) LQCS_HIP(hipGraphDestroy(sizing));
                const size_t bytes = tables.host.size();
                g->tables = device_alloc(std::max<size_t>(bytes, 256));
                tables.host.clear();
                tables.dev = static_cast<char*>(g->tables);
                capture(&tables, &g->graph);
                if (tables.host.size() != bytes)
                    throw Error(LQCS_ERR_LOGIC, "graph capture: the tables changed between the two captures");
                LQCS_HIP(hipMemcpy(g->tables, tables.host.data(), bytes, hipMemcpyHostToDevice));
            }
            LQCS_HIP(hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0));
        } catch (...) {
            if (g->exec) (void)hipGraphExecDestroy(g->exec);
            if (g->graph) (void)hipGraphDestroy(g->graph);
            if (g->buf) device_free(g->buf);
            if (g->tables) device_free(g->tables);
            delete g;
            throw;
        }
        *out = g;
    });
}

lqcs_status lqcs_graph_launch(lqcs_graph* g, lqcs_dev_state* s) {
    return guarded([&] {
        need(g, "graph");
        need(s, "state");
        if (s->d != g->d || s->nq != g->nq)
            throw Error(LQCS_ERR_LOGIC, "graph launch: the register does not hold the buffer the graph was captured on");
        ensure_ready();
        LQCS_HIP(hipGraphLaunch(g->exec, stream()));
    });
}

lqcs_status lqcs_graph_destroy(lqcs_graph* g) {
    return guarded([&] {
        if (!g) return;
        sync();
        if (g->exec) LQCS_HIP(hipGraphExecDestroy(g->exec));
        if (g->graph) LQCS_HIP(hipGraphDestroy(g->graph));
        if (g->buf) device_free(g->buf);
        if (g->tables) device_free(g->tables);
        delete g;
    });
}

lqcs_status lqcs_plan_describe(uint32_t n_qubits, const lqcs_program* prog, const lqcs_exec_options* opts,
                               uint64_t* n_steps, uint64_t* ops_per_step, uint64_t* tile_masks,
                               uint64_t* phases_per_step, uint64_t max_steps) {
    return guarded([&] {
        need(prog, "program");
        need(n_steps, "n_steps");
        if (n_qubits == 0 || n_qubits > LQCS_DEVICE_MAX_QUBITS) throw Error(LQCS_ERR_INVALID_ARGUMENT, "bad width");
        std::vector<DecodedOp> ops;
        for (uint64_t k = 0; k < prog->n_ops; ++k) ops.push_back(decode(*prog, k, n_qubits));
        const KernelScope ks(opts);
        uint32_t L = 0;
        const Plan plan = choose_plan(opts, ops, n_qubits, 1, L);
        *n_steps = plan.steps.size();
        for (uint64_t i = 0; i < plan.steps.size() && i < max_steps; ++i) {
            const PlanStep& st = plan.steps[i];
            if (ops_per_step) ops_per_step[i] = st.tiled ? st.tile_ops.size() : st.fused_ops ? st.fused_ops : 1;
            if (tile_masks) tile_masks[i] = st.tiled ? st.smask : st.fused_ops ? step_op(plan, ops, st).touched : 0;
            if (phases_per_step) {
                RTileProgram rp;
                phases_per_step[i] = (plan.steps[i].tiled && g_use_rtile &&
                                      build_rtile(plan.steps[i].tile_ops, L, plan.steps[i].smask, n_qubits, rp))
                                         ? rp.phases.size()
                                         : 0;
            }
        }
    });
}

lqcs_status lqcs_specialized_kernels(uint64_t* cached) {
    return guarded([&] {
        need(cached, "cached");
        *cached = spec_cache_size();
    });
}

lqcs_status lqcs_plan_specialized_pass(uint32_t n_qubits, const lqcs_program* prog, const lqcs_exec_options* opts,
                                       uint64_t step, char* text, size_t text_max, size_t* text_len, double* words,
                                       uint64_t words_max, uint64_t* n_words) {
    return guarded([&] {
        need(prog, "program");
        need(text_len, "text_len");
        need(n_words, "n_words");
        if (n_qubits == 0 || n_qubits > LQCS_DEVICE_MAX_QUBITS) throw Error(LQCS_ERR_INVALID_ARGUMENT, "bad width");
        std::vector<DecodedOp> ops;
        for (uint64_t k = 0; k < prog->n_ops; ++k) ops.push_back(decode(*prog, k, n_qubits));
        const KernelScope ks(opts);
        uint32_t L = 0;
        const Plan plan = choose_plan(opts, ops, n_qubits, 1, L);
        if (step >= plan.steps.size()) throw Error(LQCS_ERR_OUT_OF_RANGE, "step beyond the plan");
        const PlanStep& st = plan.steps[step];
        std::string src;
        RTileProgram rp;
        if (st.tiled && g_use_rtile && L >= 12 && build_rtile(st.tile_ops, L, st.smask, n_qubits, rp)) {
            const bool nt = (uint64_t{1} << n_qubits) * sizeof(amp_t) >= (uint64_t{1} << 30);
            src = spec_source(rp, n_qubits, L, st.smask, nt);
        }
        if (src.empty()) rp.words.clear();
        *text_len = src.size();
        *n_words = rp.words.size();
        if (text && text_max) {
            const size_t n = std::min(text_max - 1, src.size());
            std::memcpy(text, src.data(), n);
            text[n] = '\0';
        }
        if (words && words_max >= rp.words.size())
            std::memcpy(words, rp.words.data(), rp.words.size() * sizeof(amp_t));
    });
}

lqcs_status lqcs_plan_features(uint32_t n_qubits, const lqcs_program* prog, const lqcs_exec_options* opts,
                               uint32_t batch, double* features, uint64_t max_steps, uint64_t* n_steps) {
    return guarded([&] {
        need(prog, "program");
        need(n_steps, "n_steps");
        if (n_qubits == 0 || n_qubits > LQCS_DEVICE_MAX_QUBITS) throw Error(LQCS_ERR_INVALID_ARGUMENT, "bad width");
        std::vector<DecodedOp> ops;
        for (uint64_t k = 0; k < prog->n_ops; ++k) ops.push_back(decode(*prog, k, n_qubits));
        const KernelScope ks(opts);
        uint32_t L = 0;
        const Plan plan = choose_plan(opts, ops, n_qubits, std::max<uint32_t>(batch, 1), L);
        *n_steps = plan.steps.size();
        if (!features) return;
        for (uint64_t i = 0; i < plan.steps.size() && i < max_steps; ++i)
            step_features(plan, plan.steps[i], ops, n_qubits, L, std::max<uint32_t>(batch, 1),
                          features + i * LQCS_COST_FEATURES);
    });
}

lqcs_status lqcs_last_plan_choice(int32_t* plan, int32_t* low_qubits, double* predicted_ms) {
    int p = 0, l = 0;
    double ms = 0.0;
    last_plan_choice(&p, &l, &ms);
    if (plan) *plan = p;
    if (low_qubits) *low_qubits = l;
    if (predicted_ms) *predicted_ms = ms;
    return LQCS_OK;
}

lqcs_status lqcs_last_plan_tile_qubits(int32_t* tile_qubits) {
    int L = 0;
    last_plan_choice(nullptr, nullptr, nullptr, &L);
    if (tile_qubits) *tile_qubits = L;
    return LQCS_OK;
}

lqcs_status lqcs_last_plan_stats(uint64_t* passes, uint64_t* launches) {
    if (passes) *passes = g_last_passes;
    if (launches) *launches = g_last_launches;
    return LQCS_OK;
}

lqcs_status lqcs_state_norm_sq(const lqcs_dev_state* s, double* out) {
    return guarded([&] {
        need(s, "state");
        need(out, "out");
        ensure_ready();
        arena_reset();
        double* d = static_cast<double*>(arena(sizeof(double)));
        reduce_norm_sq(s->d, s->nq, 1, d);
        LQCS_HIP(hipMemcpyAsync(out, d, sizeof(double), hipMemcpyDeviceToHost, stream()));
        sync();
    });
}

lqcs_status lqcs_state_max_abs_diff(const lqcs_dev_state* a, const lqcs_dev_state* b, double* out) {
    return guarded([&] {
        need(a, "a");
        need(b, "b");
        need(out, "out");
        if (a->nq != b->nq) throw Error(LQCS_ERR_INVALID_ARGUMENT, "registers differ in width");
        ensure_ready();
        arena_reset();
        double* d = static_cast<double*>(arena(sizeof(double)));
        reduce_max_abs_diff(a->d, b->d, a->nq, d);
        LQCS_HIP(hipMemcpyAsync(out, d, sizeof(double), hipMemcpyDeviceToHost, stream()));
        sync();
    });
}

lqcs_status lqcs_state_probabilities(const lqcs_dev_state* s, double* out) {
    return guarded([&] {
        need(s, "state");
        need(out, "out");
        ensure_ready();
        arena_reset();
        double* d = static_cast<double*>(arena(sizeof(double) * amps_of(s->nq)));
        launch_probabilities(s->d, s->nq, d);
        LQCS_HIP(hipMemcpyAsync(out, d, sizeof(double) * amps_of(s->nq), hipMemcpyDeviceToHost, stream()));
        sync();
    });
}

lqcs_status lqcs_state_read(const lqcs_dev_state* s, uint64_t first, uint64_t count, double* out) {
    return guarded([&] {
        need(s, "state");
        need(out, "out");
        if (first > amps_of(s->nq) || count > amps_of(s->nq) - first)
            throw Error(LQCS_ERR_OUT_OF_RANGE, "read: amplitude range out of range");
        ensure_ready();
        if (count)
            LQCS_HIP(hipMemcpyAsync(out, s->d + first, count * sizeof(amp_t), hipMemcpyDeviceToHost, stream()));
        sync();
    });
}

lqcs_status lqcs_dm_diagonal(const lqcs_dev_state* rho, double* out) {
    return guarded([&] {
        need(rho, "rho");
        need(out, "out");
        if (rho->nq % 2) throw Error(LQCS_ERR_INVALID_ARGUMENT, "density matrix: register width must be even");
        const uint32_t n = rho->nq / 2;
        ensure_ready();
        arena_reset();
        double* d = static_cast<double*>(arena(sizeof(double) * amps_of(n)));
        launch_dm_diagonal(rho->d, n, d);
        LQCS_HIP(hipMemcpyAsync(out, d, sizeof(double) * amps_of(n), hipMemcpyDeviceToHost, stream()));
        sync();
    });
}

lqcs_status lqcs_dm_fidelity(const lqcs_dev_state* rho, const lqcs_dev_state* psi, double* out) {
    return guarded([&] {
        need(rho, "rho");
        need(psi, "psi");
        need(out, "out");
        if (rho->nq != 2 * psi->nq) throw Error(LQCS_ERR_INVALID_ARGUMENT, "fidelity: qubit count mismatch");
        ensure_ready();
        arena_reset();
        double* d = static_cast<double*>(arena(sizeof(double)));
        reduce_dm_fidelity(rho->d, psi->d, psi->nq, d);
        LQCS_HIP(hipMemcpyAsync(out, d, sizeof(double), hipMemcpyDeviceToHost, stream()));
        sync();
    });
}

lqcs_status lqcs_state_expectation_pauli(const lqcs_dev_state* s, uint64_t x_mask, uint64_t zy_mask,
                                         uint32_t n_y, double* out) {
    const double one = 1.0;
    return lqcs_state_expectation_terms(s, &x_mask, &zy_mask, &n_y, &one, 1, out);
}

lqcs_status lqcs_state_expectation_terms(const lqcs_dev_state* s, const uint64_t* x_masks,
                                         const uint64_t* zy_masks, const uint32_t* n_ys,
                                         const double* coeffs, uint64_t n_terms, double* out) {
    return guarded([&] {
        need(s, "state");
        need(out, "out");
        *out = 0.0;
        if (n_terms == 0) return;
        need(x_masks, "x_masks");
        need(zy_masks, "zy_masks");
        need(n_ys, "n_ys");
        need(coeffs, "coeffs");
        const uint64_t full = amps_of(s->nq) - 1;
        for (uint64_t t = 0; t < n_terms; ++t)
            if ((x_masks[t] | zy_masks[t]) & ~full) throw Error(LQCS_ERR_OUT_OF_RANGE, "pauli mask out of range");
        ensure_ready();
        arena_reset();
        std::vector<double> vals(2 * n_terms);
        constexpr uint64_t kChunk = 60000;  // grid.y limit
        for (uint64_t t0 = 0; t0 < n_terms; t0 += kChunk) {
            const uint32_t cnt = static_cast<uint32_t>(std::min(kChunk, n_terms - t0));
            const uint64_t* dx = stage(x_masks + t0, cnt);
            const uint64_t* dz = stage(zy_masks + t0, cnt);
            double* d = static_cast<double*>(arena(sizeof(double) * 2 * cnt));
            reduce_pauli_terms(s->d, s->nq, dx, dz, cnt, d);
            LQCS_HIP(hipMemcpyAsync(vals.data() + 2 * t0, d, sizeof(double) * 2 * cnt, hipMemcpyDeviceToHost,
                                    stream()));
        }
        sync();
        // Combine as algorithms::expectation does: e += coeff * Re(i^ny * sum).
        double e = 0.0;
        for (uint64_t t = 0; t < n_terms; ++t) {
            const double re = vals[2 * t], im = vals[2 * t + 1];
            double v;
            switch (n_ys[t] % 4) {
                case 0: v = re; break;
                case 1: v = -im; break;
                case 2: v = -re; break;
                default: v = im; break;
            }
            e += coeffs[t] * v;
        }
        *out = e;
    });
}

lqcs_status lqcs_state_sample(const lqcs_dev_state* s, const double* draws, uint64_t n_draws,
                              uint64_t* out_indices) {
    return guarded([&] {
        need(s, "state");
        if (n_draws == 0) return;
        need(draws, "draws");
        need(out_indices, "out_indices");
        ensure_ready();
        arena_reset();
        const double* dd = stage(draws, n_draws);
        uint64_t* di = static_cast<uint64_t*>(arena(sizeof(uint64_t) * n_draws));
        constexpr uint64_t kMaxGrid = 1u << 30;
        for (uint64_t o = 0; o < n_draws; o += kMaxGrid) {
            const uint64_t cnt = std::min(kMaxGrid, n_draws - o);
            launch_sample(s->d, s->nq, dd + o, cnt, di + o);
        }
        LQCS_HIP(hipMemcpyAsync(out_indices, di, sizeof(uint64_t) * n_draws, hipMemcpyDeviceToHost, stream()));
        sync();
    });
}

lqcs_status lqcs_state_collapse(lqcs_dev_state* s, uint32_t qubit, double u, int* outcome) {
    return guarded([&] {
        need(s, "state");
        if (qubit >= s->nq) throw Error(LQCS_ERR_OUT_OF_RANGE, "qubit out of range");
        ensure_ready();
        arena_reset();
        double* p1 = static_cast<double*>(arena(sizeof(double)));
        const double* du = stage(&u, 1);
        uint8_t* creg = static_cast<uint8_t*>(arena(1));
        reduce_p1(s->d, s->nq, qubit, 1, p1);
        launch_collapse(s->d, s->nq, qubit, p1, du, 1, 0, 0, creg, 1, 0, 1);
        uint8_t h = 0;
        LQCS_HIP(hipMemcpyAsync(&h, creg, 1, hipMemcpyDeviceToHost, stream()));
        sync();
        if (outcome) *outcome = h;
    });
}

lqcs_status lqcs_state_reset_qubit(lqcs_dev_state* s, uint32_t qubit, double u) {
    return guarded([&] {
        need(s, "state");
        if (qubit >= s->nq) throw Error(LQCS_ERR_OUT_OF_RANGE, "qubit out of range");
        ensure_ready();
        arena_reset();
        double* p1 = static_cast<double*>(arena(sizeof(double)));
        const double* du = stage(&u, 1);
        reduce_p1(s->d, s->nq, qubit, 1, p1);
        launch_collapse(s->d, s->nq, qubit, p1, du, 1, 0, 1, nullptr, 0, -1, 1);
    });
}

lqcs_status lqcs_state_kraus_probability(const lqcs_dev_state* s, uint32_t qubit, const double* k,
                                         double* out) {
    return guarded([&] {
        need(s, "state");
        need(k, "k");
        need(out, "out");
        if (qubit >= s->nq) throw Error(LQCS_ERR_OUT_OF_RANGE, "qubit out of range");
        ensure_ready();
        arena_reset();
        double* d = static_cast<double*>(arena(4 * sizeof(double)));
        reduce_qstats(s->d, s->nq, qubit, 1, d);
        double st[4];
        LQCS_HIP(hipMemcpyAsync(st, d, sizeof st, hipMemcpyDeviceToHost, stream()));
        sync();
        *out = kraus_branch_probability(k, st);
    });
}

}  // extern "C"
