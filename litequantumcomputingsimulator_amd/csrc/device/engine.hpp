// Engine internals shared by the C-ABI translation units.
#pragma once

#include "common.hpp"
#include "tile.hpp"

#include <functional>
#include <string>
#include <vector>

#include "../../../include/lqcs_device.h"

// A device register: 2^nq complex128 amplitudes owned by the handle.
struct lqcs_dev_state {
    lqcs_dev::amp_t* d = nullptr;
    uint32_t nq = 0;
};

namespace lqcs_dev {

void init_device(int device);
void ensure_ready();
void sync();

// One op of the stream with its tables copied out of the program pools.
struct DecodedOp {
    uint32_t kind = 0;
    std::vector<uint32_t> controls, targets;
    std::vector<amp_t> m;        // matrix (row-major) or d0,d1
    std::vector<uint64_t> perm;
    uint64_t cmask = 0;          // control bits
    uint64_t touched = 0;        // every qubit the op reads or writes
    uint64_t x = 0, zy = 0;      // Pauli rotation
    uint32_t n_y = 0;
    double theta = 0.0;
    bool diagonal = false;       // exact zeros off the diagonal
    // Batch-member restriction (trajectories): member >= 0 with only: that
    // register alone; without: every register of the batch but that one.
    int32_t member = -1;
    bool only = false;
};

DecodedOp decode(const lqcs_program& p, uint64_t k, uint32_t nq);
void run_single_op(amp_t* s, uint32_t nq, const DecodedOp& d, uint32_t batch);

struct PlanStep {
    bool tiled = false;
    uint64_t smask = 0;            // tile qubits (tiled)
    std::vector<TileOp> tile_ops;  // tiled
    uint32_t op = 0;               // stand-alone op index; from ops.size() on, Plan::fused
    uint32_t fused_ops = 0;        // ops multiplied into that fused op (dense blocks)
};
struct Plan {
    std::vector<PlanStep> steps;
    std::vector<DecodedOp> fused;  // ops the plan made (dense blocks)
};
// The op a stand-alone step runs.
inline const DecodedOp& step_op(const Plan& p, const std::vector<DecodedOp>& ops, const PlanStep& st) {
    return st.op < ops.size() ? ops[st.op] : p.fused[st.op - ops.size()];
}

struct TileChoice {
    uint32_t L = 0;    // 0: no tiling
    uint32_t low = 0;
};
TileChoice default_tile(uint32_t nq);

// Qubits that must vary inside a tile for this op (~0: op cannot be tiled).
uint64_t tile_requirement(const DecodedOp& d);
TileOp make_tile_op(const DecodedOp& d, uint64_t smask);
Plan plan_gain(const std::vector<DecodedOp>& ops, uint32_t nq, uint32_t L, uint32_t low);
Plan plan_greedy(const std::vector<DecodedOp>& ops, uint32_t nq, uint32_t L, uint32_t low);
Plan plan_dense(const std::vector<DecodedOp>& ops, uint32_t nq, uint32_t k);
void execute_plan(amp_t* s, uint32_t nq, uint32_t L, const Plan& plan,
                  const std::vector<DecodedOp>& ops, uint32_t batch);
// A plan's tile programs, ready to copy to the device: LDS-tile ops, and
// register-tile phases and op words, with each step's place in them.
struct StagedPlan {
    struct Where {
        int kind;  // 0 stand-alone, 1 LDS tile, 2 register tile
        size_t first, count, words;
    };
    std::vector<TileOp> flat;
    std::vector<RPhase> rph;
    std::vector<amp_t> rwords;
    std::vector<Where> where;
    // Per step: the specialized kernel of its register-tile program (null:
    // the interpreting kernel runs it).
    std::vector<hipFunction_t> spec;
};
// `batch` registers are passed over per launch (it picks non-temporal
// accesses for specialized passes over 1 GiB or more).
StagedPlan stage_plan(const Plan& plan, uint32_t nq, uint32_t L, uint32_t batch = 1);
// Launches every step of a staged plan whose programs are at d_ops / d_ph /
// d_rwords on the device; `timed` brackets each launch with profile events.
void launch_staged(amp_t* s, uint32_t nq, uint32_t L, const Plan& plan, const std::vector<DecodedOp>& ops,
                   uint32_t batch, const StagedPlan& sp, const TileOp* d_ops, const RPhase* d_ph,
                   const amp_t* d_rwords, bool timed);
void apply_ops(amp_t* s, uint32_t nq, const std::vector<DecodedOp>& ops, int plan_kind,
               uint32_t L, uint32_t low, uint32_t batch);
// Cost-model guided planning (planner.hip); falls back to greedy.
void apply_learned(amp_t* s, uint32_t nq, const std::vector<DecodedOp>& ops, uint32_t L,
                   uint32_t low, uint32_t batch);
// Plan for GATE_BY_GATE / GREEDY / GAIN at a fixed tile width and low count.
Plan make_plan(int kind, const std::vector<DecodedOp>& ops, uint32_t nq, uint32_t L, uint32_t low);
// The plan lqcs_state_apply runs for these options (LEARNED: the cost
// model's pick among greedy and gain plans at 2..4 low qubits -- 3..5 on 12-
// and 13-qubit tiles for specialized passes -- or greedy without a model); L
// receives the tile width (0: gate by gate).
Plan choose_plan(const lqcs_exec_options* opts, const std::vector<DecodedOp>& ops, uint32_t nq, uint32_t batch,
                 uint32_t& L);
// LQCS_COST_FEATURES values describing one planned step (planner.hip).
void step_features(const Plan& plan, const PlanStep& st, const std::vector<DecodedOp>& ops, uint32_t nq, uint32_t L, uint32_t batch,
                   double* f);
void last_plan_choice(int* plan, int* low, double* predicted_ms, int* tile_qubits = nullptr);
// Specialized register-tile passes (spec_pass.hip): on for the calls that
// select them; kernel text of a program on L = 12 or 13 tile qubits (empty:
// a form the generator does not handle); the kernels of a list of sources,
// compiled on first use; their launch (2^(L-4) threads per tile).
extern bool g_rtile_spec;
std::string spec_source(const RTileProgram& rp, uint32_t nq, uint32_t L, uint64_t smask, bool nt);
std::vector<hipFunction_t> spec_kernels(const std::vector<std::string>& sources);
void launch_spec(hipFunction_t f, amp_t* s, uint32_t nq, uint32_t L, const amp_t* d_words, uint32_t batch);
size_t spec_cache_size();
bool rtile_enabled();
// Installed cost model (mlp.hip).
bool cost_model_installed();
uint64_t cost_model_version();
std::vector<double> cost_model_predict_ms(const std::vector<double>& raw_features, uint64_t rows);
void set_last_stats(uint64_t passes, uint64_t launches);
// Runs f under the engine lock, mapping exceptions to an lqcs_status and the
// thread's last-error text.
lqcs_status run_guarded(const std::function<void()>& f);
void* device_alloc(size_t bytes);  // hipMalloc with OOM mapped to LQCS_ERR_OUT_OF_MEMORY
void device_free(void* p);

// ||K psi||^2 on one qubit from its 2x2 statistics (S00, S11, Re S01, Im S01)
// where S01 = sum conj(a0) a1; k = 2x2 row-major complex (8 doubles).
inline double kraus_branch_probability(const double* k, const double* st) {
    const double k0r = k[0], k0i = k[1], k1r = k[2], k1i = k[3];
    const double k2r = k[4], k2i = k[5], k3r = k[6], k3i = k[7];
    const double c00 = k0r * k0r + k0i * k0i + k2r * k2r + k2i * k2i;
    const double c11 = k1r * k1r + k1i * k1i + k3r * k3r + k3i * k3i;
    // conj(k0) k1 + conj(k2) k3
    const double xr = k0r * k1r + k0i * k1i + k2r * k3r + k2i * k3i;
    const double xi = k0r * k1i - k0i * k1r + k2r * k3i - k2i * k3r;
    return c00 * st[0] + c11 * st[1] + 2.0 * (xr * st[2] - xi * st[3]);
}

}  // namespace lqcs_dev
