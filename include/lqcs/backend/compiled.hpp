// Engine extension (no counterpart in the reference): a circuit lowered once
// to the device op stream and applied to resident registers many times.
#pragma once

#include <cstddef>
#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "../circuit/quantum_circuit.hpp"
#include "engine.hpp"
#include "statevector.hpp"

struct lqcs_graph;

namespace lqcs {

namespace detail {
class ProgramBuilder;
}

class CompiledCircuit {
public:
    // Lowers the unitary part of `circuit` (after greedy single-qubit fusion
    // when fuse_gates is set). Measure and Barrier are skipped; Reset throws.
    explicit CompiledCircuit(const QuantumCircuit& circuit, bool fuse_gates = true);
    ~CompiledCircuit();

    // sv <- U sv, in place on the device.
    void apply(Statevector& sv, const EngineOptions& engine = {}) const;

    // Plan the engine would run (no device needed): per HBM pass, the ops it
    // applies, its tile qubit mask (0 = stand-alone op) and register phases.
    struct PassInfo {
        std::uint64_t ops, tile_mask, phases;
    };
    std::vector<PassInfo> describe(const EngineOptions& engine = {}) const;

    // Specialized kernel of pass `step` of that plan (no device needed): the
    // kernel text the engine compiles for it and the op words it reads, as
    // (re, im) pairs; empty text when the pass runs on another kernel.
    struct SpecializedPass {
        std::string text;
        std::vector<double> words;
    };
    SpecializedPass specialized_pass(std::uint64_t step, const EngineOptions& engine = {}) const;

    // Cost-model features of every pass of that plan (kCostFeatures each,
    // row-major), for `batch` registers per launch.
    std::vector<double> features(const EngineOptions& engine = {}, std::uint32_t batch = 1) const;

    std::size_t num_qubits() const { return nq_; }
    std::size_t num_ops() const;

    // Autotuner: times every candidate tile shape on `scratch` (a register of
    // the circuit's width, overwritten) -- greedy and gain grouping at 2, 3
    // and 4 low qubits on the register tile, the other tile kernels at the
    // default shape, dense blocks of 4, 5 and 6 qubits (up to the register's
    // width), and from 12 qubits specialized passes (greedy at 4 low qubits,
    // gain at 3 on 12- and, from 17 qubits, 13-qubit tiles; a 13-qubit
    // candidate the circuit's ops refuse is skipped) -- keeping the fastest; apply() with Plan::Auto then
    // uses it. Each candidate: one untimed apply, then the minimum of `reps`
    // timed ones (kernel time from HIP events).
    struct TuneTrial {
        EngineOptions options;
        double ms;
    };
    struct TuneResult {
        EngineOptions best;
        double best_ms = 0.0;
        std::vector<TuneTrial> trials;
    };
    TuneResult autotune(Statevector& scratch, int reps = 2);
    // The tuned options (Plan::Auto when not tuned).
    EngineOptions tuned() const;

private:
    friend class CircuitGraph;
    EngineOptions resolve(const EngineOptions& e) const;
    std::shared_ptr<detail::ProgramBuilder> prog_;
    std::size_t nq_;
    std::shared_ptr<EngineOptions> tuned_;
};

// A compiled circuit's launches on one register, recorded once into a HIP
// graph and replayed: launch() applies U to the register again (queued on
// the engine stream; synchronize() waits). The Statevector object must
// outlive it; once the object holds another device register (assignment from
// another Statevector), launch() throws std::logic_error.
class CircuitGraph {
public:
    CircuitGraph(const CompiledCircuit& circuit, Statevector& sv, const EngineOptions& engine = {});
    ~CircuitGraph();
    CircuitGraph(const CircuitGraph&) = delete;
    CircuitGraph& operator=(const CircuitGraph&) = delete;
    void launch();

private:
    lqcs_graph* graph_ = nullptr;
    Statevector* sv_;
};

// Waits for all engine work.
void synchronize();

struct LaunchProfile {
    std::uint64_t tile_launches = 0;
    double tile_ms = 0.0;
    std::uint64_t other_launches = 0;
    double other_ms = 0.0;
    std::uint64_t tile_amps = 0;  // amplitudes passed over by the tiled passes
};
void profile_enable(bool on);
LaunchProfile profile_read();
// Per-launch milliseconds (and whether each was a tiled pass) since the last
// read, in launch order.
struct LaunchTimes {
    std::vector<double> ms;
    std::vector<std::uint8_t> tiled;
};
LaunchTimes profile_read_launches();

// ---- learned planning cost model (engine extension) ----
constexpr std::uint32_t kCostFeatures = 24;
struct CostModel {
    std::vector<double> params;   // packed MLP parameters (mlp_num_params(kCostFeatures))
    std::vector<double> x_mean, x_scale;
    double t_mean = 0.0, t_scale = 1.0;
};
std::uint64_t mlp_num_params(std::uint32_t n_features);
// Installs (or, with empty params, removes) the model Plan::Learned consults
// for passes on the tile kernel EngineOptions::tile_kernel names: the
// interpreting register tile (any value but 5) or specialized passes (5).
void install_cost_model(const CostModel& m, int tile_kernel = 0);
// Reads a model written by litequantumcomputingsimulator_amd.costmodel:
// "lqcs-costmodel 1 <F> <n_params>" then x_mean, x_scale, t_mean, t_scale
// and the parameters, whitespace separated.
CostModel load_cost_model(const std::string& path);
struct PlanChoice {
    EngineOptions::Plan plan;
    int low_qubits;
    double predicted_ms;  // -1 when no model decided
    int tile_qubits;      // tile width of the plan, or dense block width (0: gate by gate)
};
PlanChoice last_plan_choice();

}  // namespace lqcs
