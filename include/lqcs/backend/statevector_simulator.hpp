// Statevector simulator on the MI355X engine (reference
// include/lqcs/backend/statevector_simulator.hpp:17-42):
//  - terminal measurement: evolve once in HBM, sample shots on the device;
//  - mid-circuit measurement / reset / noise: per-shot trajectories, many
//    shots resident at once.
#pragma once

#include <cstdint>
#include <optional>
#include <vector>

#include "backend.hpp"
#include "engine.hpp"
#include "noise.hpp"
#include "statevector.hpp"

namespace lqcs {

class StatevectorSimulator final : public Backend {
public:
    struct Options {
        std::optional<std::uint64_t> seed;
        bool fuse_gates = true;
        int num_threads = 0;   // accepted for compatibility; the engine runs on the GPU
        NoiseModel noise;
        EngineOptions engine;  // engine knobs (appended after the reference's fields)
    };

    StatevectorSimulator() = default;
    explicit StatevectorSimulator(Options options) : options_(std::move(options)) {}

    Result run(const QuantumCircuit& circuit, std::size_t shots = 1024) override;
    std::string name() const override { return "statevector_simulator"; }

    Statevector run_statevector(const QuantumCircuit& circuit) const;

    // Engine extension: the basis indices the terminal-sampling path draws
    // (one per shot, before they become bitstrings).
    std::vector<std::uint64_t> sample_indices(const QuantumCircuit& circuit, std::size_t shots) const;

    const Options& options() const { return options_; }

private:
    QuantumCircuit maybe_fuse(const QuantumCircuit& circuit) const;

    Options options_;
};

}  // namespace lqcs
