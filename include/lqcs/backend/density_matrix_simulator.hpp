// Density-matrix simulator on the device: after every unitary instruction each
// channel of the noise model acts on each qubit the instruction touched
// (reference include/lqcs/backend/density_matrix_simulator.hpp:14-33).
// Terminal measurements only; memory O(4^n).
#pragma once

#include <cstdint>
#include <optional>

#include "backend.hpp"
#include "density_matrix.hpp"
#include "engine.hpp"
#include "noise.hpp"

namespace lqcs {

class DensityMatrixSimulator final : public Backend {
public:
    struct Options {
        std::optional<std::uint64_t> seed;
        NoiseModel noise;      // empty = ideal evolution
        EngineOptions engine;  // engine knobs (appended after the reference's fields)
        // Width cap: the reference's 13 by default, up to DensityMatrix::kMaxDeviceQubits.
        std::size_t max_qubits = DensityMatrix::kMaxQubits;
    };

    DensityMatrixSimulator() = default;
    explicit DensityMatrixSimulator(Options options) : options_(std::move(options)) {}

    Result run(const QuantumCircuit& circuit, std::size_t shots = 1024) override;
    std::string name() const override { return "density_matrix_simulator"; }

    // Final rho (Measure/Barrier skipped, noise applied as in run()).
    DensityMatrix run_density_matrix(const QuantumCircuit& circuit) const;

    const Options& options() const { return options_; }

private:
    Options options_;
};

}  // namespace lqcs
