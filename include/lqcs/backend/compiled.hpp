// Engine extension (no counterpart in the reference): a circuit lowered once
// to the device op stream and applied to resident registers many times.
#pragma once

#include <cstddef>
#include <cstdint>
#include <memory>
#include <vector>

#include "../circuit/quantum_circuit.hpp"
#include "engine.hpp"
#include "statevector.hpp"

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

    std::size_t num_qubits() const { return nq_; }
    std::size_t num_ops() const;

private:
    std::shared_ptr<detail::ProgramBuilder> prog_;
    std::size_t nq_;
};

// Waits for all engine work.
void synchronize();

struct LaunchProfile {
    std::uint64_t tile_launches = 0;
    double tile_ms = 0.0;
    std::uint64_t other_launches = 0;
    double other_ms = 0.0;
};
void profile_enable(bool on);
LaunchProfile profile_read();

}  // namespace lqcs
