// Engine knobs of the MI355X statevector engine (no counterpart in the
// reference; defaults reproduce its behaviour).
#pragma once

#include <cstddef>
#include <cstdint>

namespace lqcs {

struct EngineOptions {
    // Greedy: first-fit tile sets in program order. Gain: tile sets grown by
    // marginal gain. Learned: the installed cost model's pick among Greedy
    // and Gain plans at 2..4 low qubits (Greedy when no model is installed).
    // Dense: ops multiplied into unitaries over tile_qubits (1..6) qubits, one pass each.
    enum class Plan : int { GateByGate = 0, Greedy = 1, Learned = 2, Auto = 3, Gain = 4, Dense = 5 };
    Plan plan = Plan::Auto;
    int tile_qubits = 0;            // LDS tile width L (0 = engine default)
    int low_qubits = 0;             // contiguous low qubits per tile (0 = default)
    std::size_t max_qubits = 30;    // register cap; raise up to 34 on the device
    std::uint32_t max_batch = 0;    // trajectories resident at once (0 = by free HBM)
    // Noisy-trajectory kernel: 0 auto (persistent where the register fits one
    // workgroup's LDS), 1 shared tiled runs over a batch of registers in HBM,
    // 2 persistent (one workgroup per shot, at most 12 qubits).
    int trajectory_kernel = 0;
    // Tiled-pass kernel: 0 engine default, 1 LDS tile, 2 register tile,
    // 3 register tile streaming, 4 register tile with an LDS landing buffer,
    // 5 register tile specialized per pass structure (compiled at run time on
    // first use and cached for the process).
    int tile_kernel = 0;
    // Remove adjacent inverse pairs (RedundantGateElimination) before fusion
    // and planning. Off by default: the reference's simulator applies fusion only.
    bool eliminate_redundant = false;
};

}  // namespace lqcs
