// One circuit instruction (reference include/lqcs/circuit/instruction.hpp:9-13).
#pragma once

#include <vector>

#include "gate.hpp"

namespace lqcs {

struct Instruction {
    Gate gate;
    std::vector<qubit_t> qubits;  // controls first, then targets
    std::vector<clbit_t> clbits;  // Measure only
};

}  // namespace lqcs
