// Circuit pass interface (reference include/lqcs/transpiler/pass.hpp:10-15).
#pragma once

#include <string>

#include "../circuit/quantum_circuit.hpp"

namespace lqcs {

class Pass {
public:
    virtual ~Pass() = default;
    virtual QuantumCircuit run(const QuantumCircuit& circuit) const = 0;
    virtual std::string name() const = 0;
};

}  // namespace lqcs
