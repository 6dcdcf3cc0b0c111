// Pass pipeline (reference include/lqcs/transpiler/transpiler.hpp:11-31):
// passes run in the order they were added.
#pragma once

#include <memory>
#include <vector>

#include "pass.hpp"

namespace lqcs {

class Transpiler {
public:
    Transpiler& add_pass(std::unique_ptr<Pass> pass) {
        passes_.push_back(std::move(pass));
        return *this;
    }

    QuantumCircuit run(const QuantumCircuit& circuit) const {
        QuantumCircuit out = circuit;
        for (const auto& p : passes_) out = p->run(out);
        return out;
    }

    std::size_t num_passes() const { return passes_.size(); }

private:
    std::vector<std::unique_ptr<Pass>> passes_;
};

}  // namespace lqcs
