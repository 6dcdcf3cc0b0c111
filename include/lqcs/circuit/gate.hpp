// Gate model. Same enumeration order, fields and member functions as the
// reference's include/lqcs/circuit/gate.hpp:14-68 (callers and the C-ABI
// numbering depend on them).
#pragma once

#include <cstddef>
#include <string>
#include <vector>

#include "../core/types.hpp"

namespace lqcs {

enum class GateType {
    I, X, Y, Z, H, S, Sdg, T, Tdg, SX,
    RX, RY, RZ, P, U,
    SWAP, iSWAP, RXX, RYY, RZZ,
    PauliRotation,
    Unitary, Permutation,
    Measure, Reset, Barrier,
};

struct Gate {
    GateType type;
    std::vector<double> params;
    std::size_t n_controls;
    std::vector<complex_t> mat;     // Unitary: row-major base matrix
    std::vector<std::size_t> perm;  // Permutation: |l> -> |perm[l]>
    std::string pauli;              // PauliRotation: rightmost char <-> first target

    Gate(GateType t, std::vector<double> p = {}, std::size_t nc = 0, std::vector<complex_t> m = {},
         std::vector<std::size_t> pm = {}, std::string pl = {})
        : type(t), params(std::move(p)), n_controls(nc), mat(std::move(m)), perm(std::move(pm)),
          pauli(std::move(pl)) {}

    struct PauliMasks {
        std::size_t x = 0;
        std::size_t zy = 0;
        unsigned n_y = 0;
    };
    PauliMasks pauli_masks() const;

    std::size_t base_qubits() const;
    std::size_t num_qubits() const { return n_controls + base_qubits(); }
    bool is_unitary() const;
    bool is_diagonal() const;

    std::vector<complex_t> base_matrix() const;  // 2^k x 2^k, no controls
    std::vector<complex_t> matrix() const;       // with controls on the low local bits

    Gate inverse() const;
    std::string name() const;
};

}  // namespace lqcs
