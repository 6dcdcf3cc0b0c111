// Density matrix rho (2^n x 2^n), |0...0><0...0| at construction
// (reference include/lqcs/backend/density_matrix.hpp:17-44).
//
// rho lives in HBM as a 2n-qubit register: element <r|rho|c> at index
// r | (c << n), i.e. the ket index on the low n qubits and the bra index on
// the high n. A unitary acts as U on the ket qubits and conj(U) on the bra
// qubits; a single-qubit Kraus channel acts as its 4x4 superoperator
// sum_i K_i (x) conj(K_i) on the qubit pair (q, q + n). Everything runs on the
// same device kernels as Statevector.
#pragma once

#include <cstddef>
#include <vector>

#include "../circuit/instruction.hpp"
#include "../core/types.hpp"
#include "noise.hpp"
#include "statevector.hpp"

struct lqcs_dev_state;

namespace lqcs {

class DensityMatrix {
public:
    static constexpr std::size_t kMaxQubits = 13;  // the reference's cap: 4^13 x 16 B = 1 GiB
    static constexpr std::size_t kMaxDeviceQubits = 16;  // 4^16 x 16 B = 64 GiB of HBM

    explicit DensityMatrix(std::size_t num_qubits);
    // Engine extension: a matrix wider than kMaxQubits (up to kMaxDeviceQubits).
    DensityMatrix(std::size_t num_qubits, std::size_t max_qubits);
    ~DensityMatrix();
    DensityMatrix(const DensityMatrix& other);
    DensityMatrix& operator=(const DensityMatrix& other);
    DensityMatrix(DensityMatrix&& other) noexcept;
    DensityMatrix& operator=(DensityMatrix&& other) noexcept;

    std::size_t num_qubits() const { return num_qubits_; }
    std::size_t dim() const { return std::size_t{1} << num_qubits_; }

    // <row|rho|col>
    complex_t element(std::size_t row, std::size_t col) const;

    // rho -> U rho U^dag for any unitary instruction (controlled, permutation).
    void apply_instruction(const Instruction& inst);
    // rho -> sum_i K_i rho K_i^dag on qubit q.
    void apply_channel(const KrausChannel& channel, qubit_t q);

    double trace() const;
    double purity() const;                      // Tr(rho^2)
    std::vector<double> probabilities() const;  // the diagonal
    double fidelity(const Statevector& psi) const;  // <psi|rho|psi>

    // Engine access: the 2n-qubit device register.
    lqcs_dev_state* device() const { return dev_; }

private:
    std::size_t num_qubits_;
    lqcs_dev_state* dev_ = nullptr;
};

}  // namespace lqcs
