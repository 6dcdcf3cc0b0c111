#include "lowering.hpp"

#include <new>
#include <stdexcept>

namespace lqcs::detail {

void check(lqcs_status status) {
    if (status == LQCS_OK) return;
    const std::string msg = lqcs_dev_last_error();
    switch (status) {
        case LQCS_ERR_INVALID_ARGUMENT: throw std::invalid_argument(msg);
        case LQCS_ERR_OUT_OF_RANGE: throw std::out_of_range(msg);
        case LQCS_ERR_LOGIC: throw std::logic_error(msg);
        case LQCS_ERR_OUT_OF_MEMORY: throw std::runtime_error("out of device memory: " + msg);
        case LQCS_ERR_NO_DEVICE: throw std::runtime_error("no usable MI355X device: " + msg);
        default: throw std::runtime_error("device error: " + msg);
    }
}

PauliString parse_pauli(std::string_view pauli, std::size_t n) {
    if (pauli.size() != n)
        throw std::invalid_argument("expectation_value: pauli string length must equal num_qubits");
    PauliString p;
    for (std::size_t j = 0; j < n; ++j) {
        const uint64_t bit = uint64_t{1} << (n - 1 - j);
        switch (pauli[j]) {
            case 'I': break;
            case 'X': p.x |= bit; break;
            case 'Z': p.zy |= bit; break;
            case 'Y': p.x |= bit; p.zy |= bit; ++p.n_y; break;
            default: throw std::invalid_argument("expectation_value: pauli must contain only I/X/Y/Z");
        }
    }
    return p;
}

uint32_t ProgramBuilder::push(uint32_t kind, const std::vector<qubit_t>& ctl, const std::vector<qubit_t>& tgt,
                              const std::vector<complex_t>& data) {
    lqcs_op op{};
    op.kind = kind;
    op.n_controls = static_cast<uint32_t>(ctl.size());
    op.n_targets = static_cast<uint32_t>(tgt.size());
    op.qubit_offset = static_cast<uint32_t>(qubits_.size());
    qubits_.insert(qubits_.end(), ctl.begin(), ctl.end());
    qubits_.insert(qubits_.end(), tgt.begin(), tgt.end());
    op.data_offset = cplx_.size() / 2;
    for (const complex_t& z : data) {
        cplx_.push_back(z.real());
        cplx_.push_back(z.imag());
    }
    ops_.push_back(op);
    return static_cast<uint32_t>(ops_.size() - 1);
}

uint32_t ProgramBuilder::add_pauli_rotation(uint64_t x, uint64_t zy, uint32_t n_y, double theta) {
    lqcs_op op{};
    op.kind = LQCS_OP_PAULI_ROTATION;
    op.qubit_offset = static_cast<uint32_t>(qubits_.size());
    op.x_mask = x;
    op.zy_mask = zy;
    op.n_y = n_y;
    op.theta = theta;
    ops_.push_back(op);
    return static_cast<uint32_t>(ops_.size() - 1);
}

uint32_t ProgramBuilder::add_instruction(const Instruction& inst) {
    const Gate& g = inst.gate;
    if (!g.is_unitary()) throw std::logic_error("lowering: " + g.name() + " is not a unitary op");
    const std::size_t nc = g.n_controls;
    const std::vector<qubit_t> ctl(inst.qubits.begin(), inst.qubits.begin() + nc);
    const std::vector<qubit_t> tgt(inst.qubits.begin() + nc, inst.qubits.end());
    const std::size_t k = g.base_qubits();

    if (g.type == GateType::Permutation) {
        const uint32_t idx = push(LQCS_OP_CONTROLLED_PERMUTATION, ctl, tgt, {});
        ops_[idx].data_offset = perm_.size();
        for (std::size_t v : g.perm) perm_.push_back(v);
        return idx;
    }
    if (g.type == GateType::PauliRotation && nc == 0) {
        const Gate::PauliMasks pm = g.pauli_masks();
        uint64_t x = 0, zy = 0;
        for (std::size_t j = 0; j < k; ++j) {
            if (pm.x & (std::size_t{1} << j)) x |= uint64_t{1} << tgt[j];
            if (pm.zy & (std::size_t{1} << j)) zy |= uint64_t{1} << tgt[j];
        }
        return add_pauli_rotation(x, zy, pm.n_y, g.params.at(0));
    }
    const std::vector<complex_t> m = g.base_matrix();
    if (nc == 0 && k == 1) {
        if (g.is_diagonal()) return push(LQCS_OP_DIAGONAL_SINGLE, {}, tgt, {m[0], m[3]});
        return push(LQCS_OP_SINGLE, {}, tgt, m);
    }
    if (nc == 0 && g.type == GateType::SWAP) return push(LQCS_OP_SWAP, {}, tgt, {});
    if (nc == 0 && k == 2) return push(LQCS_OP_TWO_QUBIT, {}, tgt, m);
    if (nc == 1 && k == 1) return push(LQCS_OP_CONTROLLED_SINGLE, ctl, tgt, m);
    return push(LQCS_OP_CONTROLLED_UNITARY, ctl, tgt, m);
}

uint32_t ProgramBuilder::add_instruction_conj(const Instruction& inst, qubit_t shift) {
    const Gate& g = inst.gate;
    if (!g.is_unitary()) throw std::logic_error("lowering: " + g.name() + " is not a unitary op");
    const std::size_t nc = g.n_controls;
    std::vector<qubit_t> q(inst.qubits);
    for (qubit_t& v : q) v += shift;
    const std::vector<qubit_t> ctl(q.begin(), q.begin() + nc);
    const std::vector<qubit_t> tgt(q.begin() + nc, q.end());
    const std::size_t k = g.base_qubits();

    if (g.type == GateType::Permutation) {  // real 0/1 matrix: its own conjugate
        const uint32_t idx = push(LQCS_OP_CONTROLLED_PERMUTATION, ctl, tgt, {});
        ops_[idx].data_offset = perm_.size();
        for (std::size_t v : g.perm) perm_.push_back(v);
        return idx;
    }
    if (g.type == GateType::PauliRotation && nc == 0) {
        // conj(exp(-i t/2 P)) = exp(i t/2 conj(P)) and conj(P) = (-1)^{#Y} P.
        const Gate::PauliMasks pm = g.pauli_masks();
        uint64_t x = 0, zy = 0;
        for (std::size_t j = 0; j < k; ++j) {
            if (pm.x & (std::size_t{1} << j)) x |= uint64_t{1} << tgt[j];
            if (pm.zy & (std::size_t{1} << j)) zy |= uint64_t{1} << tgt[j];
        }
        const double t = g.params.at(0);
        return add_pauli_rotation(x, zy, pm.n_y, (pm.n_y % 2) ? t : -t);
    }
    std::vector<complex_t> m = g.base_matrix();
    for (complex_t& z : m) z = std::conj(z);
    if (nc == 0 && k == 1) {
        if (g.is_diagonal()) return push(LQCS_OP_DIAGONAL_SINGLE, {}, tgt, {m[0], m[3]});
        return push(LQCS_OP_SINGLE, {}, tgt, m);
    }
    if (nc == 0 && g.type == GateType::SWAP) return push(LQCS_OP_SWAP, {}, tgt, {});
    if (nc == 0 && k == 2) return push(LQCS_OP_TWO_QUBIT, {}, tgt, m);
    if (nc == 1 && k == 1) return push(LQCS_OP_CONTROLLED_SINGLE, ctl, tgt, m);
    return push(LQCS_OP_CONTROLLED_UNITARY, ctl, tgt, m);
}

uint32_t ProgramBuilder::add_two_qubit(qubit_t q0, qubit_t q1, const std::vector<complex_t>& m) {
    if (m.size() != 16) throw std::invalid_argument("lowering: a two-qubit op needs 16 entries");
    return push(LQCS_OP_TWO_QUBIT, {}, {q0, q1}, m);
}

lqcs_program ProgramBuilder::program() const {
    lqcs_program p{};
    p.ops = ops_.data();
    p.n_ops = ops_.size();
    p.qubits = qubits_.data();
    p.n_qubits = qubits_.size();
    p.cplx = cplx_.data();
    p.n_cplx = cplx_.size() / 2;
    p.perm = perm_.data();
    p.n_perm = perm_.size();
    return p;
}

PauliString generator_of(const Instruction& inst) {
    const Gate& g = inst.gate;
    if (g.n_controls != 0) throw std::invalid_argument("adjoint: controlled parametric gates are not supported");
    PauliString p;
    switch (g.type) {
        case GateType::RX: p.x = uint64_t{1} << inst.qubits[0]; break;
        case GateType::RY:
            p.x = p.zy = uint64_t{1} << inst.qubits[0];
            p.n_y = 1;
            break;
        case GateType::RZ: p.zy = uint64_t{1} << inst.qubits[0]; break;
        case GateType::PauliRotation: {
            const Gate::PauliMasks pm = g.pauli_masks();
            for (std::size_t j = 0; j < inst.qubits.size(); ++j) {
                if (pm.x & (std::size_t{1} << j)) p.x |= uint64_t{1} << inst.qubits[j];
                if (pm.zy & (std::size_t{1} << j)) p.zy |= uint64_t{1} << inst.qubits[j];
            }
            p.n_y = pm.n_y;
            break;
        }
        default: throw std::invalid_argument("adjoint: parameter gate must be RX, RY, RZ or PauliRotation");
    }
    return p;
}

}  // namespace lqcs::detail
