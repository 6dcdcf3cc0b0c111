// Circuit IR, parametric circuits, results and noise channels: the host-side
// data model of the lqcs API. Validation rules and exception types follow the
// reference (src/circuit/quantum_circuit.cpp, parametric_circuit.cpp,
// src/result/result.cpp, src/backend/noise.cpp) because callers test them.
#include <algorithm>
#include <cmath>
#include <sstream>
#include <stdexcept>
#include <string>

#include "lqcs/backend/noise.hpp"
#include "lqcs/circuit/parametric_circuit.hpp"
#include "lqcs/circuit/quantum_circuit.hpp"
#include "lqcs/result/result.hpp"

namespace lqcs {
namespace {

// U U^dag = I to 1e-10 (quantum_circuit.cpp:15-31)
void check_unitary(std::span<const complex_t> m, std::size_t dim) {
    for (std::size_t r = 0; r < dim; ++r)
        for (std::size_t c = 0; c < dim; ++c) {
            complex_t acc{0, 0};
            for (std::size_t l = 0; l < dim; ++l) acc += m[r * dim + l] * std::conj(m[c * dim + l]);
            if (std::abs(acc - complex_t{r == c ? 1.0 : 0.0, 0.0}) > 1e-10)
                throw std::invalid_argument("unitary: matrix is not unitary (U*Udag != I)");
        }
}

void check_pauli(std::string_view p, const char* who) {
    for (char ch : p)
        if (ch != 'I' && ch != 'X' && ch != 'Y' && ch != 'Z')
            throw std::invalid_argument(std::string(who) + ": pauli must contain only I/X/Y/Z");
}

}  // namespace

// ---- QuantumCircuit ------------------------------------------------------------
QuantumCircuit::QuantumCircuit(std::size_t num_qubits, std::size_t num_clbits)
    : num_qubits_(num_qubits), num_clbits_(num_clbits) {
    if (num_qubits == 0) throw std::invalid_argument("QuantumCircuit: num_qubits must be >= 1");
}

void QuantumCircuit::check_qubit(qubit_t q) const {
    if (q >= num_qubits_)
        throw std::out_of_range("qubit index " + std::to_string(q) + " out of range (num_qubits = " +
                                std::to_string(num_qubits_) + ")");
}

void QuantumCircuit::check_clbit(clbit_t c) const {
    if (c >= num_clbits_)
        throw std::out_of_range("clbit index " + std::to_string(c) + " out of range (num_clbits = " +
                                std::to_string(num_clbits_) + ")");
}

QuantumCircuit& QuantumCircuit::add_gate(Gate gate, std::vector<qubit_t> qubits) {
    if (qubits.size() != gate.num_qubits())
        throw std::invalid_argument(gate.name() + ": expects " + std::to_string(gate.num_qubits()) +
                                    " qubits, got " + std::to_string(qubits.size()));
    std::vector<qubit_t> sorted = qubits;
    for (qubit_t q : sorted) check_qubit(q);
    std::sort(sorted.begin(), sorted.end());
    const auto dup = std::adjacent_find(sorted.begin(), sorted.end());
    if (dup != sorted.end()) throw std::invalid_argument(gate.name() + ": duplicate qubit " + std::to_string(*dup));
    instructions_.push_back({std::move(gate), std::move(qubits), {}});
    return *this;
}

QuantumCircuit& QuantumCircuit::append(Instruction inst) {
    if (inst.gate.type == GateType::Measure) {
        if (inst.qubits.size() != 1 || inst.clbits.size() != 1)
            throw std::invalid_argument("append: malformed measure instruction");
        check_clbit(inst.clbits[0]);
        check_qubit(inst.qubits[0]);
        instructions_.push_back(std::move(inst));
        return *this;
    }
    return add_gate(std::move(inst.gate), std::move(inst.qubits));
}

#define LQCS_FIXED_1Q(fn, T) \
    QuantumCircuit& QuantumCircuit::fn(qubit_t q) { return add_gate({GateType::T}, {q}); }
LQCS_FIXED_1Q(i, I)
LQCS_FIXED_1Q(x, X)
LQCS_FIXED_1Q(y, Y)
LQCS_FIXED_1Q(z, Z)
LQCS_FIXED_1Q(h, H)
LQCS_FIXED_1Q(s, S)
LQCS_FIXED_1Q(sdg, Sdg)
LQCS_FIXED_1Q(t, T)
LQCS_FIXED_1Q(tdg, Tdg)
LQCS_FIXED_1Q(sx, SX)
#undef LQCS_FIXED_1Q

QuantumCircuit& QuantumCircuit::rx(double a, qubit_t q) { return add_gate({GateType::RX, {a}}, {q}); }
QuantumCircuit& QuantumCircuit::ry(double a, qubit_t q) { return add_gate({GateType::RY, {a}}, {q}); }
QuantumCircuit& QuantumCircuit::rz(double a, qubit_t q) { return add_gate({GateType::RZ, {a}}, {q}); }
QuantumCircuit& QuantumCircuit::p(double a, qubit_t q) { return add_gate({GateType::P, {a}}, {q}); }
QuantumCircuit& QuantumCircuit::u(double th, double ph, double la, qubit_t q) {
    return add_gate({GateType::U, {th, ph, la}}, {q});
}

QuantumCircuit& QuantumCircuit::cx(qubit_t c, qubit_t t) { return add_gate({GateType::X, {}, 1}, {c, t}); }
QuantumCircuit& QuantumCircuit::cy(qubit_t c, qubit_t t) { return add_gate({GateType::Y, {}, 1}, {c, t}); }
QuantumCircuit& QuantumCircuit::cz(qubit_t c, qubit_t t) { return add_gate({GateType::Z, {}, 1}, {c, t}); }
QuantumCircuit& QuantumCircuit::ch(qubit_t c, qubit_t t) { return add_gate({GateType::H, {}, 1}, {c, t}); }
QuantumCircuit& QuantumCircuit::cp(double a, qubit_t c, qubit_t t) {
    return add_gate({GateType::P, {a}, 1}, {c, t});
}
QuantumCircuit& QuantumCircuit::crx(double a, qubit_t c, qubit_t t) {
    return add_gate({GateType::RX, {a}, 1}, {c, t});
}
QuantumCircuit& QuantumCircuit::cry(double a, qubit_t c, qubit_t t) {
    return add_gate({GateType::RY, {a}, 1}, {c, t});
}
QuantumCircuit& QuantumCircuit::crz(double a, qubit_t c, qubit_t t) {
    return add_gate({GateType::RZ, {a}, 1}, {c, t});
}
QuantumCircuit& QuantumCircuit::ccx(qubit_t c1, qubit_t c2, qubit_t t) {
    return add_gate({GateType::X, {}, 2}, {c1, c2, t});
}
QuantumCircuit& QuantumCircuit::cswap(qubit_t c, qubit_t a, qubit_t b) {
    return add_gate({GateType::SWAP, {}, 1}, {c, a, b});
}

QuantumCircuit& QuantumCircuit::mcx(std::span<const qubit_t> controls, qubit_t t) {
    if (controls.empty()) throw std::invalid_argument("mcx: needs at least one control qubit");
    std::vector<qubit_t> qs(controls.begin(), controls.end());
    qs.push_back(t);
    return add_gate({GateType::X, {}, controls.size()}, std::move(qs));
}
QuantumCircuit& QuantumCircuit::mcx(std::initializer_list<qubit_t> controls, qubit_t t) {
    return mcx(std::span<const qubit_t>(controls.begin(), controls.size()), t);
}
QuantumCircuit& QuantumCircuit::mcp(double a, std::span<const qubit_t> controls, qubit_t t) {
    if (controls.empty()) throw std::invalid_argument("mcp: needs at least one control qubit");
    std::vector<qubit_t> qs(controls.begin(), controls.end());
    qs.push_back(t);
    return add_gate({GateType::P, {a}, controls.size()}, std::move(qs));
}
QuantumCircuit& QuantumCircuit::mcp(double a, std::initializer_list<qubit_t> controls, qubit_t t) {
    return mcp(a, std::span<const qubit_t>(controls.begin(), controls.size()), t);
}

QuantumCircuit& QuantumCircuit::swap(qubit_t a, qubit_t b) { return add_gate({GateType::SWAP}, {a, b}); }
QuantumCircuit& QuantumCircuit::iswap(qubit_t a, qubit_t b) { return add_gate({GateType::iSWAP}, {a, b}); }
QuantumCircuit& QuantumCircuit::rxx(double t, qubit_t a, qubit_t b) { return add_gate({GateType::RXX, {t}}, {a, b}); }
QuantumCircuit& QuantumCircuit::ryy(double t, qubit_t a, qubit_t b) { return add_gate({GateType::RYY, {t}}, {a, b}); }
QuantumCircuit& QuantumCircuit::rzz(double t, qubit_t a, qubit_t b) { return add_gate({GateType::RZZ, {t}}, {a, b}); }

QuantumCircuit& QuantumCircuit::rp(double theta, std::string_view pauli, std::span<const qubit_t> qubits) {
    if (pauli.empty() || pauli.size() != qubits.size())
        throw std::invalid_argument("rp: pauli length must equal qubit count and be non-empty");
    check_pauli(pauli, "rp");
    return add_gate(Gate{GateType::PauliRotation, {theta}, 0, {}, {}, std::string(pauli)},
                    std::vector<qubit_t>(qubits.begin(), qubits.end()));
}
QuantumCircuit& QuantumCircuit::rp(double theta, std::string_view pauli, std::initializer_list<qubit_t> qubits) {
    return rp(theta, pauli, std::span<const qubit_t>(qubits.begin(), qubits.size()));
}

QuantumCircuit& QuantumCircuit::unitary(std::span<const complex_t> m, std::span<const qubit_t> qubits) {
    const std::size_t dim = std::size_t{1} << qubits.size();
    if (m.size() != dim * dim) throw std::invalid_argument("unitary: matrix size must be 2^k x 2^k for k qubits");
    check_unitary(m, dim);
    return add_gate(Gate{GateType::Unitary, {}, 0, std::vector<complex_t>(m.begin(), m.end())},
                    std::vector<qubit_t>(qubits.begin(), qubits.end()));
}
QuantumCircuit& QuantumCircuit::unitary(std::span<const complex_t> m, std::initializer_list<qubit_t> qubits) {
    return unitary(m, std::span<const qubit_t>(qubits.begin(), qubits.size()));
}

QuantumCircuit& QuantumCircuit::permutation(std::span<const std::size_t> table, std::span<const qubit_t> qubits) {
    const std::size_t dim = std::size_t{1} << qubits.size();
    if (table.size() != dim) throw std::invalid_argument("permutation: table size must be 2^k for k qubits");
    std::vector<bool> hit(dim, false);
    for (std::size_t v : table) {
        if (v >= dim || hit[v]) throw std::invalid_argument("permutation: table is not a bijection");
        hit[v] = true;
    }
    return add_gate(Gate{GateType::Permutation, {}, 0, {}, std::vector<std::size_t>(table.begin(), table.end())},
                    std::vector<qubit_t>(qubits.begin(), qubits.end()));
}
QuantumCircuit& QuantumCircuit::permutation(std::span<const std::size_t> table, std::initializer_list<qubit_t> qubits) {
    return permutation(table, std::span<const qubit_t>(qubits.begin(), qubits.size()));
}

QuantumCircuit& QuantumCircuit::measure(qubit_t q, clbit_t c) {
    check_qubit(q);
    check_clbit(c);
    instructions_.push_back({Gate{GateType::Measure}, {q}, {c}});
    return *this;
}

QuantumCircuit& QuantumCircuit::measure_all() {
    num_clbits_ = std::max(num_clbits_, num_qubits_);
    for (qubit_t q = 0; q < num_qubits_; ++q) measure(q, q);
    return *this;
}

QuantumCircuit& QuantumCircuit::reset(qubit_t q) {
    check_qubit(q);
    instructions_.push_back({Gate{GateType::Reset}, {q}, {}});
    return *this;
}

QuantumCircuit& QuantumCircuit::barrier() {
    instructions_.push_back({Gate{GateType::Barrier}, {}, {}});
    return *this;
}

QuantumCircuit& QuantumCircuit::compose(const QuantumCircuit& other, std::span<const qubit_t> qubit_map) {
    std::vector<qubit_t> map(qubit_map.begin(), qubit_map.end());
    if (map.empty()) {
        if (other.num_qubits() > num_qubits_) throw std::invalid_argument("compose: sub-circuit has more qubits");
        for (std::size_t q = 0; q < other.num_qubits(); ++q) map.push_back(static_cast<qubit_t>(q));
    }
    if (map.size() != other.num_qubits())
        throw std::invalid_argument("compose: qubit_map size must equal sub-circuit qubit count");
    std::vector<qubit_t> sorted = map;
    for (qubit_t q : sorted) check_qubit(q);
    std::sort(sorted.begin(), sorted.end());
    if (std::adjacent_find(sorted.begin(), sorted.end()) != sorted.end())
        throw std::invalid_argument("compose: duplicate qubit in map");
    if (other.num_clbits() > num_clbits_) throw std::invalid_argument("compose: sub-circuit has more clbits");
    for (const Instruction& in : other.instructions()) {
        Instruction mapped = in;
        for (qubit_t& q : mapped.qubits) q = map[q];
        instructions_.push_back(std::move(mapped));
    }
    return *this;
}

QuantumCircuit QuantumCircuit::inverse() const {
    QuantumCircuit inv(num_qubits_, num_clbits_);
    for (auto it = instructions_.rbegin(); it != instructions_.rend(); ++it) {
        if (it->gate.type == GateType::Barrier) {
            inv.barrier();
            continue;
        }
        if (!it->gate.is_unitary())
            throw std::invalid_argument("inverse: circuit contains non-unitary operation '" + it->gate.name() + "'");
        inv.instructions_.push_back({it->gate.inverse(), it->qubits, {}});
    }
    return inv;
}

QuantumCircuit QuantumCircuit::power(int n) const {
    QuantumCircuit out(num_qubits_, num_clbits_);
    const QuantumCircuit unit = n < 0 ? inverse() : *this;
    for (int r = 0; r < std::abs(n); ++r) out.compose(unit);
    return out;
}

std::size_t QuantumCircuit::depth() const {
    std::vector<std::size_t> level(num_qubits_, 0);
    for (const Instruction& in : instructions_) {
        if (in.gate.type == GateType::Barrier) continue;
        std::size_t d = 0;
        for (qubit_t q : in.qubits) d = std::max(d, level[q]);
        for (qubit_t q : in.qubits) level[q] = d + 1;
    }
    return *std::max_element(level.begin(), level.end());
}

// Plain listing, one instruction per line.
std::string QuantumCircuit::draw() const {
    std::ostringstream os;
    os << "circuit(" << num_qubits_ << " qubits, " << num_clbits_ << " clbits)\n";
    for (const Instruction& in : instructions_) {
        os << "  " << in.gate.name();
        if (!in.gate.params.empty()) {
            os << "(";
            for (std::size_t k = 0; k < in.gate.params.size(); ++k) os << (k ? "," : "") << in.gate.params[k];
            os << ")";
        }
        for (qubit_t q : in.qubits) os << " q" << q;
        for (clbit_t c : in.clbits) os << " -> c" << c;
        os << "\n";
    }
    return os.str();
}

// ---- ParametricCircuit ---------------------------------------------------------
ParametricCircuit::ParametricCircuit(std::size_t num_qubits, std::size_t num_clbits)
    : template_(num_qubits, num_clbits) {}

#define LQCS_PC_FIXED(fn) \
    ParametricCircuit& ParametricCircuit::fn(qubit_t q) { template_.fn(q); return *this; }
LQCS_PC_FIXED(x)
LQCS_PC_FIXED(y)
LQCS_PC_FIXED(z)
LQCS_PC_FIXED(h)
LQCS_PC_FIXED(s)
LQCS_PC_FIXED(sdg)
LQCS_PC_FIXED(t)
LQCS_PC_FIXED(sx)
#undef LQCS_PC_FIXED
ParametricCircuit& ParametricCircuit::cx(qubit_t c, qubit_t t) { template_.cx(c, t); return *this; }
ParametricCircuit& ParametricCircuit::cy(qubit_t c, qubit_t t) { template_.cy(c, t); return *this; }
ParametricCircuit& ParametricCircuit::cz(qubit_t c, qubit_t t) { template_.cz(c, t); return *this; }
ParametricCircuit& ParametricCircuit::swap(qubit_t a, qubit_t b) { template_.swap(a, b); return *this; }
ParametricCircuit& ParametricCircuit::ccx(qubit_t a, qubit_t b, qubit_t t) { template_.ccx(a, b, t); return *this; }
ParametricCircuit& ParametricCircuit::barrier() { template_.barrier(); return *this; }

std::size_t ParametricCircuit::add_param_gate(Gate gate, std::vector<qubit_t> qubits) {
    const std::size_t at = template_.num_instructions();
    template_.append({std::move(gate), std::move(qubits), {}});
    param_inst_.push_back(at);
    return param_inst_.size() - 1;
}

std::size_t ParametricCircuit::rx(qubit_t q) { return add_param_gate({GateType::RX, {0.0}}, {q}); }
std::size_t ParametricCircuit::ry(qubit_t q) { return add_param_gate({GateType::RY, {0.0}}, {q}); }
std::size_t ParametricCircuit::rz(qubit_t q) { return add_param_gate({GateType::RZ, {0.0}}, {q}); }

std::size_t ParametricCircuit::rp(std::string_view pauli, std::span<const qubit_t> qubits) {
    if (pauli.empty() || pauli.size() != qubits.size())
        throw std::invalid_argument("ParametricCircuit::rp: pauli length must equal qubit count");
    check_pauli(pauli, "rp");
    return add_param_gate(Gate{GateType::PauliRotation, {0.0}, 0, {}, {}, std::string(pauli)},
                          std::vector<qubit_t>(qubits.begin(), qubits.end()));
}
std::size_t ParametricCircuit::rp(std::string_view pauli, std::initializer_list<qubit_t> qubits) {
    return rp(pauli, std::span<const qubit_t>(qubits.begin(), qubits.size()));
}

QuantumCircuit ParametricCircuit::bind(std::span<const double> values) const {
    if (values.size() != num_parameters())
        throw std::invalid_argument("bind: number of values must equal num_parameters()");
    std::vector<std::size_t> slot(template_.num_instructions(), SIZE_MAX);
    for (std::size_t p = 0; p < param_inst_.size(); ++p) slot[param_inst_[p]] = p;
    QuantumCircuit out(template_.num_qubits(), template_.num_clbits());
    const auto insts = template_.instructions();
    for (std::size_t i = 0; i < insts.size(); ++i) {
        Instruction copy = insts[i];
        if (slot[i] != SIZE_MAX) copy.gate.params[0] = values[slot[i]];
        out.append(std::move(copy));
    }
    return out;
}

// ---- Result ----------------------------------------------------------------------
std::map<std::string, double> Result::probabilities() const {
    std::map<std::string, double> out;
    if (shots_ == 0) return out;
    for (const auto& [k, n] : counts_) out[k] = static_cast<double>(n) / static_cast<double>(shots_);
    return out;
}

void Result::print_counts(std::ostream& os) const {
    os << "{";
    bool first = true;
    for (const auto& [k, n] : counts_) {
        os << (first ? "" : ", ") << "\"" << k << "\": " << n;
        first = false;
    }
    os << "}\n";
}

// ---- Kraus channels ----------------------------------------------------------------
namespace {
void check_probability(double p, const char* who) {
    if (p < 0.0 || p > 1.0) throw std::invalid_argument(std::string(who) + ": probability must be in [0, 1]");
}
}  // namespace

KrausChannel::KrausChannel(std::vector<std::array<complex_t, 4>> kraus_ops) : ops(std::move(kraus_ops)) {
    if (ops.empty()) throw std::invalid_argument("KrausChannel: needs at least one operator");
    complex_t s00{0, 0}, s01{0, 0}, s10{0, 0}, s11{0, 0};
    for (const auto& k : ops) {  // sum K^dag K
        s00 += std::conj(k[0]) * k[0] + std::conj(k[2]) * k[2];
        s01 += std::conj(k[0]) * k[1] + std::conj(k[2]) * k[3];
        s10 += std::conj(k[1]) * k[0] + std::conj(k[3]) * k[2];
        s11 += std::conj(k[1]) * k[1] + std::conj(k[3]) * k[3];
    }
    const double tol = 1e-10;
    if (std::abs(s00 - 1.0) > tol || std::abs(s01) > tol || std::abs(s10) > tol || std::abs(s11 - 1.0) > tol)
        throw std::invalid_argument("KrausChannel: completeness violated (sum K_i^dag K_i != I)");
}

KrausChannel KrausChannel::depolarizing(double p) {
    check_probability(p, "depolarizing");
    const double a = std::sqrt(1.0 - 3.0 * p / 4.0), b = std::sqrt(p / 4.0);
    const complex_t i{0, 1};
    return KrausChannel({{a, 0, 0, a}, {0, b, b, 0}, {0, -b * i, b * i, 0}, {b, 0, 0, -b}});
}

KrausChannel KrausChannel::amplitude_damping(double g) {
    check_probability(g, "amplitude_damping");
    return KrausChannel({{1, 0, 0, std::sqrt(1.0 - g)}, {0, std::sqrt(g), 0, 0}});
}

KrausChannel KrausChannel::phase_damping(double g) {
    check_probability(g, "phase_damping");
    return KrausChannel({{1, 0, 0, std::sqrt(1.0 - g)}, {0, 0, 0, std::sqrt(g)}});
}

KrausChannel KrausChannel::bit_flip(double p) {
    check_probability(p, "bit_flip");
    const double a = std::sqrt(1.0 - p), b = std::sqrt(p);
    return KrausChannel({{a, 0, 0, a}, {0, b, b, 0}});
}

KrausChannel KrausChannel::phase_flip(double p) {
    check_probability(p, "phase_flip");
    const double a = std::sqrt(1.0 - p), b = std::sqrt(p);
    return KrausChannel({{a, 0, 0, a}, {b, 0, 0, -b}});
}

}  // namespace lqcs
