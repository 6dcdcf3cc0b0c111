// Gate model: host-side matrices that feed every device operation.
// The entries follow the gate definitions of the reference
// (src/circuit/gate.cpp:94-201) expression for expression, so the doubles
// handed to the device are the ones the reference's kernels would see.
#include <bit>
#include <cmath>
#include <numbers>
#include <stdexcept>

#include "lqcs/circuit/gate.hpp"

namespace lqcs {
namespace {

constexpr complex_t kI{0.0, 1.0};
const double kHalfSqrt2 = std::numbers::sqrt2 / 2.0;
constexpr complex_t kIPow[4] = {{1, 0}, {0, 1}, {-1, 0}, {0, -1}};

complex_t expi(double a) { return {std::cos(a), std::sin(a)}; }

std::vector<complex_t> dagger(const std::vector<complex_t>& m, std::size_t dim) {
    std::vector<complex_t> out(dim * dim);
    for (std::size_t r = 0; r < dim; ++r)
        for (std::size_t c = 0; c < dim; ++c) out[c * dim + r] = std::conj(m[r * dim + c]);
    return out;
}

}  // namespace

std::size_t Gate::base_qubits() const {
    switch (type) {
        case GateType::SWAP:
        case GateType::iSWAP:
        case GateType::RXX:
        case GateType::RYY:
        case GateType::RZZ:
            return 2;
        case GateType::Unitary:
            return static_cast<std::size_t>(
                std::countr_zero(static_cast<std::size_t>(std::sqrt(static_cast<double>(mat.size())))));
        case GateType::Permutation:
            return static_cast<std::size_t>(std::countr_zero(perm.size()));
        case GateType::PauliRotation:
            return pauli.size();
        case GateType::Barrier:
            return 0;
        default:
            return 1;
    }
}

Gate::PauliMasks Gate::pauli_masks() const {
    PauliMasks m;
    const std::size_t k = pauli.size();
    for (std::size_t j = 0; j < k; ++j) {
        const std::size_t bit = std::size_t{1} << j;
        switch (pauli[k - 1 - j]) {  // local bit j <-> j-th char from the right
            case 'I': break;
            case 'X': m.x |= bit; break;
            case 'Y': m.x |= bit; m.zy |= bit; ++m.n_y; break;
            case 'Z': m.zy |= bit; break;
            default: throw std::invalid_argument("PauliRotation: pauli must contain only I/X/Y/Z");
        }
    }
    return m;
}

bool Gate::is_unitary() const {
    return type != GateType::Measure && type != GateType::Reset && type != GateType::Barrier;
}

bool Gate::is_diagonal() const {
    switch (type) {
        case GateType::I: case GateType::Z: case GateType::S: case GateType::Sdg:
        case GateType::T: case GateType::Tdg: case GateType::RZ: case GateType::P:
        case GateType::RZZ:
            return true;
        case GateType::PauliRotation:
            return pauli_masks().x == 0;
        default:
            return false;
    }
}

std::vector<complex_t> Gate::base_matrix() const {
    if (!is_unitary()) throw std::logic_error("Gate::base_matrix: " + name() + " has no matrix");
    switch (type) {
        case GateType::I: return {1, 0, 0, 1};
        case GateType::X: return {0, 1, 1, 0};
        case GateType::Y: return {0, -kI, kI, 0};
        case GateType::Z: return {1, 0, 0, -1};
        case GateType::H: return {kHalfSqrt2, kHalfSqrt2, kHalfSqrt2, -kHalfSqrt2};
        case GateType::S: return {1, 0, 0, kI};
        case GateType::Sdg: return {1, 0, 0, -kI};
        case GateType::T: return {1, 0, 0, expi(std::numbers::pi / 4)};
        case GateType::Tdg: return {1, 0, 0, expi(-std::numbers::pi / 4)};
        case GateType::SX:
            return {complex_t{0.5, 0.5}, complex_t{0.5, -0.5}, complex_t{0.5, -0.5}, complex_t{0.5, 0.5}};
        case GateType::RX: {
            const double h = params.at(0) / 2;
            return {std::cos(h), -kI * std::sin(h), -kI * std::sin(h), std::cos(h)};
        }
        case GateType::RY: {
            const double h = params.at(0) / 2;
            return {std::cos(h), -std::sin(h), std::sin(h), std::cos(h)};
        }
        case GateType::RZ: {
            const double h = params.at(0) / 2;
            return {expi(-h), 0, 0, expi(h)};
        }
        case GateType::P: return {1, 0, 0, expi(params.at(0))};
        case GateType::U: {
            const double h = params.at(0) / 2, phi = params.at(1), lam = params.at(2);
            return {std::cos(h), -expi(lam) * std::sin(h), expi(phi) * std::sin(h),
                    expi(phi + lam) * std::cos(h)};
        }
        case GateType::SWAP:
        case GateType::iSWAP: {
            const complex_t off = type == GateType::SWAP ? complex_t{1, 0} : kI;
            return {1, 0, 0, 0, 0, 0, off, 0, 0, off, 0, 0, 0, 0, 0, 1};
        }
        case GateType::RXX:
        case GateType::RYY: {
            const double h = params.at(0) / 2;
            const complex_t c = std::cos(h), is = kI * std::sin(h);
            const complex_t corner = type == GateType::RXX ? -is : is;
            return {c, 0, 0, corner, 0, c, -is, 0, 0, -is, c, 0, corner, 0, 0, c};
        }
        case GateType::RZZ: {
            const double h = params.at(0) / 2;
            return {expi(-h), 0, 0, 0, 0, expi(h), 0, 0, 0, 0, expi(h), 0, 0, 0, 0, expi(-h)};
        }
        case GateType::PauliRotation: {
            // exp(-i theta/2 P) = cos(theta/2) I - i sin(theta/2) P with
            // P|col> = i^ny (-1)^popcount(col & zy) |col ^ x>
            const std::size_t dim = std::size_t{1} << pauli.size();
            const PauliMasks pm = pauli_masks();
            const double h = params.at(0) / 2;
            const complex_t c = std::cos(h);
            const complex_t mis{0.0, -std::sin(h)};
            std::vector<complex_t> out(dim * dim, complex_t{0, 0});
            for (std::size_t col = 0; col < dim; ++col) {
                out[col * dim + col] += c;
                const double sign = (std::popcount(col & pm.zy) % 2 == 0) ? 1.0 : -1.0;
                out[(col ^ pm.x) * dim + col] += mis * kIPow[pm.n_y % 4] * sign;
            }
            return out;
        }
        case GateType::Unitary:
            return mat;
        case GateType::Permutation: {
            const std::size_t dim = perm.size();
            std::vector<complex_t> out(dim * dim, complex_t{0, 0});
            for (std::size_t l = 0; l < dim; ++l) out[perm[l] * dim + l] = 1;
            return out;
        }
        default:
            throw std::logic_error("Gate::base_matrix: unhandled gate type");
    }
}

std::vector<complex_t> Gate::matrix() const {
    const std::vector<complex_t> base = base_matrix();
    if (n_controls == 0) return base;
    const std::size_t bdim = std::size_t{1} << base_qubits();
    const std::size_t dim = std::size_t{1} << num_qubits();
    const std::size_t all = (std::size_t{1} << n_controls) - 1;
    std::vector<complex_t> out(dim * dim, complex_t{0, 0});
    for (std::size_t col = 0; col < dim; ++col) {
        if ((col & all) != all) {
            out[col * dim + col] = 1;
            continue;
        }
        for (std::size_t tr = 0; tr < bdim; ++tr)
            out[((tr << n_controls) | all) * dim + col] = base[tr * bdim + (col >> n_controls)];
    }
    return out;
}

Gate Gate::inverse() const {
    if (!is_unitary()) throw std::logic_error("Gate::inverse: " + name() + " is not invertible");
    Gate g = *this;
    switch (type) {
        case GateType::I: case GateType::X: case GateType::Y: case GateType::Z:
        case GateType::H: case GateType::SWAP:
            return g;
        case GateType::RX: case GateType::RY: case GateType::RZ: case GateType::P:
        case GateType::RXX: case GateType::RYY: case GateType::RZZ: case GateType::PauliRotation:
            g.params[0] = -g.params[0];
            return g;
        case GateType::S: g.type = GateType::Sdg; return g;
        case GateType::Sdg: g.type = GateType::S; return g;
        case GateType::T: g.type = GateType::Tdg; return g;
        case GateType::Tdg: g.type = GateType::T; return g;
        case GateType::U:
            g.params = {-params.at(0), -params.at(2), -params.at(1)};
            return g;
        case GateType::SX:
        case GateType::iSWAP:
            g.type = GateType::Unitary;
            g.mat = dagger(base_matrix(), std::size_t{1} << base_qubits());
            return g;
        case GateType::Unitary:
            g.mat = dagger(mat, std::size_t{1} << base_qubits());
            return g;
        case GateType::Permutation:
            for (std::size_t l = 0; l < perm.size(); ++l) g.perm[perm[l]] = l;
            return g;
        default:
            throw std::logic_error("Gate::inverse: unhandled gate type");
    }
}

std::string Gate::name() const {
    static const char* const kNames[] = {
        "id", "x", "y", "z", "h", "s", "sdg", "t", "tdg", "sx", "rx", "ry", "rz", "p", "u",
        "swap", "iswap", "rxx", "ryy", "rzz", "rp", "unitary", "perm", "measure", "reset", "barrier"};
    std::string base = kNames[static_cast<int>(type)];
    if (type == GateType::PauliRotation) base = "rp(" + pauli + ")";
    if (n_controls == 0) return base;
    if (n_controls == 1) return "c" + base;
    if (n_controls == 2) return "cc" + base;
    return "mc" + base;
}

}  // namespace lqcs
