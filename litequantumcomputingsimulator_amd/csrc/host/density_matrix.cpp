// Density matrix and density-matrix simulator on the device engine. Public
// behaviour (caps, exceptions, channel order, terminal-only measurement, RNG
// consumption, bitstring format) follows src/backend/density_matrix.cpp and
// src/backend/density_matrix_simulator.cpp of the reference; every operation
// on rho runs on the GPU through lqcs_device.h.
#include <algorithm>
#include <map>
#include <numeric>
#include <optional>
#include <stdexcept>
#include <string>
#include <utility>

#include "lowering.hpp"
#include "lqcs/backend/density_matrix_simulator.hpp"
#include "lqcs/core/bit_utils.hpp"
#include "lqcs/core/random.hpp"

namespace lqcs {

using detail::check;

namespace {

lqcs_exec_options exec_of(const EngineOptions& e) {
    lqcs_exec_options o{};
    o.plan = static_cast<int32_t>(e.plan);
    o.tile_qubits = e.tile_qubits;
    o.low_qubits = e.low_qubits;
    o.tile_kernel = e.tile_kernel;
    return o;
}

// sum_i K_i (x) conj(K_i) on (ket q, bra q): local index l = ket bit + 2 * bra bit,
// S[l][l'] = sum_i K_i[l & 1][l' & 1] * conj(K_i[l >> 1][l' >> 1]) (density_matrix.cpp:68-84
// applies K_i on the ket qubit and conj(K_i) on the bra qubit, then sums over i).
std::vector<complex_t> superoperator(const KrausChannel& ch) {
    std::vector<complex_t> s(16, complex_t{0, 0});
    for (const auto& k : ch.ops)
        for (int l = 0; l < 4; ++l)
            for (int lp = 0; lp < 4; ++lp)
                s[4 * l + lp] += k[2 * (l & 1) + (lp & 1)] * std::conj(k[2 * (l >> 1) + (lp >> 1)]);
    return s;
}

// 4x4 product a * b (row-major).
std::vector<complex_t> mul4(const std::vector<complex_t>& a, const std::vector<complex_t>& b) {
    std::vector<complex_t> c(16, complex_t{0, 0});
    for (int i = 0; i < 4; ++i)
        for (int k = 0; k < 4; ++k)
            for (int j = 0; j < 4; ++j) c[4 * i + j] += a[4 * i + k] * b[4 * k + j];
    return c;
}

void check_qubits(const Instruction& inst, std::size_t n) {
    for (qubit_t q : inst.qubits)
        if (q >= n) throw std::out_of_range("DensityMatrix: qubit index out of range");
}

void add_unitary(detail::ProgramBuilder& b, const Instruction& inst, std::size_t n) {
    check_qubits(inst, n);
    b.add_instruction(inst);                                  // ket side: U
    b.add_instruction_conj(inst, static_cast<qubit_t>(n));    // bra side: conj(U)
}

void add_channel(detail::ProgramBuilder& b, const KrausChannel& ch, qubit_t q, std::size_t n) {
    b.add_two_qubit(q, static_cast<qubit_t>(q + n), superoperator(ch));
}

void apply_program(lqcs_dev_state* dev, const detail::ProgramBuilder& b, const EngineOptions& e) {
    if (b.size() == 0) return;
    const lqcs_program prog = b.program();
    const lqcs_exec_options o = exec_of(e);
    check(lqcs_state_apply(dev, &prog, &o));
}

}  // namespace

// ---- DensityMatrix ----------------------------------------------------------------------
DensityMatrix::DensityMatrix(std::size_t num_qubits) : DensityMatrix(num_qubits, kMaxQubits) {}

DensityMatrix::DensityMatrix(std::size_t num_qubits, std::size_t max_qubits) : num_qubits_(num_qubits) {
    const std::size_t cap = std::min(max_qubits, kMaxDeviceQubits);
    if (num_qubits == 0 || num_qubits > cap)
        throw std::invalid_argument("DensityMatrix: num_qubits must be in [1, " + std::to_string(cap) +
                                    "] (memory is O(4^n))");
    check(lqcs_state_create(static_cast<uint32_t>(2 * num_qubits), &dev_));  // |0...0><0...0|
}

DensityMatrix::~DensityMatrix() {
    if (dev_) lqcs_state_destroy(dev_);
}

DensityMatrix::DensityMatrix(const DensityMatrix& o) : num_qubits_(o.num_qubits_) {
    check(lqcs_state_create(static_cast<uint32_t>(2 * num_qubits_), &dev_));
    check(lqcs_state_copy(dev_, o.dev_));
}

DensityMatrix& DensityMatrix::operator=(const DensityMatrix& o) {
    if (this != &o) {
        DensityMatrix tmp(o);
        *this = std::move(tmp);
    }
    return *this;
}

DensityMatrix::DensityMatrix(DensityMatrix&& o) noexcept : num_qubits_(o.num_qubits_), dev_(o.dev_) {
    o.dev_ = nullptr;
}

DensityMatrix& DensityMatrix::operator=(DensityMatrix&& o) noexcept {
    if (this != &o) {
        if (dev_) lqcs_state_destroy(dev_);
        num_qubits_ = o.num_qubits_;
        dev_ = o.dev_;
        o.dev_ = nullptr;
    }
    return *this;
}

complex_t DensityMatrix::element(std::size_t row, std::size_t col) const {
    if (row >= dim() || col >= dim()) throw std::out_of_range("DensityMatrix::element: index out of range");
    double z[2];
    check(lqcs_state_read(dev_, (uint64_t(col) << num_qubits_) | row, 1, z));
    return {z[0], z[1]};
}

void DensityMatrix::apply_instruction(const Instruction& inst) {
    if (!inst.gate.is_unitary())
        throw std::invalid_argument("DensityMatrix: '" + inst.gate.name() + "' is not unitary");
    detail::ProgramBuilder b;
    add_unitary(b, inst, num_qubits_);
    apply_program(dev_, b, EngineOptions{});
}

void DensityMatrix::apply_channel(const KrausChannel& channel, qubit_t q) {
    if (q >= num_qubits_) throw std::out_of_range("apply_channel: qubit index out of range");
    detail::ProgramBuilder b;
    add_channel(b, channel, q, num_qubits_);
    apply_program(dev_, b, EngineOptions{});
}

double DensityMatrix::trace() const {
    const std::vector<double> d = probabilities();
    double tr = 0.0;
    for (double v : d) tr += v;
    return tr;
}

double DensityMatrix::purity() const {
    double s = 0.0;
    check(lqcs_state_norm_sq(dev_, &s));  // sum_rc |rho_rc|^2 = Tr(rho^2) for Hermitian rho
    return s;
}

std::vector<double> DensityMatrix::probabilities() const {
    std::vector<double> p(dim());
    check(lqcs_dm_diagonal(dev_, p.data()));
    return p;
}

double DensityMatrix::fidelity(const Statevector& psi) const {
    if (psi.num_qubits() != num_qubits_) throw std::invalid_argument("fidelity: qubit count mismatch");
    double f = 0.0;
    check(lqcs_dm_fidelity(dev_, psi.device(), &f));
    return f;
}

// ---- DensityMatrixSimulator ----------------------------------------------------------------
// density_matrix_simulator.cpp:13-28: Reset rejected, Measure/Barrier skipped,
// every channel on every touched qubit after each unitary, channel-major. The
// whole circuit becomes one op stream, planned into tiled passes. With noise,
// everything that acts on one (ket q, bra q) pair alone -- a single-qubit gate
// with its channels, S * (U (x) conj(U)), or the channels after a multi-qubit
// gate -- is collected into one pending 4x4 per qubit and multiplied by the
// next such block on that pair; a pending block is emitted just before the
// next op that touches q, or at the end. The pair's blocks commute with every
// op on other qubits, so the order of operations on rho is unchanged.
DensityMatrix DensityMatrixSimulator::run_density_matrix(const QuantumCircuit& circuit) const {
    const std::size_t n = circuit.num_qubits();
    DensityMatrix rho(n, options_.max_qubits);
    // The channels act on disjoint (ket, bra) pairs for different qubits, so
    // their channel-major order reduces to one composed superoperator per
    // touched qubit: S = S_last ... S_first.
    const auto& channels = options_.noise.channels();
    std::vector<complex_t> chain;
    for (const KrausChannel& ch : channels) chain = chain.empty() ? superoperator(ch) : mul4(superoperator(ch), chain);
    detail::ProgramBuilder b;
    std::vector<std::vector<complex_t>> pending(n);  // empty: nothing pending on that pair
    auto push = [&](qubit_t q, const std::vector<complex_t>& m) {
        pending[q] = pending[q].empty() ? m : mul4(m, pending[q]);
    };
    auto flush = [&](qubit_t q) {
        if (pending[q].empty()) return;
        b.add_two_qubit(q, static_cast<qubit_t>(q + n), pending[q]);
        pending[q].clear();
    };
    for (const Instruction& inst : circuit.instructions()) {
        if (inst.gate.type == GateType::Reset)
            throw std::invalid_argument("DensityMatrixSimulator: Reset is not supported yet");
        const Gate& g = inst.gate;
        if (!g.is_unitary()) continue;
        if (!chain.empty() && g.n_controls == 0 && g.base_qubits() == 1 && g.type != GateType::Permutation) {
            // A noisy single-qubit gate: S * (U (x) conj(U)) on (q, q + n).
            check_qubits(inst, n);
            const std::vector<complex_t> u = g.base_matrix();
            std::vector<complex_t> m(16);
            for (int l = 0; l < 4; ++l)
                for (int lp = 0; lp < 4; ++lp)
                    m[4 * l + lp] = u[2 * (l & 1) + (lp & 1)] * std::conj(u[2 * (l >> 1) + (lp >> 1)]);
            push(inst.qubits[0], mul4(chain, m));
            continue;
        }
        check_qubits(inst, n);
        for (qubit_t q : inst.qubits) flush(q);
        add_unitary(b, inst, n);
        if (!chain.empty())
            for (qubit_t q : inst.qubits) push(q, chain);
    }
    for (qubit_t q = 0; q < n; ++q) flush(q);
    apply_program(rho.device(), b, options_.engine);
    return rho;
}

// density_matrix_simulator.cpp:30-82: terminal measurements only; the diagonal
// is the outcome distribution, sampled with the statevector path's cumulative
// search and RNG.
Result DensityMatrixSimulator::run(const QuantumCircuit& circuit, std::size_t shots) {
    if (shots == 0) throw std::invalid_argument("run: shots must be >= 1");
    bool seen_measure = false;
    for (const Instruction& inst : circuit.instructions()) {
        if (inst.gate.type == GateType::Measure)
            seen_measure = true;
        else if (seen_measure && inst.gate.is_unitary())
            throw std::invalid_argument("DensityMatrixSimulator: mid-circuit measurement is not supported");
    }
    if (!seen_measure) throw std::invalid_argument("run: circuit has no measurements; use run_density_matrix()");

    const DensityMatrix rho = run_density_matrix(circuit);
    std::vector<std::optional<qubit_t>> source(circuit.num_clbits());
    for (const Instruction& inst : circuit.instructions())
        if (inst.gate.type == GateType::Measure) source[inst.clbits[0]] = inst.qubits[0];

    std::vector<double> cum = rho.probabilities();
    std::partial_sum(cum.begin(), cum.end(), cum.begin());
    const double total = cum.back();
    Rng rng(options_.seed);
    std::map<std::string, std::size_t> counts;
    const std::size_t nc = circuit.num_clbits();
    std::string key(nc, '0');
    for (std::size_t s = 0; s < shots; ++s) {
        const double r = rng.uniform() * total;
        const std::size_t idx = static_cast<std::size_t>(std::upper_bound(cum.begin(), cum.end(), r) - cum.begin());
        for (std::size_t c = 0; c < nc; ++c)
            key[nc - 1 - c] = (source[c] && bits::test_bit(idx, *source[c])) ? '1' : '0';
        ++counts[key];
    }
    return Result(std::move(counts), shots);
}

}  // namespace lqcs
