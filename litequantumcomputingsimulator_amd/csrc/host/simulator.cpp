// Statevector container and simulator on the device engine. The public
// behaviour (path selection, exceptions, RNG consumption, bitstring format)
// follows src/backend/statevector_simulator.cpp and statevector.cpp of the
// reference; every amplitude operation runs on the GPU through lqcs_device.h.
#include <cstring>
#include <fstream>
#include <map>
#include <optional>
#include <stdexcept>
#include <string>
#include <utility>

#include "lowering.hpp"
#include "lqcs/backend/compiled.hpp"
#include "lqcs/backend/statevector_simulator.hpp"
#include "lqcs/core/random.hpp"
#include "lqcs/transpiler/gate_fusion.hpp"
#include "lqcs/transpiler/redundant_elimination.hpp"

namespace lqcs {

using detail::check;

// ---- Statevector ------------------------------------------------------------------
Statevector::Statevector(std::size_t num_qubits) : Statevector(num_qubits, kMaxQubits) {}

Statevector::Statevector(std::size_t num_qubits, std::size_t max_qubits) : num_qubits_(num_qubits) {
    if (num_qubits == 0) throw std::invalid_argument("Statevector: num_qubits must be >= 1");
    const std::size_t cap = std::min<std::size_t>(max_qubits, LQCS_DEVICE_MAX_QUBITS);
    if (num_qubits > cap) {
        const double gib = num_qubits >= 30 ? 16.0 * static_cast<double>(std::size_t{1} << (num_qubits - 30))
                                            : 16.0 / static_cast<double>(std::size_t{1} << (30 - num_qubits));
        throw std::invalid_argument("Statevector: " + std::to_string(num_qubits) + " qubits would need " +
                                    std::to_string(gib) + " GiB (max " + std::to_string(cap) + " qubits)");
    }
    check(lqcs_state_create(static_cast<uint32_t>(num_qubits), &dev_));
    fresh_ = Fresh::Device;
}

Statevector::~Statevector() {
    if (dev_) lqcs_state_destroy(dev_);
}

Statevector::Statevector(const Statevector& o) : num_qubits_(o.num_qubits_) {
    check(lqcs_state_create(static_cast<uint32_t>(num_qubits_), &dev_));
    check(lqcs_state_copy(dev_, o.device()));
    fresh_ = Fresh::Device;
}

Statevector& Statevector::operator=(const Statevector& o) {
    if (this != &o) {
        Statevector tmp(o);
        *this = std::move(tmp);
    }
    return *this;
}

Statevector::Statevector(Statevector&& o) noexcept
    : num_qubits_(o.num_qubits_), dev_(o.dev_), host_(std::move(o.host_)), fresh_(o.fresh_),
      host_exported_(o.host_exported_) {
    o.dev_ = nullptr;
}

Statevector& Statevector::operator=(Statevector&& o) noexcept {
    if (this != &o) {
        if (dev_) lqcs_state_destroy(dev_);
        num_qubits_ = o.num_qubits_;
        dev_ = o.dev_;
        host_ = std::move(o.host_);
        fresh_ = o.fresh_;
        host_exported_ = o.host_exported_;
        o.dev_ = nullptr;
    }
    return *this;
}

void Statevector::make_host_current() const {
    if (host_.size() != size()) host_.resize(size());
    if (fresh_ == Fresh::Device) {
        check(lqcs_state_download(dev_, reinterpret_cast<double*>(host_.data())));
        fresh_ = Fresh::Both;
    }
}

complex_t* Statevector::data() {
    make_host_current();
    fresh_ = Fresh::Host;  // caller may write through the pointer
    host_exported_ = true;
    return host_.data();
}

const complex_t* Statevector::data() const {
    make_host_current();
    return host_.data();
}

lqcs_dev_state* Statevector::device() const {
    if (fresh_ == Fresh::Host || host_exported_) {
        check(lqcs_state_upload(dev_, reinterpret_cast<const double*>(host_.data())));
        fresh_ = Fresh::Both;
    }
    return dev_;
}

void Statevector::mark_device_newer() {
    // Host views handed out before are stale now: the next host access
    // downloads into the same buffer (and hands it out again).
    fresh_ = Fresh::Device;
    host_exported_ = false;
}

void Statevector::set_basis(std::size_t index) {
    if (index >= size()) throw std::out_of_range("set_basis: index out of range");
    check(lqcs_state_set_basis(dev_, index));
    host_exported_ = false;
    fresh_ = Fresh::Device;
}

double Statevector::norm() const {
    double s = 0.0;
    check(lqcs_state_norm_sq(device(), &s));
    return std::sqrt(s);
}

double Statevector::max_abs_diff(const Statevector& other) const {
    double d = 0.0;
    check(lqcs_state_max_abs_diff(device(), other.device(), &d));
    return d;
}

std::vector<double> Statevector::probabilities() const {
    std::vector<double> p(size());
    check(lqcs_state_probabilities(device(), p.data()));
    return p;
}

double Statevector::expectation_value(std::string_view pauli) const {
    const detail::PauliString p = detail::parse_pauli(pauli, num_qubits_);
    double v = 0.0;
    check(lqcs_state_expectation_pauli(device(), p.x, p.zy, p.n_y, &v));
    return v;
}

// ---- StatevectorSimulator ------------------------------------------------------------
namespace {

lqcs_exec_options exec_options(const EngineOptions& e) {
    lqcs_exec_options o{};
    o.plan = static_cast<int32_t>(e.plan);
    o.tile_qubits = e.tile_qubits;
    o.low_qubits = e.low_qubits;
    o.tile_kernel = e.tile_kernel;
    return o;
}

// evolve (statevector_simulator.cpp:197-207): Measure/Barrier skipped, Reset rejected.
Statevector evolve(const QuantumCircuit& c, const EngineOptions& eng) {
    detail::ProgramBuilder b;
    for (const Instruction& in : c.instructions()) {
        if (in.gate.type == GateType::Reset)
            throw std::invalid_argument("run_statevector: Reset is non-deterministic; use run() instead");
        if (in.gate.is_unitary()) b.add_instruction(in);
    }
    Statevector sv(c.num_qubits(), eng.max_qubits);
    const lqcs_program prog = b.program();
    const lqcs_exec_options o = exec_options(eng);
    check(lqcs_state_apply(sv.device(), &prog, &o));
    sv.mark_device_newer();
    return sv;
}

std::string key_of(const std::vector<std::optional<qubit_t>>& src, uint64_t idx) {
    std::string key(src.size(), '0');
    for (std::size_t c = 0; c < src.size(); ++c)
        if (src[c] && ((idx >> *src[c]) & 1u)) key[src.size() - 1 - c] = '1';
    return key;
}

}  // namespace

QuantumCircuit StatevectorSimulator::maybe_fuse(const QuantumCircuit& circuit) const {
    const QuantumCircuit c = options_.engine.eliminate_redundant ? RedundantGateElimination().run(circuit) : circuit;
    if (!options_.fuse_gates) return c;
    return SingleQubitGateFusion().run(c);
}

Statevector StatevectorSimulator::run_statevector(const QuantumCircuit& circuit) const {
    if (!options_.noise.empty())
        throw std::invalid_argument(
            "run_statevector: noisy simulation is stochastic (one trajectory is not the quantum state); use "
            "run() or DensityMatrixSimulator");
    return evolve(maybe_fuse(circuit), options_.engine);
}

std::vector<std::uint64_t> StatevectorSimulator::sample_indices(const QuantumCircuit& circuit,
                                                                std::size_t shots) const {
    if (shots == 0) throw std::invalid_argument("run: shots must be >= 1");
    const Statevector sv = evolve(maybe_fuse(circuit), options_.engine);
    Rng rng(options_.seed);
    std::vector<double> draws(shots);
    for (double& u : draws) u = rng.uniform();
    std::vector<std::uint64_t> idx(shots);
    check(lqcs_state_sample(sv.device(), draws.data(), shots, idx.data()));
    return idx;
}

Result StatevectorSimulator::run(const QuantumCircuit& circuit, std::size_t shots) {
    if (shots == 0) throw std::invalid_argument("run: shots must be >= 1");
    const bool noisy = !options_.noise.empty();
    // Noise is attached to gates, so fusion would move it: off when noisy.
    const QuantumCircuit opt = noisy ? circuit : maybe_fuse(circuit);

    bool seen_measure = false, per_shot = noisy;
    for (const Instruction& in : opt.instructions()) {
        if (in.gate.type == GateType::Measure) seen_measure = true;
        else if (in.gate.type == GateType::Reset) per_shot = true;
        else if (seen_measure && in.gate.is_unitary()) per_shot = true;
    }
    if (!seen_measure)
        throw std::invalid_argument("run: circuit has no measurements; use run_statevector() to inspect the final state");

    Rng rng(options_.seed);
    const std::size_t ncl = opt.num_clbits();
    std::map<std::string, std::size_t> counts;

    if (!per_shot) {
        // run_sampling (:88-119): last Measure on a clbit wins.
        std::vector<std::optional<qubit_t>> src(ncl);
        for (const Instruction& in : opt.instructions())
            if (in.gate.type == GateType::Measure) src[in.clbits[0]] = in.qubits[0];
        const Statevector sv = evolve(opt, options_.engine);
        std::vector<double> draws(shots);
        for (double& u : draws) u = rng.uniform();
        std::vector<uint64_t> idx(shots);
        check(lqcs_state_sample(sv.device(), draws.data(), shots, idx.data()));
        for (uint64_t i : idx) ++counts[key_of(src, i)];
        return Result(std::move(counts), shots);
    }

    // Per-shot trajectories (:156-190).
    detail::ProgramBuilder b;
    std::vector<lqcs_traj_step> steps;
    std::vector<uint32_t> noise_qubits;
    uint64_t D = 0;
    const uint32_t nch = static_cast<uint32_t>(options_.noise.channels().size());
    for (const Instruction& in : opt.instructions()) {
        lqcs_traj_step st{};
        if (in.gate.type == GateType::Measure) {
            st.kind = LQCS_TOP_MEASURE;
            st.qubit = in.qubits[0];
            st.clbit = in.clbits[0];
            D += 1;
        } else if (in.gate.type == GateType::Reset) {
            st.kind = LQCS_TOP_RESET;
            st.qubit = in.qubits[0];
            D += 1;
        } else if (in.gate.type == GateType::Barrier) {
            continue;
        } else {
            st.kind = LQCS_TOP_UNITARY;
            st.op_index = b.add_instruction(in);
            st.noise_qubit_offset = static_cast<uint32_t>(noise_qubits.size());
            st.n_noise_qubits = nch ? static_cast<uint32_t>(in.qubits.size()) : 0;
            if (nch) noise_qubits.insert(noise_qubits.end(), in.qubits.begin(), in.qubits.end());
            D += uint64_t(nch) * st.n_noise_qubits;
        }
        steps.push_back(st);
    }
    std::vector<double> draws(D * shots);
    for (double& u : draws) u = rng.uniform();
    std::vector<std::vector<double>> kraus(nch);
    std::vector<lqcs_channel> chans(nch);
    for (uint32_t c = 0; c < nch; ++c) {
        for (const auto& k : options_.noise.channels()[c].ops)
            for (const complex_t& z : k) {
                kraus[c].push_back(z.real());
                kraus[c].push_back(z.imag());
            }
        chans[c].n_ops = static_cast<uint32_t>(options_.noise.channels()[c].ops.size());
        chans[c].ops = kraus[c].data();
    }
    std::vector<uint8_t> clbits(shots * ncl);
    const lqcs_program prog = b.program();
    check(lqcs_trajectories_run(static_cast<uint32_t>(opt.num_qubits()), static_cast<uint32_t>(ncl), &prog,
                                steps.data(), steps.size(), noise_qubits.data(), chans.data(), nch, draws.data(),
                                D, shots, options_.engine.max_batch,
                                static_cast<uint32_t>(options_.engine.trajectory_kernel), clbits.data()));
    std::string key(ncl, '0');
    for (std::size_t s = 0; s < shots; ++s) {
        for (std::size_t c = 0; c < ncl; ++c) key[ncl - 1 - c] = clbits[s * ncl + c] ? '1' : '0';
        ++counts[key];
    }
    return Result(std::move(counts), shots);
}

// ---- CompiledCircuit ---------------------------------------------------------------
CompiledCircuit::CompiledCircuit(const QuantumCircuit& circuit, bool fuse_gates)
    : prog_(std::make_shared<detail::ProgramBuilder>()), nq_(circuit.num_qubits()) {
    const QuantumCircuit c = fuse_gates ? SingleQubitGateFusion().run(circuit) : circuit;
    for (const Instruction& in : c.instructions()) {
        if (in.gate.type == GateType::Reset)
            throw std::invalid_argument("run_statevector: Reset is non-deterministic; use run() instead");
        if (in.gate.is_unitary()) prog_->add_instruction(in);
    }
}

CompiledCircuit::~CompiledCircuit() = default;

std::size_t CompiledCircuit::num_ops() const { return prog_->size(); }

EngineOptions CompiledCircuit::resolve(const EngineOptions& e) const {
    if (e.plan != EngineOptions::Plan::Auto || !tuned_) return e;
    EngineOptions t = *tuned_;
    t.max_qubits = e.max_qubits;
    t.max_batch = e.max_batch;
    t.trajectory_kernel = e.trajectory_kernel;
    return t;
}

EngineOptions CompiledCircuit::tuned() const { return tuned_ ? *tuned_ : EngineOptions{}; }

CompiledCircuit::TuneResult CompiledCircuit::autotune(Statevector& scratch, int reps) {
    if (scratch.num_qubits() != nq_) throw std::invalid_argument("autotune: register width differs from the circuit");
    reps = std::max(1, reps);
    std::vector<EngineOptions> cands;
    auto add = [&](EngineOptions::Plan plan, int low, int kernel, int width = 0) {
        EngineOptions e;
        e.plan = plan;
        e.low_qubits = low;
        e.tile_kernel = kernel;
        e.tile_qubits = width;
        cands.push_back(e);
    };
    for (auto plan : {EngineOptions::Plan::Greedy, EngineOptions::Plan::Gain})
        for (int low : {2, 3, 4}) add(plan, low, 2);
    for (int kernel : {1, 3, 4}) add(EngineOptions::Plan::Greedy, 0, kernel);
    for (int k : {4, 5, 6})
        if (static_cast<std::size_t>(k) <= nq_) add(EngineOptions::Plan::Dense, 0, 0, k);
    // Specialized passes (their kernels compile during the untimed apply).
    if (nq_ >= 12) {
        add(EngineOptions::Plan::Greedy, 4, 5, 12);
        add(EngineOptions::Plan::Gain, 3, 5, 12);
        if (nq_ >= 17) add(EngineOptions::Plan::Gain, 3, 5, 13);
    }
    TuneResult r;
    for (const EngineOptions& e : cands) {
        try {
            apply(scratch, e);  // untimed
        } catch (const std::invalid_argument&) {
            if (e.tile_qubits == 13) continue;  // 13-qubit tiles refuse Pauli rotations wider than 4 qubits
            throw;
        }
        double best = -1.0;
        for (int k = 0; k < reps; ++k) {
            synchronize();
            profile_enable(true);
            (void)profile_read();
            apply(scratch, e);
            const LaunchProfile p = profile_read();
            profile_enable(false);
            const double ms = p.tile_ms + p.other_ms;
            if (best < 0.0 || ms < best) best = ms;
        }
        r.trials.push_back({e, best});
        if (r.trials.size() == 1 || best < r.best_ms) {
            r.best_ms = best;
            r.best = e;
        }
    }
    tuned_ = std::make_shared<EngineOptions>(r.best);
    return r;
}

CircuitGraph::CircuitGraph(const CompiledCircuit& circuit, Statevector& sv, const EngineOptions& engine) : sv_(&sv) {
    if (sv.num_qubits() != circuit.nq_) throw std::invalid_argument("CircuitGraph: register width differs from the circuit");
    const lqcs_program prog = circuit.prog_->program();
    const lqcs_exec_options o = exec_options(circuit.resolve(engine));
    check(lqcs_graph_create(sv.device(), &prog, &o, &graph_));
}

CircuitGraph::~CircuitGraph() { (void)lqcs_graph_destroy(graph_); }

void CircuitGraph::launch() {
    check(lqcs_graph_launch(graph_, sv_->device()));
    sv_->mark_device_newer();
}

void CompiledCircuit::apply(Statevector& sv, const EngineOptions& engine) const {
    if (sv.num_qubits() != nq_) throw std::invalid_argument("apply: register width differs from the circuit");
    const lqcs_program prog = prog_->program();
    const lqcs_exec_options o = exec_options(resolve(engine));
    check(lqcs_state_apply(sv.device(), &prog, &o));
    sv.mark_device_newer();
}

std::vector<CompiledCircuit::PassInfo> CompiledCircuit::describe(const EngineOptions& engine) const {
    const lqcs_program prog = prog_->program();
    const lqcs_exec_options o = exec_options(engine);
    uint64_t n = 0;
    check(lqcs_plan_describe(static_cast<uint32_t>(nq_), &prog, &o, &n, nullptr, nullptr, nullptr, 0));
    std::vector<uint64_t> ops(n), masks(n), phases(n);
    check(lqcs_plan_describe(static_cast<uint32_t>(nq_), &prog, &o, &n, ops.data(), masks.data(), phases.data(), n));
    std::vector<PassInfo> out(n);
    for (uint64_t i = 0; i < n; ++i) out[i] = {ops[i], masks[i], phases[i]};
    return out;
}

CompiledCircuit::SpecializedPass CompiledCircuit::specialized_pass(std::uint64_t step, const EngineOptions& engine) const {
    const lqcs_program prog = prog_->program();
    const lqcs_exec_options o = exec_options(engine);
    size_t len = 0;
    uint64_t nw = 0;
    check(lqcs_plan_specialized_pass(static_cast<uint32_t>(nq_), &prog, &o, step, nullptr, 0, &len, nullptr, 0, &nw));
    SpecializedPass out;
    std::vector<char> text(len + 1);
    out.words.resize(2 * nw);
    check(lqcs_plan_specialized_pass(static_cast<uint32_t>(nq_), &prog, &o, step, text.data(), text.size(), &len,
                                     out.words.data(), nw, &nw));
    out.text.assign(text.data(), len);
    return out;
}

std::vector<double> CompiledCircuit::features(const EngineOptions& engine, std::uint32_t batch) const {
    const lqcs_program prog = prog_->program();
    const lqcs_exec_options o = exec_options(engine);
    uint64_t n = 0;
    check(lqcs_plan_features(static_cast<uint32_t>(nq_), &prog, &o, batch, nullptr, 0, &n));
    std::vector<double> f(n * kCostFeatures);
    check(lqcs_plan_features(static_cast<uint32_t>(nq_), &prog, &o, batch, f.data(), n, &n));
    return f;
}

void synchronize() { check(lqcs_dev_synchronize()); }

LaunchTimes profile_read_launches() {
    uint64_t n = 0;
    LaunchTimes t;
    // one read drains the recording: size generously, then trim
    t.ms.resize(1 << 20);
    t.tiled.resize(1 << 20);
    check(lqcs_profile_read_launches(t.ms.data(), t.tiled.data(), t.ms.size(), &n));
    if (n > t.ms.size()) throw std::length_error("profile_read_launches: more than 2^20 launches recorded");
    t.ms.resize(n);
    t.tiled.resize(n);
    return t;
}

std::uint64_t mlp_num_params(std::uint32_t n_features) {
    uint64_t n = 0;
    check(lqcs_mlp_num_params(n_features, &n));
    return n;
}

void install_cost_model(const CostModel& m, int tile_kernel) {
    const int32_t slot = tile_kernel == LQCS_TILE_SPECIALIZED ? LQCS_TILE_SPECIALIZED : LQCS_TILE_REGISTER;
    if (m.params.empty()) {
        check(lqcs_costmodel_install_kernel(slot, kCostFeatures, nullptr, nullptr, nullptr, 0.0, 1.0));
        return;
    }
    if (m.params.size() != mlp_num_params(kCostFeatures) || m.x_mean.size() != kCostFeatures ||
        m.x_scale.size() != kCostFeatures)
        throw std::invalid_argument("install_cost_model: sizes do not match the feature count");
    check(lqcs_costmodel_install_kernel(slot, kCostFeatures, m.params.data(), m.x_mean.data(), m.x_scale.data(),
                                        m.t_mean, m.t_scale));
}

CostModel load_cost_model(const std::string& path) {
    std::ifstream in(path);
    if (!in) throw std::invalid_argument("load_cost_model: cannot open " + path);
    std::string tag;
    int version = 0;
    std::uint32_t F = 0;
    std::uint64_t np = 0;
    in >> tag >> version >> F >> np;
    if (!in || tag != "lqcs-costmodel" || version != 1 || F != kCostFeatures || np != mlp_num_params(F))
        throw std::invalid_argument("load_cost_model: not a version-1 model with " + std::to_string(kCostFeatures) +
                                    " features: " + path);
    CostModel m;
    m.x_mean.resize(F);
    m.x_scale.resize(F);
    m.params.resize(np);
    for (double& v : m.x_mean) in >> v;
    for (double& v : m.x_scale) in >> v;
    in >> m.t_mean >> m.t_scale;
    for (double& v : m.params) in >> v;
    if (!in) throw std::invalid_argument("load_cost_model: truncated file " + path);
    return m;
}

PlanChoice last_plan_choice() {
    int32_t plan = 0, low = 0, L = 0;
    double ms = 0.0;
    check(lqcs_last_plan_choice(&plan, &low, &ms));
    check(lqcs_last_plan_tile_qubits(&L));
    return {static_cast<EngineOptions::Plan>(plan), low, ms, L};
}

void profile_enable(bool on) { check(lqcs_profile_enable(on ? 1 : 0)); }

LaunchProfile profile_read() {
    LaunchProfile p;
    check(lqcs_profile_read(&p.tile_launches, &p.tile_ms, &p.other_launches, &p.other_ms, &p.tile_amps));
    return p;
}

// ---- greedy single-qubit fusion (src/transpiler/gate_fusion.cpp:32-68) ----------------
QuantumCircuit SingleQubitGateFusion::run(const QuantumCircuit& circuit) const {
    using M2 = std::array<complex_t, 4>;
    const std::size_t nq = circuit.num_qubits();
    QuantumCircuit out(nq, circuit.num_clbits());
    std::vector<std::optional<M2>> pending(nq);
    auto near_identity = [](const M2& m) {
        return std::abs(m[0] - 1.0) < 1e-12 && std::abs(m[1]) < 1e-12 && std::abs(m[2]) < 1e-12 &&
               std::abs(m[3] - 1.0) < 1e-12;
    };
    auto flush = [&](qubit_t q) {
        if (!pending[q]) return;
        if (!near_identity(*pending[q]))
            out.append({Gate{GateType::Unitary, {}, 0, {pending[q]->begin(), pending[q]->end()}}, {q}, {}});
        pending[q].reset();
    };
    for (const Instruction& in : circuit.instructions()) {
        const Gate& g = in.gate;
        if (g.is_unitary() && g.n_controls == 0 && g.base_qubits() == 1) {
            const auto m = g.base_matrix();
            const M2 a{m[0], m[1], m[2], m[3]};
            std::optional<M2>& acc = pending[in.qubits[0]];
            if (!acc) {
                acc = a;
            } else {  // the later gate multiplies from the left
                const M2& b = *acc;
                acc = M2{a[0] * b[0] + a[1] * b[2], a[0] * b[1] + a[1] * b[3], a[2] * b[0] + a[3] * b[2],
                         a[2] * b[1] + a[3] * b[3]};
            }
            continue;
        }
        if (g.type == GateType::Barrier) {
            for (qubit_t q = 0; q < nq; ++q) flush(q);
        } else {
            for (qubit_t q : in.qubits) flush(q);
        }
        out.append(in);
    }
    for (qubit_t q = 0; q < nq; ++q) flush(q);
    return out;
}

}  // namespace lqcs
