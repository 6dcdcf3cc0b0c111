// _pylqcs: the Python face of the engine, with the names and call shapes of
// the reference's pylqcs bindings (python/src/bindings.cpp:31-444) for the
// statevector path, plus engine extensions (adjoint gradients, sample
// indices, plan statistics). Long-running calls release the GIL.
#include <pybind11/complex.h>
#include <pybind11/functional.h>
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <algorithm>
#include <cstring>
#include <optional>
#include <vector>

#include "lqcs/lqcs.hpp"
#include "lqcs_device.h"

namespace py = pybind11;
using namespace py::literals;
using namespace lqcs;

namespace {

using CArray = py::array_t<std::complex<double>, py::array::c_style | py::array::forcecast>;

std::vector<complex_t> to_matrix(const CArray& a) {
    const auto* p = static_cast<const complex_t*>(a.data());
    return {p, p + a.size()};
}

algorithms::Hamiltonian to_h(const std::vector<std::pair<double, std::string>>& terms) {
    algorithms::Hamiltonian h;
    for (const auto& [c, p] : terms) h.push_back({c, p});
    return h;
}

EngineOptions::Plan plan_of(const std::string& s) {
    if (s == "gate_by_gate") return EngineOptions::Plan::GateByGate;
    if (s == "greedy") return EngineOptions::Plan::Greedy;
    if (s == "learned") return EngineOptions::Plan::Learned;
    if (s == "auto") return EngineOptions::Plan::Auto;
    if (s == "gain") return EngineOptions::Plan::Gain;
    if (s == "dense") return EngineOptions::Plan::Dense;
    throw std::invalid_argument("plan must be gate_by_gate, greedy, gain, dense, learned or auto");
}

int trajectory_kernel_of(const std::string& s) {
    if (s == "auto") return 0;
    if (s == "shared") return 1;
    if (s == "persistent") return 2;
    throw std::invalid_argument("trajectory_kernel must be auto, shared or persistent");
}

int tile_kernel_of(const std::string& s) {
    if (s == "default") return 0;
    if (s == "lds") return 1;
    if (s == "register") return 2;
    if (s == "stream") return 3;
    if (s == "land") return 4;
    if (s == "specialized") return 5;
    throw std::invalid_argument("tile_kernel must be default, lds, register, stream, land or specialized");
}

py::dict options_dict(const EngineOptions& e) {
    static const char* plans[] = {"gate_by_gate", "greedy", "learned", "auto", "gain", "dense"};
    static const char* kernels[] = {"default", "lds", "register", "stream", "land", "specialized"};
    const int p = static_cast<int>(e.plan);
    return py::dict("plan"_a = std::string(p >= 0 && p <= 5 ? plans[p] : "?"), "tile_qubits"_a = e.tile_qubits,
                    "low_qubits"_a = e.low_qubits,
                    "tile_kernel"_a = std::string(e.tile_kernel >= 0 && e.tile_kernel <= 5 ? kernels[e.tile_kernel] : "?"));
}

}  // namespace

PYBIND11_MODULE(_pylqcs, m) {
    m.doc() = "MI355X statevector engine with the pylqcs API";

    py::enum_<GateType>(m, "GateType")
        .value("I", GateType::I).value("X", GateType::X).value("Y", GateType::Y).value("Z", GateType::Z)
        .value("H", GateType::H).value("S", GateType::S).value("Sdg", GateType::Sdg).value("T", GateType::T)
        .value("Tdg", GateType::Tdg).value("SX", GateType::SX).value("RX", GateType::RX)
        .value("RY", GateType::RY).value("RZ", GateType::RZ).value("P", GateType::P).value("U", GateType::U)
        .value("SWAP", GateType::SWAP).value("iSWAP", GateType::iSWAP).value("RXX", GateType::RXX)
        .value("RYY", GateType::RYY).value("RZZ", GateType::RZZ)
        .value("PauliRotation", GateType::PauliRotation).value("Unitary", GateType::Unitary)
        .value("Permutation", GateType::Permutation).value("Measure", GateType::Measure)
        .value("Reset", GateType::Reset).value("Barrier", GateType::Barrier);

    py::class_<Gate>(m, "Gate")
        .def(py::init<GateType, std::vector<double>, std::size_t>(), "type"_a,
             "params"_a = std::vector<double>{}, "n_controls"_a = 0)
        .def_readonly("type", &Gate::type)
        .def_readonly("params", &Gate::params)
        .def_readonly("n_controls", &Gate::n_controls)
        .def("num_qubits", &Gate::num_qubits)
        .def("name", &Gate::name)
        .def("inverse", &Gate::inverse)
        .def("matrix", [](const Gate& g) {
            const auto v = g.matrix();
            const py::ssize_t d = py::ssize_t{1} << g.num_qubits();
            py::array_t<std::complex<double>> a({d, d});
            std::copy(v.begin(), v.end(), static_cast<std::complex<double>*>(a.mutable_data()));
            return a;
        })
        .def("__repr__", [](const Gate& g) { return "<Gate " + g.name() + ">"; });

    auto ri = py::return_value_policy::reference_internal;
    py::class_<QuantumCircuit>(m, "QuantumCircuit")
        .def(py::init<std::size_t, std::size_t>(), "num_qubits"_a, "num_clbits"_a = 0)
        .def("i", &QuantumCircuit::i, ri).def("x", &QuantumCircuit::x, ri).def("y", &QuantumCircuit::y, ri)
        .def("z", &QuantumCircuit::z, ri).def("h", &QuantumCircuit::h, ri).def("s", &QuantumCircuit::s, ri)
        .def("sdg", &QuantumCircuit::sdg, ri).def("t", &QuantumCircuit::t, ri)
        .def("tdg", &QuantumCircuit::tdg, ri).def("sx", &QuantumCircuit::sx, ri)
        .def("rx", &QuantumCircuit::rx, ri).def("ry", &QuantumCircuit::ry, ri).def("rz", &QuantumCircuit::rz, ri)
        .def("p", &QuantumCircuit::p, ri).def("u", &QuantumCircuit::u, ri)
        .def("cx", &QuantumCircuit::cx, ri).def("cy", &QuantumCircuit::cy, ri).def("cz", &QuantumCircuit::cz, ri)
        .def("ch", &QuantumCircuit::ch, ri).def("cp", &QuantumCircuit::cp, ri)
        .def("crx", &QuantumCircuit::crx, ri).def("cry", &QuantumCircuit::cry, ri)
        .def("crz", &QuantumCircuit::crz, ri).def("ccx", &QuantumCircuit::ccx, ri)
        .def("cswap", &QuantumCircuit::cswap, ri)
        .def("mcx", [](QuantumCircuit& qc, std::vector<qubit_t> c, qubit_t t) -> QuantumCircuit& { return qc.mcx(c, t); },
             ri, "controls"_a, "target"_a)
        .def("mcp",
             [](QuantumCircuit& qc, double lam, std::vector<qubit_t> c, qubit_t t) -> QuantumCircuit& {
                 return qc.mcp(lam, c, t);
             },
             ri, "lam"_a, "controls"_a, "target"_a)
        .def("swap", &QuantumCircuit::swap, ri).def("iswap", &QuantumCircuit::iswap, ri)
        .def("rxx", &QuantumCircuit::rxx, ri).def("ryy", &QuantumCircuit::ryy, ri)
        .def("rzz", &QuantumCircuit::rzz, ri)
        .def("rp",
             [](QuantumCircuit& qc, double th, std::string_view p, std::vector<qubit_t> q) -> QuantumCircuit& {
                 return qc.rp(th, p, q);
             },
             ri, "theta"_a, "pauli"_a, "qubits"_a)
        .def("unitary",
             [](QuantumCircuit& qc, const CArray& mat, std::vector<qubit_t> q) -> QuantumCircuit& {
                 return qc.unitary(to_matrix(mat), q);
             },
             ri, "matrix"_a, "qubits"_a)
        .def("permutation",
             [](QuantumCircuit& qc, std::vector<std::size_t> table, std::vector<qubit_t> q) -> QuantumCircuit& {
                 return qc.permutation(table, q);
             },
             ri, "table"_a, "qubits"_a)
        .def("controlled_permutation",
             [](QuantumCircuit& qc, std::vector<std::size_t> table, std::vector<qubit_t> controls,
                std::vector<qubit_t> targets) -> QuantumCircuit& {
                 std::vector<qubit_t> qs = controls;
                 qs.insert(qs.end(), targets.begin(), targets.end());
                 return qc.append({Gate{GateType::Permutation, {}, controls.size(), {}, table}, qs, {}});
             },
             ri, "table"_a, "controls"_a, "targets"_a)
        .def("controlled_unitary",
             [](QuantumCircuit& qc, const CArray& mat, std::vector<qubit_t> controls,
                std::vector<qubit_t> targets) -> QuantumCircuit& {
                 std::vector<qubit_t> qs = controls;
                 qs.insert(qs.end(), targets.begin(), targets.end());
                 return qc.append({Gate{GateType::Unitary, {}, controls.size(), to_matrix(mat)}, qs, {}});
             },
             ri, "matrix"_a, "controls"_a, "targets"_a)
        .def("measure", &QuantumCircuit::measure, ri)
        .def("measure_all", &QuantumCircuit::measure_all, ri)
        .def("reset", &QuantumCircuit::reset, ri)
        .def("barrier", &QuantumCircuit::barrier, ri)
        .def("compose",
             [](QuantumCircuit& qc, const QuantumCircuit& o, std::vector<qubit_t> map) -> QuantumCircuit& {
                 return qc.compose(o, map);
             },
             ri, "other"_a, "qubit_map"_a = std::vector<qubit_t>{})
        .def("inverse", &QuantumCircuit::inverse)
        .def("power", &QuantumCircuit::power)
        .def("num_qubits", &QuantumCircuit::num_qubits)
        .def("num_clbits", &QuantumCircuit::num_clbits)
        .def("num_instructions", &QuantumCircuit::num_instructions)
        .def("depth", &QuantumCircuit::depth)
        .def("draw", &QuantumCircuit::draw)
        .def("__str__", &QuantumCircuit::draw)
        .def("__repr__", [](const QuantumCircuit& qc) {
            return "<QuantumCircuit qubits=" + std::to_string(qc.num_qubits()) +
                   " clbits=" + std::to_string(qc.num_clbits()) +
                   " instructions=" + std::to_string(qc.num_instructions()) + ">";
        });

    py::class_<Statevector>(m, "Statevector")
        .def(py::init<std::size_t>(), "num_qubits"_a)
        .def("num_qubits", &Statevector::num_qubits)
        .def("norm", &Statevector::norm)
        .def("max_abs_diff", &Statevector::max_abs_diff, "other"_a, "max |a_i - b_i|, on the device")
        .def("__len__", &Statevector::size)
        .def("__getitem__",
             [](const Statevector& sv, std::size_t i) {
                 if (i >= sv.size()) throw py::index_error();
                 return static_cast<std::complex<double>>(sv[i]);
             })
        .def("probabilities",
             [](const Statevector& sv) {
                 const auto p = sv.probabilities();
                 py::array_t<double> a(static_cast<py::ssize_t>(p.size()));
                 std::copy(p.begin(), p.end(), a.mutable_data());
                 return a;
             })
        .def("expectation_value", &Statevector::expectation_value, "pauli"_a)
        .def("set_basis", &Statevector::set_basis, "index"_a)
        .def("copy_from",
             [](Statevector& dst, const Statevector& src) {
                 if (dst.num_qubits() != src.num_qubits()) throw std::invalid_argument("copy_from: widths differ");
                 if (lqcs_state_copy(dst.device(), src.device()) != LQCS_OK)
                     throw std::runtime_error(lqcs_dev_last_error());
                 dst.mark_device_newer();
             },
             "source"_a)
        .def("to_numpy", [](py::object self) {
            auto& sv = self.cast<Statevector&>();
            return py::array_t<std::complex<double>>({static_cast<py::ssize_t>(sv.size())},
                                                     {static_cast<py::ssize_t>(sizeof(complex_t))},
                                                     reinterpret_cast<std::complex<double>*>(sv.data()), self);
        });

    py::class_<Result>(m, "Result")
        .def("counts", &Result::counts)
        .def("probabilities", &Result::probabilities)
        .def("shots", &Result::shots)
        .def("__repr__", [](const Result& r) {
            std::string s = "<Result shots=" + std::to_string(r.shots()) + " {";
            bool first = true;
            for (const auto& [k, v] : r.counts()) {
                s += (first ? "'" : ", '") + k + "': " + std::to_string(v);
                first = false;
            }
            return s + "}>";
        });

    py::class_<KrausChannel>(m, "KrausChannel")
        .def(py::init([](const std::vector<CArray>& ops) {
                 std::vector<std::array<complex_t, 4>> k;
                 for (const auto& a : ops) {
                     if (a.size() != 4) throw std::invalid_argument("KrausChannel: operators must be 2x2");
                     const auto v = to_matrix(a);
                     k.push_back({v[0], v[1], v[2], v[3]});
                 }
                 return KrausChannel(std::move(k));
             }),
             "ops"_a)
        .def_static("depolarizing", &KrausChannel::depolarizing, "p"_a)
        .def_static("amplitude_damping", &KrausChannel::amplitude_damping, "gamma"_a)
        .def_static("phase_damping", &KrausChannel::phase_damping, "gamma"_a)
        .def_static("bit_flip", &KrausChannel::bit_flip, "p"_a)
        .def_static("phase_flip", &KrausChannel::phase_flip, "p"_a)
        .def("ops", [](const KrausChannel& k) {
            std::vector<std::vector<std::complex<double>>> out;
            for (const auto& o : k.ops) out.push_back({o[0], o[1], o[2], o[3]});
            return out;
        });

    py::class_<NoiseModel>(m, "NoiseModel")
        .def(py::init<>())
        .def("add_all_qubit_channel", &NoiseModel::add_all_qubit_channel, ri, "channel"_a)
        .def("empty", &NoiseModel::empty);

    py::class_<StatevectorSimulator>(m, "StatevectorSimulator")
        .def(py::init([](std::optional<std::uint64_t> seed, bool fuse_gates, int num_threads, NoiseModel noise,
                         const std::string& plan, int tile_qubits, int low_qubits, std::size_t max_qubits,
                         std::uint32_t max_batch, bool eliminate_redundant, const std::string& trajectory_kernel,
                         const std::string& tile_kernel) {
                 StatevectorSimulator::Options o;
                 o.engine.trajectory_kernel = trajectory_kernel_of(trajectory_kernel);
                 o.engine.tile_kernel = tile_kernel_of(tile_kernel);
                 o.engine.eliminate_redundant = eliminate_redundant;
                 o.seed = seed;
                 o.fuse_gates = fuse_gates;
                 o.num_threads = num_threads;
                 o.noise = std::move(noise);
                 o.engine.plan = plan_of(plan);
                 o.engine.tile_qubits = tile_qubits;
                 o.engine.low_qubits = low_qubits;
                 o.engine.max_qubits = max_qubits;
                 o.engine.max_batch = max_batch;
                 return StatevectorSimulator(std::move(o));
             }),
             "seed"_a = py::none(), "fuse_gates"_a = true, "num_threads"_a = 0, "noise"_a = NoiseModel{},
             "plan"_a = "auto", "tile_qubits"_a = 0, "low_qubits"_a = 0, "max_qubits"_a = 30, "max_batch"_a = 0,
             "eliminate_redundant"_a = false, "trajectory_kernel"_a = "auto", "tile_kernel"_a = "default")
        .def("run", &StatevectorSimulator::run, "circuit"_a, "shots"_a = 1024,
             py::call_guard<py::gil_scoped_release>())
        .def("run_statevector", &StatevectorSimulator::run_statevector, "circuit"_a,
             py::call_guard<py::gil_scoped_release>())
        .def("sample_indices",
             [](const StatevectorSimulator& s, const QuantumCircuit& c, std::size_t shots) {
                 std::vector<std::uint64_t> v;
                 {
                     py::gil_scoped_release nogil;
                     v = s.sample_indices(c, shots);
                 }
                 py::array_t<std::uint64_t> a(static_cast<py::ssize_t>(v.size()));
                 std::copy(v.begin(), v.end(), a.mutable_data());
                 return a;
             },
             "circuit"_a, "shots"_a = 1024)
        .def("name", &StatevectorSimulator::name);

    // Circuit passes (C++ API of the reference's include/lqcs/transpiler/*.hpp).
    py::class_<SingleQubitGateFusion>(m, "SingleQubitGateFusion")
        .def(py::init<>())
        .def("run", &SingleQubitGateFusion::run, "circuit"_a)
        .def("name", &SingleQubitGateFusion::name);
    py::class_<RedundantGateElimination>(m, "RedundantGateElimination")
        .def(py::init<>())
        .def("run", &RedundantGateElimination::run, "circuit"_a)
        .def("name", &RedundantGateElimination::name);

    // OpenQASM 2.0 (reference python/src/bindings.cpp:410-413).
    auto io_mod = m.def_submodule("io", "OpenQASM import/export");
    io_mod.def("to_qasm", &io::to_qasm, "circuit"_a);
    io_mod.def("from_qasm", [](const std::string& src) { return io::from_qasm(src); }, "source"_a);

    // Density-matrix backend (reference python/src/bindings.cpp:416-443), on the device.
    py::class_<DensityMatrix>(m, "DensityMatrix")
        .def(py::init<std::size_t>(), "num_qubits"_a)
        .def(py::init<std::size_t, std::size_t>(), "num_qubits"_a, "max_qubits"_a)
        .def("num_qubits", &DensityMatrix::num_qubits)
        .def("dim", &DensityMatrix::dim)
        .def("element", &DensityMatrix::element, "row"_a, "col"_a)
        .def("apply_channel", &DensityMatrix::apply_channel, "channel"_a, "qubit"_a,
             py::call_guard<py::gil_scoped_release>())
        .def("trace", &DensityMatrix::trace, py::call_guard<py::gil_scoped_release>())
        .def("purity", &DensityMatrix::purity, py::call_guard<py::gil_scoped_release>())
        .def("fidelity", &DensityMatrix::fidelity, "psi"_a, py::call_guard<py::gil_scoped_release>())
        .def("probabilities", [](const DensityMatrix& rho) {
            std::vector<double> p;
            {
                py::gil_scoped_release nogil;
                p = rho.probabilities();
            }
            py::array_t<double> a(static_cast<py::ssize_t>(p.size()));
            std::copy(p.begin(), p.end(), a.mutable_data());
            return a;
        })
        // Engine extension: rho as a dim x dim matrix (a copy).
        .def("to_numpy", [](const DensityMatrix& rho) {
            const auto d = static_cast<py::ssize_t>(rho.dim());
            py::array_t<std::complex<double>> a({d, d});
            {
                py::gil_scoped_release nogil;
                // Device order is r | (c << n): row-major in (c, r), i.e. the transpose.
                if (lqcs_state_read(rho.device(), 0, static_cast<uint64_t>(d) * d,
                                    reinterpret_cast<double*>(a.mutable_data())) != LQCS_OK)
                    throw std::runtime_error(lqcs_dev_last_error());
            }
            return a.attr("T").attr("copy")();
        });

    py::class_<DensityMatrixSimulator>(m, "DensityMatrixSimulator")
        .def(py::init([](std::optional<std::uint64_t> seed, NoiseModel noise, const std::string& plan,
                         int tile_qubits, int low_qubits, std::size_t max_qubits, const std::string& tile_kernel) {
                 DensityMatrixSimulator::Options o;
                 o.seed = seed;
                 o.noise = std::move(noise);
                 o.engine.plan = plan_of(plan);
                 o.engine.tile_qubits = tile_qubits;
                 o.engine.low_qubits = low_qubits;
                 o.engine.tile_kernel = tile_kernel_of(tile_kernel);
                 o.max_qubits = max_qubits;
                 return DensityMatrixSimulator(std::move(o));
             }),
             "seed"_a = py::none(), "noise"_a = NoiseModel{}, "plan"_a = "auto", "tile_qubits"_a = 0,
             "low_qubits"_a = 0, "max_qubits"_a = DensityMatrix::kMaxQubits, "tile_kernel"_a = "default")
        .def("run", &DensityMatrixSimulator::run, "circuit"_a, "shots"_a = 1024,
             py::call_guard<py::gil_scoped_release>())
        .def("run_density_matrix", &DensityMatrixSimulator::run_density_matrix, "circuit"_a,
             py::call_guard<py::gil_scoped_release>())
        .def("name", &DensityMatrixSimulator::name);

    py::class_<ParametricCircuit>(m, "ParametricCircuit")
        .def(py::init<std::size_t, std::size_t>(), "num_qubits"_a, "num_clbits"_a = 0)
        .def("x", &ParametricCircuit::x, ri).def("y", &ParametricCircuit::y, ri)
        .def("z", &ParametricCircuit::z, ri).def("h", &ParametricCircuit::h, ri)
        .def("s", &ParametricCircuit::s, ri).def("sdg", &ParametricCircuit::sdg, ri)
        .def("t", &ParametricCircuit::t, ri).def("sx", &ParametricCircuit::sx, ri)
        .def("cx", &ParametricCircuit::cx, ri).def("cy", &ParametricCircuit::cy, ri)
        .def("cz", &ParametricCircuit::cz, ri).def("swap", &ParametricCircuit::swap, ri)
        .def("ccx", &ParametricCircuit::ccx, ri).def("barrier", &ParametricCircuit::barrier, ri)
        .def("rx", &ParametricCircuit::rx, "qubit"_a)
        .def("ry", &ParametricCircuit::ry, "qubit"_a)
        .def("rz", &ParametricCircuit::rz, "qubit"_a)
        .def("rp", [](ParametricCircuit& pc, std::string_view p, std::vector<qubit_t> q) { return pc.rp(p, q); },
             "pauli"_a, "qubits"_a)
        .def("bind", [](const ParametricCircuit& pc, std::vector<double> v) { return pc.bind(v); }, "values"_a)
        .def("num_parameters", &ParametricCircuit::num_parameters)
        .def("num_qubits", &ParametricCircuit::num_qubits)
        .def("num_clbits", &ParametricCircuit::num_clbits);

    auto alg = m.def_submodule("algorithms", "quantum algorithms on the statevector engine");
    alg.def("qft", &algorithms::qft, "n"_a, "do_swaps"_a = true);
    alg.def("qft_inverse", &algorithms::qft_inverse, "n"_a, "do_swaps"_a = true);
    alg.def(
        "expectation",
        [](const Statevector& sv, const std::vector<std::pair<double, std::string>>& terms) {
            return algorithms::expectation(sv, to_h(terms));
        },
        "statevector"_a, "hamiltonian"_a);
    alg.def(
        "parameter_shift_gradient",
        [](const std::vector<std::pair<double, std::string>>& terms, const ParametricCircuit& ansatz,
           std::vector<double> params) {
            const auto h = to_h(terms);
            py::gil_scoped_release nogil;
            return algorithms::parameter_shift_gradient(h, ansatz, params);
        },
        "hamiltonian"_a, "ansatz"_a, "params"_a);
    alg.def(
        "adjoint_gradient",
        [](const std::vector<std::pair<double, std::string>>& terms, const ParametricCircuit& ansatz,
           std::vector<double> params) {
            const auto h = to_h(terms);
            double e = 0.0;
            std::vector<double> g;
            {
                py::gil_scoped_release nogil;
                g = algorithms::adjoint_gradient(h, ansatz, params, &e);
            }
            return py::make_tuple(e, g);
        },
        "hamiltonian"_a, "ansatz"_a, "params"_a, "returns (energy, gradient)");
    alg.def(
        "adjoint_gradient_batched",
        [](const std::vector<std::vector<std::pair<double, std::string>>>& hams,
           const std::vector<const ParametricCircuit*>& ansatze, const std::vector<std::vector<double>>& params) {
            if (hams.size() != ansatze.size() || params.size() != ansatze.size())
                throw std::invalid_argument("adjoint_gradient_batched: list lengths differ");
            std::vector<algorithms::Hamiltonian> hs;
            for (const auto& t : hams) hs.push_back(to_h(t));
            std::vector<algorithms::GradientJob> jobs;
            for (std::size_t i = 0; i < ansatze.size(); ++i) jobs.push_back({ansatze[i], &hs[i], params[i]});
            std::vector<algorithms::GradientResult> r;
            {
                py::gil_scoped_release nogil;
                r = algorithms::adjoint_gradient_batched(jobs);
            }
            py::list out;
            for (const auto& x : r) out.append(py::make_tuple(x.energy, x.grad));
            return out;
        },
        "hamiltonians"_a, "ansatze"_a, "params"_a, "list of (energy, gradient)");
    py::class_<algorithms::AdjointPlan>(alg, "AdjointPlan",
                                        "Adjoint gradients of fixed ansatz/Hamiltonian pairs at changing "
                                        "parameter values: the device launch sequence is captured once in a "
                                        "graph and replayed by run().")
        .def(py::init([](const std::vector<std::vector<std::pair<double, std::string>>>& hams,
                         const std::vector<const ParametricCircuit*>& ansatze) {
                 if (hams.size() != ansatze.size()) throw std::invalid_argument("AdjointPlan: list lengths differ");
                 std::vector<algorithms::Hamiltonian> hs;
                 for (const auto& t : hams) hs.push_back(to_h(t));
                 std::vector<algorithms::GradientJob> jobs;
                 for (std::size_t i = 0; i < ansatze.size(); ++i)
                     jobs.push_back({ansatze[i], &hs[i], std::vector<double>(ansatze[i]->num_parameters(), 0.0)});
                 py::gil_scoped_release nogil;
                 return std::make_unique<algorithms::AdjointPlan>(jobs);
             }),
             "hamiltonians"_a, "ansatze"_a)
        .def_property_readonly("num_jobs", &algorithms::AdjointPlan::num_jobs)
        .def(
            "run",
            [](algorithms::AdjointPlan& plan, const std::vector<std::vector<double>>& params) {
                std::vector<algorithms::GradientResult> r;
                {
                    py::gil_scoped_release nogil;
                    r = plan.run(params);
                }
                py::list out;
                for (const auto& x : r) out.append(py::make_tuple(x.energy, x.grad));
                return out;
            },
            "params"_a, "list of (energy, gradient), one per job");
    py::class_<algorithms::VQEResult>(alg, "VQEResult")
        .def_readonly("energy", &algorithms::VQEResult::energy)
        .def_readonly("parameters", &algorithms::VQEResult::parameters)
        .def_readonly("iterations", &algorithms::VQEResult::iterations)
        .def_readonly("history", &algorithms::VQEResult::history);
    // The ansatz is a Python callable, so the GIL stays held while vqe runs.
    alg.def(
        "vqe",
        [](const std::vector<std::pair<double, std::string>>& terms,
           const std::function<QuantumCircuit(std::vector<double>)>& ansatz, std::size_t n_params,
           std::size_t max_iterations, double tol, std::vector<double> initial_parameters) {
            const algorithms::Ansatz call = [&ansatz](std::span<const double> th) {
                return ansatz(std::vector<double>(th.begin(), th.end()));
            };
            return algorithms::vqe(to_h(terms), call, n_params,
                                   {max_iterations, tol, std::move(initial_parameters)});
        },
        "hamiltonian"_a, "ansatz"_a, "n_params"_a, "max_iterations"_a = 100, "tol"_a = 1e-9,
        "initial_parameters"_a = std::vector<double>{});
    alg.def(
        "vqe_gradient_descent",
        [](const std::vector<std::pair<double, std::string>>& terms, const ParametricCircuit& ansatz,
           std::size_t max_iterations, double learning_rate, double tol, std::vector<double> initial_parameters) {
            const auto h = to_h(terms);
            py::gil_scoped_release nogil;
            return algorithms::vqe_gradient_descent(
                h, ansatz, {max_iterations, learning_rate, tol, std::move(initial_parameters)});
        },
        "hamiltonian"_a, "ansatz"_a, "max_iterations"_a = 200, "learning_rate"_a = 0.1, "tol"_a = 1e-9,
        "initial_parameters"_a = std::vector<double>{});

    py::class_<CompiledCircuit>(m, "CompiledCircuit")
        .def(py::init<const QuantumCircuit&, bool>(), "circuit"_a, "fuse_gates"_a = true)
        .def("apply",
             [](const CompiledCircuit& c, Statevector& sv, const std::string& plan, int tile_qubits,
                int low_qubits, const std::string& tile_kernel) {
                 EngineOptions e;
                 e.plan = plan_of(plan);
                 e.tile_qubits = tile_qubits;
                 e.low_qubits = low_qubits;
                 e.tile_kernel = tile_kernel_of(tile_kernel);
                 e.max_qubits = LQCS_DEVICE_MAX_QUBITS;
                 py::gil_scoped_release nogil;
                 c.apply(sv, e);
             },
             "statevector"_a, "plan"_a = "auto", "tile_qubits"_a = 0, "low_qubits"_a = 0, "tile_kernel"_a = "default")
        .def("autotune",
             [](CompiledCircuit& c, Statevector& scratch, int reps) {
                 CompiledCircuit::TuneResult r;
                 {
                     py::gil_scoped_release nogil;
                     r = c.autotune(scratch, reps);
                 }
                 py::list trials;
                 for (const auto& t : r.trials) trials.append(py::make_tuple(options_dict(t.options), t.ms));
                 return py::dict("best"_a = options_dict(r.best), "best_ms"_a = r.best_ms, "trials"_a = trials);
             },
             "scratch"_a, "reps"_a = 2,
             "time every candidate tile shape on scratch (overwritten); apply(plan='auto') then uses the fastest")
        .def("tuned", [](const CompiledCircuit& c) { return options_dict(c.tuned()); })
        .def("num_qubits", &CompiledCircuit::num_qubits)
        .def("num_ops", &CompiledCircuit::num_ops)
        .def("describe",
             [](const CompiledCircuit& c, const std::string& plan, int tile_qubits, int low_qubits,
                const std::string& tile_kernel) {
                 EngineOptions e;
                 e.plan = plan_of(plan);
                 e.tile_qubits = tile_qubits;
                 e.low_qubits = low_qubits;
                 e.tile_kernel = tile_kernel_of(tile_kernel);
                 py::list out;
                 for (const auto& p : c.describe(e))
                     out.append(py::dict("ops"_a = p.ops, "tile_mask"_a = p.tile_mask, "phases"_a = p.phases));
                 return out;
             },
             "plan"_a = "auto", "tile_qubits"_a = 0, "low_qubits"_a = 0, "tile_kernel"_a = "default")
        .def("specialized_pass",
             [](const CompiledCircuit& c, std::uint64_t step, const std::string& plan, int tile_qubits,
                int low_qubits, const std::string& tile_kernel) {
                 EngineOptions e;
                 e.plan = plan_of(plan);
                 e.tile_qubits = tile_qubits;
                 e.low_qubits = low_qubits;
                 e.tile_kernel = tile_kernel_of(tile_kernel);
                 const CompiledCircuit::SpecializedPass sp = c.specialized_pass(step, e);
                 py::array_t<double> words({static_cast<py::ssize_t>(sp.words.size() / 2), py::ssize_t{2}});
                 if (!sp.words.empty()) std::memcpy(words.mutable_data(), sp.words.data(), sp.words.size() * sizeof(double));
                 return py::make_tuple(sp.text, words);
             },
             "step"_a, "plan"_a = "auto", "tile_qubits"_a = 0, "low_qubits"_a = 0, "tile_kernel"_a = "default",
             "kernel text and op words of the specialized kernel of one planned pass (no device needed)")
        .def("features",
             [](const CompiledCircuit& c, const std::string& plan, int tile_qubits, int low_qubits,
                std::uint32_t batch, const std::string& tile_kernel) {
                 EngineOptions e;
                 e.plan = plan_of(plan);
                 e.tile_qubits = tile_qubits;
                 e.low_qubits = low_qubits;
                 e.tile_kernel = tile_kernel_of(tile_kernel);
                 std::vector<double> f;
                 {
                     py::gil_scoped_release nogil;
                     f = c.features(e, batch);
                 }
                 const py::ssize_t rows = static_cast<py::ssize_t>(f.size() / kCostFeatures);
                 py::array_t<double> out({rows, static_cast<py::ssize_t>(kCostFeatures)});
                 std::memcpy(out.mutable_data(), f.data(), f.size() * sizeof(double));
                 return out;
             },
             "plan"_a = "auto", "tile_qubits"_a = 0, "low_qubits"_a = 0, "batch"_a = 1, "tile_kernel"_a = "default",
             "cost-model features of every pass of the plan (rows x COST_FEATURES)");
    py::class_<CircuitGraph>(m, "CircuitGraph")
        .def(py::init([](const CompiledCircuit& c, Statevector& sv, const std::string& plan, int tile_qubits,
                         int low_qubits, const std::string& tile_kernel) {
                 EngineOptions e;
                 e.plan = plan_of(plan);
                 e.tile_qubits = tile_qubits;
                 e.low_qubits = low_qubits;
                 e.tile_kernel = tile_kernel_of(tile_kernel);
                 return std::make_unique<CircuitGraph>(c, sv, e);
             }),
             "circuit"_a, "statevector"_a, "plan"_a = "auto", "tile_qubits"_a = 0, "low_qubits"_a = 0,
             "tile_kernel"_a = "default", py::keep_alive<1, 3>())
        .def("launch", [](CircuitGraph& g) {
            py::gil_scoped_release nogil;
            g.launch();
        }, "replay the captured launches (sv <- U sv), queued on the engine stream");
    m.def("statevector", [](std::size_t n) { return Statevector(n, LQCS_DEVICE_MAX_QUBITS); }, "num_qubits"_a,
          "register of up to MAX_DEVICE_QUBITS qubits (beyond the 30-qubit reference cap)");
    m.def("synchronize", [] {
        py::gil_scoped_release nogil;
        synchronize();
    });
    m.def("profile_enable", &profile_enable, "on"_a);
    m.def("set_tile_kernel", [](const std::string& k) {
        const int code = k == "specialized" ? 4 : k == "land" ? 3 : k == "stream" ? 2 : k == "register" ? 1
                         : k == "lds" ? 0 : -1;
        if (code < 0)
            throw std::invalid_argument("tile kernel must be 'register', 'stream', 'land', 'specialized' or 'lds'");
        if (lqcs_set_tile_kernel(code) != LQCS_OK) throw std::runtime_error(lqcs_dev_last_error());
    }, "kernel"_a);
    m.def("specialized_kernels_cached", [] {
        uint64_t n = 0;
        if (lqcs_specialized_kernels(&n) != LQCS_OK) throw std::runtime_error(lqcs_dev_last_error());
        return n;
    }, "specialized pass kernels compiled and loaded in this process");
    m.def("set_dense_mfma_min_targets", [](uint32_t k) {
        if (lqcs_set_dense_mfma_min_targets(k) != LQCS_OK) throw std::invalid_argument(lqcs_dev_last_error());
    }, "min_targets"_a);
    m.def("profile_read", [] {
        const LaunchProfile p = profile_read();
        return py::dict("tile_launches"_a = p.tile_launches, "tile_ms"_a = p.tile_ms,
                        "other_launches"_a = p.other_launches, "other_ms"_a = p.other_ms,
                        "tile_amps"_a = p.tile_amps);
    });

    m.def("profile_read_launches", [] {
        const LaunchTimes t = profile_read_launches();
        py::array_t<double> ms(static_cast<py::ssize_t>(t.ms.size()));
        py::array_t<std::uint8_t> tiled(static_cast<py::ssize_t>(t.tiled.size()));
        std::memcpy(ms.mutable_data(), t.ms.data(), t.ms.size() * sizeof(double));
        std::memcpy(tiled.mutable_data(), t.tiled.data(), t.tiled.size());
        return py::make_tuple(ms, tiled);
    }, "per-launch milliseconds and tiled flags since the last read");

    // ---- learned planning cost model ----
    m.attr("COST_FEATURES") = kCostFeatures;
    m.def("mlp_num_params", &mlp_num_params, "n_features"_a);
    m.def(
        "mlp_forward",
        [](py::array_t<double, py::array::c_style | py::array::forcecast> params,
           py::array_t<double, py::array::c_style | py::array::forcecast> x) {
            if (x.ndim() != 2) throw std::invalid_argument("x must be 2-D");
            const auto rows = static_cast<std::uint64_t>(x.shape(0));
            const auto F = static_cast<std::uint32_t>(x.shape(1));
            if (static_cast<std::uint64_t>(params.size()) != mlp_num_params(F))
                throw std::invalid_argument("params size does not match the feature count");
            py::array_t<double> y(static_cast<py::ssize_t>(rows));
            {
                py::gil_scoped_release nogil;
                if (lqcs_mlp_forward(F, params.data(), x.data(), rows, y.mutable_data()) != LQCS_OK)
                    throw std::runtime_error(lqcs_dev_last_error());
            }
            return y;
        },
        "params"_a, "x"_a, "network output per row, evaluated on the device");
    m.def(
        "mlp_loss_grad",
        [](py::array_t<double, py::array::c_style | py::array::forcecast> params,
           py::array_t<double, py::array::c_style | py::array::forcecast> x,
           py::array_t<double, py::array::c_style | py::array::forcecast> t) {
            if (x.ndim() != 2 || t.size() != x.shape(0)) throw std::invalid_argument("x must be rows x F, t rows");
            const auto rows = static_cast<std::uint64_t>(x.shape(0));
            const auto F = static_cast<std::uint32_t>(x.shape(1));
            if (static_cast<std::uint64_t>(params.size()) != mlp_num_params(F))
                throw std::invalid_argument("params size does not match the feature count");
            py::array_t<double> g(params.size());
            double loss = 0.0;
            {
                py::gil_scoped_release nogil;
                if (lqcs_mlp_loss_grad(F, params.data(), x.data(), t.data(), rows, &loss, g.mutable_data()) != LQCS_OK)
                    throw std::runtime_error(lqcs_dev_last_error());
            }
            return py::make_tuple(loss, g);
        },
        "params"_a, "x"_a, "t"_a, "(mean squared error, gradient), on the device");
    m.def(
        "mlp_train",
        [](py::array_t<double, py::array::c_style | py::array::forcecast> params,
           py::array_t<double, py::array::c_style | py::array::forcecast> x,
           py::array_t<double, py::array::c_style | py::array::forcecast> t, std::uint32_t epochs,
           std::uint32_t batch_rows, double lr, std::uint64_t seed) {
            if (x.ndim() != 2 || t.size() != x.shape(0)) throw std::invalid_argument("x must be rows x F, t rows");
            const auto rows = static_cast<std::uint64_t>(x.shape(0));
            const auto F = static_cast<std::uint32_t>(x.shape(1));
            if (static_cast<std::uint64_t>(params.size()) != mlp_num_params(F))
                throw std::invalid_argument("params size does not match the feature count");
            py::array_t<double> p(params.size());
            std::memcpy(p.mutable_data(), params.data(), params.size() * sizeof(double));
            double loss = 0.0;
            {
                py::gil_scoped_release nogil;
                if (lqcs_mlp_train(F, p.mutable_data(), x.data(), t.data(), rows, epochs, batch_rows, lr, seed,
                                   &loss) != LQCS_OK)
                    throw std::runtime_error(lqcs_dev_last_error());
            }
            return py::make_tuple(p, loss);
        },
        "params"_a, "x"_a, "t"_a, "epochs"_a, "batch_rows"_a = 256, "lr"_a = 1e-3, "seed"_a = 0,
        "Adam on the device; returns (trained params, final loss)");
    m.def(
        "costmodel_install",
        [](std::vector<double> params, std::vector<double> x_mean, std::vector<double> x_scale, double t_mean,
           double t_scale, const std::string& tile_kernel) {
            install_cost_model(CostModel{params, x_mean, x_scale, t_mean, t_scale}, tile_kernel_of(tile_kernel));
        },
        "params"_a, "x_mean"_a, "x_scale"_a, "t_mean"_a, "t_scale"_a, "tile_kernel"_a = "register",
        "install the model plan='learned' consults for passes on this tile kernel (register or specialized)");
    m.def("costmodel_uninstall",
          [](const std::string& tile_kernel) { install_cost_model(CostModel{}, tile_kernel_of(tile_kernel)); },
          "tile_kernel"_a = "register");
    m.def("costmodel_load",
          [](const std::string& path, const std::string& tile_kernel) {
              install_cost_model(load_cost_model(path), tile_kernel_of(tile_kernel));
          },
          "path"_a, "tile_kernel"_a = "register");
    m.def("last_plan_choice", [] {
        const PlanChoice c = last_plan_choice();
        static const char* names[] = {"gate_by_gate", "greedy", "learned", "auto", "gain", "dense"};
        const int k = static_cast<int>(c.plan);
        return py::dict("plan"_a = std::string(k >= 0 && k <= 5 ? names[k] : "?"), "low_qubits"_a = c.low_qubits,
                        "predicted_ms"_a = c.predicted_ms, "tile_qubits"_a = c.tile_qubits);
    });

    m.def("device_info", [] {
        char name[256] = {0};
        int cus = 0;
        std::uint64_t hbm = 0;
        if (lqcs_dev_info(name, sizeof name, &cus, &hbm) != LQCS_OK)
            throw std::runtime_error(lqcs_dev_last_error());
        return py::dict("name"_a = std::string(name), "compute_units"_a = cus, "hbm_bytes"_a = hbm);
    });
    m.def("last_plan_stats", [] {
        std::uint64_t passes = 0, launches = 0;
        lqcs_last_plan_stats(&passes, &launches);
        return py::dict("passes"_a = passes, "launches"_a = launches);
    });
    m.attr("MAX_DEVICE_QUBITS") = LQCS_DEVICE_MAX_QUBITS;
}
