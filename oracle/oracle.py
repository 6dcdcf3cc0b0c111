"""ctypes access to the checkers: the C restatement (oracle/_build) and the
reference compiled from its sources (oracle/_ref).

TEST INFRASTRUCTURE: imported only by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg. The product never imports this module.

`Circuit` records a circuit once and replays it into (a) the oracle's flat
instruction format and (b) the product's pylqcs.QuantumCircuit builder, so a
parity test states each circuit a single time. The builder-to-instruction
mapping restates src/circuit/quantum_circuit.cpp of the reference (which
gate type, how many controls, which qubit order each method records).
"""
from __future__ import annotations

import ctypes as C
import math
import os
from typing import Sequence

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "_build", "liblqcs_oracle.so")
REF_SO = os.path.join(HERE, "_ref", "libref_lqcs.so")

# GateType numbering (include/lqcs/circuit/gate.hpp:14-28)
(I, X, Y, Z, H, S, SDG, T, TDG, SX, RX, RY, RZ, P, U, SWAP, ISWAP, RXX, RYY, RZZ,
 PAULI_ROTATION, UNITARY, PERMUTATION, MEASURE, RESET, BARRIER) = range(26)


class OrcC(C.Structure):
    _fields_ = [("re", C.c_double), ("im", C.c_double)]


class OrcInst(C.Structure):
    _fields_ = [
        ("type", C.c_int32), ("n_controls", C.c_uint32), ("n_qubits", C.c_uint32),
        ("n_params", C.c_uint32), ("qubits", C.POINTER(C.c_uint32)),
        ("params", C.POINTER(C.c_double)), ("mat_len", C.c_uint64), ("mat", C.POINTER(OrcC)),
        ("perm_len", C.c_uint64), ("perm", C.POINTER(C.c_uint64)), ("pauli_len", C.c_uint32),
        ("pauli", C.c_char_p), ("clbit", C.c_uint32),
    ]


class OrcCircuit(C.Structure):
    _fields_ = [("n_qubits", C.c_uint32), ("n_clbits", C.c_uint32), ("n_inst", C.c_uint64),
                ("inst", C.POINTER(OrcInst))]


class OrcChannel(C.Structure):
    _fields_ = [("n_ops", C.c_uint32), ("ops", C.POINTER(OrcC))]


class OrcNoise(C.Structure):
    _fields_ = [("n_channels", C.c_uint32), ("channels", C.POINTER(OrcChannel))]


class OrcTerm(C.Structure):
    _fields_ = [("coeff", C.c_double), ("pauli", C.c_char_p)]


_libs: dict[str, C.CDLL] = {}


def _load(path: str) -> C.CDLL:
    if path not in _libs:
        if not os.path.exists(path):
            raise FileNotFoundError(f"checker library not built: {path}")
        _libs[path] = C.CDLL(path)
    return _libs[path]


def oracle_lib() -> C.CDLL:
    lib = _load(ORACLE_SO)
    lib.orc_norm.restype = C.c_double
    lib.orc_expectation_value.restype = C.c_double
    lib.orc_expectation.restype = C.c_double
    lib.orc_rng_uniform.restype = C.c_double
    lib.orc_rng_next.restype = C.c_uint64
    lib.orc_draws_per_shot.restype = C.c_uint64
    lib.orc_fuse.restype = C.c_void_p
    lib.orc_fused_circuit.restype = C.POINTER(OrcCircuit)
    lib.orc_fused_circuit.argtypes = [C.c_void_p]
    lib.orc_fused_free.argtypes = [C.c_void_p]
    return lib


def ref_available() -> bool:
    return os.path.exists(REF_SO)


def ref_lib() -> C.CDLL:
    lib = _load(REF_SO)
    lib.ref_last_error.restype = C.c_char_p
    return lib


class Circuit:
    """Records builder calls (names and argument order of QuantumCircuit)."""

    def __init__(self, n_qubits: int, n_clbits: int = 0):
        self.n_qubits = n_qubits
        self.n_clbits = n_clbits
        self.declared_clbits = n_clbits
        self.calls: list[tuple[str, tuple]] = []
        self.insts: list[dict] = []

    # -- recording helpers
    def _g(self, call, typ, qubits, params=(), nc=0, mat=None, perm=None, pauli=None, clbit=0):
        self.calls.append(call)
        self.insts.append(dict(type=typ, qubits=list(qubits), params=list(params), nc=nc,
                               mat=None if mat is None else np.asarray(mat, complex).reshape(-1),
                               perm=None if perm is None else list(perm), pauli=pauli, clbit=clbit))
        return self

    def __getattr__(self, name):
        fixed = {"i": I, "x": X, "y": Y, "z": Z, "h": H, "s": S, "sdg": SDG, "t": T, "tdg": TDG, "sx": SX}
        rot = {"rx": RX, "ry": RY, "rz": RZ, "p": P}
        ctl = {"cx": (X, 0), "cy": (Y, 0), "cz": (Z, 0), "ch": (H, 0), "cp": (P, 1), "crx": (RX, 1),
               "cry": (RY, 1), "crz": (RZ, 1)}
        two = {"swap": (SWAP, 0), "iswap": (ISWAP, 0), "rxx": (RXX, 1), "ryy": (RYY, 1), "rzz": (RZZ, 1)}
        if name in fixed:
            return lambda q: self._g((name, (q,)), fixed[name], [q])
        if name in rot:
            return lambda a, q: self._g((name, (a, q)), rot[name], [q], [a])
        if name in ctl:
            typ, npar = ctl[name]
            if npar:
                return lambda a, c, t: self._g((name, (a, c, t)), typ, [c, t], [a], nc=1)
            return lambda c, t: self._g((name, (c, t)), typ, [c, t], nc=1)
        if name in two:
            typ, npar = two[name]
            if npar:
                return lambda a, q0, q1: self._g((name, (a, q0, q1)), typ, [q0, q1], [a])
            return lambda q0, q1: self._g((name, (q0, q1)), typ, [q0, q1])
        raise AttributeError(name)

    def u(self, th, ph, la, q):
        return self._g(("u", (th, ph, la, q)), U, [q], [th, ph, la])

    def ccx(self, c1, c2, t):
        return self._g(("ccx", (c1, c2, t)), X, [c1, c2, t], nc=2)

    def cswap(self, c, a, b):
        return self._g(("cswap", (c, a, b)), SWAP, [c, a, b], nc=1)

    def mcx(self, controls, t):
        return self._g(("mcx", (list(controls), t)), X, [*controls, t], nc=len(controls))

    def mcp(self, lam, controls, t):
        return self._g(("mcp", (lam, list(controls), t)), P, [*controls, t], [lam], nc=len(controls))

    def rp(self, theta, pauli, qubits):
        return self._g(("rp", (theta, pauli, list(qubits))), PAULI_ROTATION, qubits, [theta], pauli=pauli)

    def unitary(self, mat, qubits):
        return self._g(("unitary", (np.asarray(mat, complex), list(qubits))), UNITARY, qubits, mat=mat)

    def controlled_unitary(self, mat, controls, targets):
        return self._g(("controlled_unitary", (np.asarray(mat, complex), list(controls), list(targets))),
                       UNITARY, [*controls, *targets], nc=len(controls), mat=mat)

    def permutation(self, table, qubits):
        return self._g(("permutation", (list(table), list(qubits))), PERMUTATION, qubits, perm=table)

    def controlled_permutation(self, table, controls, targets):
        return self._g(("controlled_permutation", (list(table), list(controls), list(targets))), PERMUTATION,
                       [*controls, *targets], nc=len(controls), perm=table)

    def measure(self, q, c):
        return self._g(("measure", (q, c)), MEASURE, [q], clbit=c)

    def measure_all(self):
        self.n_clbits = max(self.n_clbits, self.n_qubits)
        self.calls.append(("measure_all", ()))
        for q in range(self.n_qubits):
            self.insts.append(dict(type=MEASURE, qubits=[q], params=[], nc=0, mat=None, perm=None,
                                   pauli=None, clbit=q))
        return self

    def reset(self, q):
        return self._g(("reset", (q,)), RESET, [q])

    def barrier(self):
        return self._g(("barrier", ()), BARRIER, [])

    # -- product replay
    def to_product(self, lq):
        qc = lq.QuantumCircuit(self.n_qubits, self.declared_clbits)
        for name, args in self.calls:
            getattr(qc, name)(*args)
        return qc

    # -- oracle format
    def to_orc(self):
        keep = []
        insts = (OrcInst * max(1, len(self.insts)))()
        for k, d in enumerate(self.insts):
            q = (C.c_uint32 * max(1, len(d["qubits"])))(*d["qubits"])
            pr = (C.c_double * max(1, len(d["params"])))(*d["params"])
            keep += [q, pr]
            it = insts[k]
            it.type = d["type"]
            it.n_controls = d["nc"]
            it.n_qubits = len(d["qubits"])
            it.n_params = len(d["params"])
            it.qubits = q
            it.params = pr
            if d["mat"] is not None:
                m = (OrcC * len(d["mat"]))(*[OrcC(z.real, z.imag) for z in d["mat"]])
                keep.append(m)
                it.mat_len = len(d["mat"])
                it.mat = m
            if d["perm"] is not None:
                pm = (C.c_uint64 * len(d["perm"]))(*d["perm"])
                keep.append(pm)
                it.perm_len = len(d["perm"])
                it.perm = pm
            if d["pauli"] is not None:
                b = d["pauli"].encode()
                keep.append(b)
                it.pauli_len = len(b)
                it.pauli = b
            it.clbit = d["clbit"]
        circ = OrcCircuit(self.n_qubits, self.n_clbits, len(self.insts), insts)
        keep.append(insts)
        return circ, keep


def noise_struct(channels: Sequence[Sequence[Sequence[complex]]]):
    """channels: list of channels, each a list of 2x2 (4 entries) Kraus ops."""
    keep = []
    arr = (OrcChannel * max(1, len(channels)))()
    for i, ch in enumerate(channels):
        flat = [complex(z) for op in ch for z in np.asarray(op, complex).reshape(-1)]
        ops = (OrcC * len(flat))(*[OrcC(z.real, z.imag) for z in flat])
        keep.append(ops)
        arr[i].n_ops = len(ch)
        arr[i].ops = ops
    nz = OrcNoise(len(channels), arr)
    keep.append(arr)
    return nz, keep


def depolarizing(p):
    a, b = math.sqrt(1 - 3 * p / 4), math.sqrt(p / 4)
    return [[a, 0, 0, a], [0, b, b, 0], [0, -1j * b, 1j * b, 0], [b, 0, 0, -b]]


def amplitude_damping(g):
    return [[1, 0, 0, math.sqrt(1 - g)], [0, math.sqrt(g), 0, 0]]


def bit_flip(p):
    a, b = math.sqrt(1 - p), math.sqrt(p)
    return [[a, 0, 0, a], [0, b, b, 0]]


def phase_damping(g):
    return [[1, 0, 0, math.sqrt(1 - g)], [0, 0, 0, math.sqrt(g)]]


def _cbuf(n):
    return np.zeros(n, dtype=np.complex128)


# ---- oracle calls ---------------------------------------------------------------------
def run_statevector(circ: Circuit, fuse: bool = True) -> np.ndarray:
    lib = oracle_lib()
    c, keep = circ.to_orc()
    out = _cbuf(1 << circ.n_qubits)
    rc = lib.orc_run_statevector(C.byref(c), int(fuse), out.ctypes.data_as(C.POINTER(OrcC)))
    if rc != 0:
        raise ValueError("run_statevector: Reset is non-deterministic")
    return out


def run_sampling(circ: Circuit, seed: int, shots: int, fuse: bool = True):
    """(indices, cumulative) of the terminal-sampling path."""
    lib = oracle_lib()
    c, keep = circ.to_orc()
    idx = np.zeros(shots, dtype=np.uint64)
    cum = np.zeros(1 << circ.n_qubits, dtype=np.float64)
    rc = lib.orc_run_sampling(C.byref(c), int(fuse), C.c_uint64(seed), C.c_uint64(shots),
                              idx.ctypes.data_as(C.POINTER(C.c_uint64)),
                              cum.ctypes.data_as(C.POINTER(C.c_double)))
    if rc != 0:
        raise ValueError(f"orc_run_sampling rc={rc}")
    return idx, cum


def run_counts(circ: Circuit, seed: int, shots: int, fuse: bool = True, channels=()):
    """StatevectorSimulator::run counts as {bitstring: count}."""
    lib = oracle_lib()
    c, keep = circ.to_orc()
    nz, keep2 = noise_struct(channels)
    out = np.zeros(shots * max(1, circ.n_clbits), dtype=np.uint8)
    rc = lib.orc_run(C.byref(c), C.byref(nz) if channels else None, int(fuse), C.c_uint64(seed),
                     C.c_uint64(shots), out.ctypes.data_as(C.POINTER(C.c_uint8)))
    if rc < 0:
        raise ValueError(f"orc_run rc={rc}")
    counts: dict[str, int] = {}
    nc = circ.n_clbits
    for s in range(shots):
        bits = out[s * nc:(s + 1) * nc]
        key = "".join("1" if bits[nc - 1 - k] else "0" for k in range(nc))
        counts[key] = counts.get(key, 0) + 1
    return counts


def expectation_value(state: np.ndarray, n_qubits: int, pauli: str) -> float:
    lib = oracle_lib()
    s = np.ascontiguousarray(state, dtype=np.complex128)
    return lib.orc_expectation_value(s.ctypes.data_as(C.POINTER(OrcC)), C.c_uint32(n_qubits), pauli.encode())


def _terms(h):
    arr = (OrcTerm * max(1, len(h)))()
    keep = []
    for i, (c, p) in enumerate(h):
        b = p.encode()
        keep.append(b)
        arr[i].coeff = c
        arr[i].pauli = b
    return arr, keep


def expectation(state: np.ndarray, n_qubits: int, h) -> float:
    lib = oracle_lib()
    s = np.ascontiguousarray(state, dtype=np.complex128)
    arr, keep = _terms(h)
    return lib.orc_expectation(s.ctypes.data_as(C.POINTER(OrcC)), C.c_uint32(n_qubits), arr, C.c_uint64(len(h)))


def parameter_shift_gradient(template: Circuit, param_inst: Sequence[int], theta, h) -> np.ndarray:
    lib = oracle_lib()
    c, keep = template.to_orc()
    pi = np.asarray(param_inst, dtype=np.uint64)
    th = np.asarray(theta, dtype=np.float64)
    arr, keep2 = _terms(h)
    g = np.zeros(len(th), dtype=np.float64)
    lib.orc_parameter_shift_gradient(C.byref(c), pi.ctypes.data_as(C.POINTER(C.c_uint64)), C.c_uint64(len(th)),
                                     th.ctypes.data_as(C.POINTER(C.c_double)), arr, C.c_uint64(len(h)),
                                     g.ctypes.data_as(C.POINTER(C.c_double)))
    return g


def bound_energy(template: Circuit, param_inst: Sequence[int], theta, h) -> float:
    lib = oracle_lib()
    c, keep = template.to_orc()
    pi = np.asarray(param_inst, dtype=np.uint64)
    th = np.asarray(theta, dtype=np.float64)
    arr, keep2 = _terms(h)
    e = C.c_double(0)
    lib.orc_bound_energy(C.byref(c), pi.ctypes.data_as(C.POINTER(C.c_uint64)), C.c_uint64(len(th)),
                         th.ctypes.data_as(C.POINTER(C.c_double)), arr, C.c_uint64(len(h)), C.byref(e))
    return e.value


def vqe_rotosolve(template: Circuit, param_inst: Sequence[int], h, max_iterations: int = 100,
                  tol: float = 1e-9, init=None):
    """algorithms::vqe (src/algorithms/vqe.cpp:19-63) over the restated
    bound_energy: per coordinate, energies at s and s +- pi/2 fix the sinusoid
    E(t) = a + b cos(t - phi); the coordinate moves to its minimum folded into
    (-pi, pi] (std::remainder). Returns (energy, params, iterations, history)."""
    n = len(param_inst)
    theta = np.zeros(n) if init is None or len(init) == 0 else np.array(init, dtype=np.float64)
    if len(theta) != n:
        raise ValueError("initial_parameters size must equal n_params")
    e = bound_energy(template, param_inst, theta, h)
    history, it = [], 0
    for sweep in range(max_iterations):
        start = e
        for k in range(n):
            s0 = theta[k]
            theta[k] = s0 + math.pi / 2
            ep = bound_energy(template, param_inst, theta, h)
            theta[k] = s0 - math.pi / 2
            em = bound_energy(template, param_inst, theta, h)
            theta[k] = math.remainder(s0 - math.pi / 2 - math.atan2(2.0 * e - ep - em, ep - em), 2.0 * math.pi)
            e = bound_energy(template, param_inst, theta, h)
        history.append(e)
        it = sweep + 1
        if start - e < tol:
            break
    return e, theta, it, history


def vqe_gradient_descent(template: Circuit, param_inst: Sequence[int], h, max_iterations: int = 200,
                         learning_rate: float = 0.1, tol: float = 1e-9, init=None):
    """algorithms::vqe_gradient_descent (src/algorithms/vqe.cpp:96-133) over
    the restated parameter-shift gradient: theta -= lr * g, record E(theta),
    stop once |g| < tol. Returns (energy, params, iterations, history)."""
    n = len(param_inst)
    theta = np.zeros(n) if init is None or len(init) == 0 else np.array(init, dtype=np.float64)
    history, it = [], 0
    for i in range(max_iterations):
        g = parameter_shift_gradient(template, param_inst, theta, h)
        theta = theta - learning_rate * g
        history.append(bound_energy(template, param_inst, theta, h))
        it = i + 1
        if math.sqrt(float(np.dot(g, g))) < tol:
            break
    e = history[-1] if history else bound_energy(template, param_inst, theta, h)
    return e, theta, it, history


def density_probabilities(circ: Circuit, channels=()) -> np.ndarray:
    lib = oracle_lib()
    c, keep = circ.to_orc()
    nz, keep2 = noise_struct(channels)
    rho = _cbuf(1 << (2 * circ.n_qubits))
    lib.orc_run_density_matrix(C.byref(c), C.byref(nz) if channels else None, rho.ctypes.data_as(C.POINTER(OrcC)))
    d = 1 << circ.n_qubits
    return np.array([rho[(i << circ.n_qubits) | i].real for i in range(d)])


def density_matrix(circ: Circuit, channels=()) -> np.ndarray:
    """rho as a 2^n x 2^n matrix, rho[r, c] = <r|rho|c>."""
    lib = oracle_lib()
    c, keep = circ.to_orc()
    nz, keep2 = noise_struct(channels)
    vec = _cbuf(1 << (2 * circ.n_qubits))
    lib.orc_run_density_matrix(C.byref(c), C.byref(nz) if channels else None, vec.ctypes.data_as(C.POINTER(OrcC)))
    d = 1 << circ.n_qubits
    return vec.reshape(d, d).T.copy()  # vector index (col << n) | row


def mt19937_64_outputs(seed: int, n: int) -> np.ndarray:
    lib = oracle_lib()
    st = (C.c_uint64 * 313)()
    lib.orc_rng_seed(st, C.c_uint64(seed))
    return np.array([lib.orc_rng_next(st) for _ in range(n)], dtype=np.uint64)


def uniform_stream(seed: int, n: int) -> np.ndarray:
    lib = oracle_lib()
    out = np.zeros(n, dtype=np.float64)
    lib.orc_uniform_stream(C.c_uint64(seed), C.c_uint64(n), out.ctypes.data_as(C.POINTER(C.c_double)))
    return out


def base_matrix(circ: Circuit, k: int = 0) -> np.ndarray:
    lib = oracle_lib()
    c, keep = circ.to_orc()
    out = _cbuf(4096)
    dim = lib.orc_base_matrix(C.byref(c.inst[k]), out.ctypes.data_as(C.POINTER(OrcC)))
    return out[: dim * dim].reshape(dim, dim)


# ---- reference calls (oracle/_ref) ---------------------------------------------------------
def _ref_check(rc):
    if rc != 0:
        raise RuntimeError(f"reference error {rc}: {ref_lib().ref_last_error().decode()}")


def ref_run_statevector(circ: Circuit, fuse: bool = True) -> np.ndarray:
    c, keep = circ.to_orc()
    out = _cbuf(1 << circ.n_qubits)
    _ref_check(ref_lib().ref_run_statevector(C.byref(c), int(fuse), out.ctypes.data_as(C.POINTER(OrcC))))
    return out


def ref_run_counts(circ: Circuit, seed: int, shots: int, fuse: bool = True, channels=()):
    c, keep = circ.to_orc()
    nz, keep2 = noise_struct(channels)
    buf = C.create_string_buffer(1 << 22)
    _ref_check(ref_lib().ref_run_counts(C.byref(c), C.byref(nz) if channels else None, int(fuse),
                                        C.c_uint64(seed), C.c_uint64(shots), buf, C.c_uint64(len(buf))))
    counts = {}
    for line in buf.value.decode().splitlines():
        k, v = line.split("=")
        counts[k] = int(v)
    return counts


def ref_expectation(state, n_qubits, h) -> float:
    s = np.ascontiguousarray(state, dtype=np.complex128)
    arr, keep = _terms(h)
    out = C.c_double(0)
    _ref_check(ref_lib().ref_expectation(s.ctypes.data_as(C.POINTER(OrcC)), C.c_uint32(n_qubits), arr,
                                         C.c_uint64(len(h)), C.byref(out)))
    return out.value


def ref_parameter_shift_gradient(template: Circuit, param_inst, theta, h) -> np.ndarray:
    c, keep = template.to_orc()
    pi = np.asarray(param_inst, dtype=np.uint64)
    th = np.asarray(theta, dtype=np.float64)
    arr, keep2 = _terms(h)
    g = np.zeros(len(th), dtype=np.float64)
    _ref_check(ref_lib().ref_parameter_shift_gradient(
        C.byref(c), pi.ctypes.data_as(C.POINTER(C.c_uint64)), C.c_uint64(len(th)),
        th.ctypes.data_as(C.POINTER(C.c_double)), arr, C.c_uint64(len(h)), g.ctypes.data_as(C.POINTER(C.c_double))))
    return g


def ref_density_probabilities(circ: Circuit, channels=()) -> np.ndarray:
    c, keep = circ.to_orc()
    nz, keep2 = noise_struct(channels)
    out = np.zeros(1 << circ.n_qubits, dtype=np.float64)
    _ref_check(ref_lib().ref_density_probabilities(C.byref(c), C.byref(nz) if channels else None,
                                                   out.ctypes.data_as(C.POINTER(C.c_double))))
    return out


def ref_to_qasm(circ: Circuit) -> str:
    c, keep = circ.to_orc()
    buf = C.create_string_buffer(1 << 22)
    _ref_check(ref_lib().ref_to_qasm(C.byref(c), buf, C.c_uint64(len(buf))))
    return buf.value.decode()


def ref_qasm_reparse(text: str) -> tuple[str, int]:
    """(to_qasm(from_qasm(text)), instruction count) in the reference."""
    buf = C.create_string_buffer(1 << 22)
    n = C.c_uint64(0)
    _ref_check(ref_lib().ref_qasm_reparse(text.encode(), buf, C.c_uint64(len(buf)), C.byref(n)))
    return buf.value.decode(), n.value


def ref_eliminate_qasm(circ: Circuit) -> tuple[str, int]:
    """RedundantGateElimination::run in the reference, as (QASM, instruction count)."""
    c, keep = circ.to_orc()
    buf = C.create_string_buffer(1 << 22)
    n = C.c_uint64(0)
    _ref_check(ref_lib().ref_eliminate_qasm(C.byref(c), buf, C.c_uint64(len(buf)), C.byref(n)))
    return buf.value.decode(), n.value


def ref_qft_qasm(n: int, do_swaps: bool = True, inverse: bool = False) -> str:
    buf = C.create_string_buffer(1 << 22)
    _ref_check(ref_lib().ref_qft_qasm(C.c_uint32(n), int(do_swaps), int(inverse), buf, C.c_uint64(len(buf))))
    return buf.value.decode()


def _ref_vqe(fn, template: Circuit, param_inst, h, max_iterations, extra, init):
    c, keep = template.to_orc()
    pi = np.asarray(param_inst, dtype=np.uint64)
    arr, keep2 = _terms(h)
    n = len(pi)
    ini = None if init is None or len(init) == 0 else np.asarray(init, dtype=np.float64)
    e = C.c_double(0)
    params = np.zeros(n)
    it = C.c_uint64(0)
    hist = np.zeros(max(1, max_iterations))
    dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))  # noqa: E731
    _ref_check(fn(C.byref(c), pi.ctypes.data_as(C.POINTER(C.c_uint64)), C.c_uint64(n), arr, C.c_uint64(len(h)),
                  C.c_uint64(max_iterations), *[C.c_double(x) for x in extra], dp(ini) if ini is not None else None,
                  C.byref(e), dp(params), C.byref(it), dp(hist)))
    return e.value, params, it.value, list(hist[: it.value])


def ref_vqe_rotosolve(template: Circuit, param_inst, h, max_iterations=100, tol=1e-9, init=None):
    return _ref_vqe(ref_lib().ref_vqe_rotosolve, template, param_inst, h, max_iterations, (tol,), init)


def ref_vqe_gradient_descent(template: Circuit, param_inst, h, max_iterations=200, learning_rate=0.1, tol=1e-9,
                             init=None):
    return _ref_vqe(ref_lib().ref_vqe_gradient_descent, template, param_inst, h, max_iterations,
                    (learning_rate, tol), init)


def ref_uniform_stream(seed: int, n: int) -> np.ndarray:
    out = np.zeros(n, dtype=np.float64)
    ref_lib().ref_uniform_stream(C.c_uint64(seed), C.c_uint64(n), out.ctypes.data_as(C.POINTER(C.c_double)))
    return out


def ref_set_threads(n: int) -> None:
    ref_lib().ref_set_threads(int(n))
