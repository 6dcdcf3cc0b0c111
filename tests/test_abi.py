"""CPU checks of the boundary: the C-ABI libraries load and export every
entry point include/lqcs_device.h declares; the host layer's validation and
exception types match the reference; host-computed gate matrices equal the
oracle's (which the reference pins). No device compute here."""
import ctypes
import os
import re

import numpy as np
import pytest

from litequantumcomputingsimulator_amd import LIB_DIR
from oracle import oracle as O
from tests import cases

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "lqcs_device.h")).read()
    return sorted(set(re.findall(r"^(?:lqcs_status|const char\*)\s+(lqcs_\w+)\s*\(", text, re.M)))


def test_device_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(os.path.join(LIB_DIR, "liblqcs_device.so"))
    names = declared_symbols()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_host_library_loads():
    ctypes.CDLL(os.path.join(LIB_DIR, "liblqcs.so"))


def test_last_error_and_no_device_path():
    lib = ctypes.CDLL(os.path.join(LIB_DIR, "liblqcs_device.so"))
    lib.lqcs_dev_last_error.restype = ctypes.c_char_p
    st = ctypes.c_void_p()
    rc = lib.lqcs_state_create(ctypes.c_uint32(0), ctypes.byref(st))
    assert rc == 1  # LQCS_ERR_INVALID_ARGUMENT: width checked before any device call
    assert b"num_qubits" in lib.lqcs_dev_last_error()
    rc = lib.lqcs_state_create(ctypes.c_uint32(40), ctypes.byref(st))
    assert rc == 1


@pytest.fixture(scope="module")
def lq():
    from litequantumcomputingsimulator_amd import pylqcs

    return pylqcs


def test_builder_validation_matches_reference_types(lq):
    qc = lq.QuantumCircuit(2)
    with pytest.raises(IndexError):
        qc.h(5)  # out_of_range
    with pytest.raises(ValueError):
        qc.cx(1, 1)  # duplicate qubit
    with pytest.raises(ValueError):
        qc.unitary(np.array([[1, 1], [0, 1]], complex), [0])  # not unitary
    with pytest.raises(ValueError):
        qc.permutation([0, 0, 1, 2], [0, 1])  # not a bijection
    with pytest.raises(ValueError):
        qc.rp(0.5, "XAB", [0, 1])
    with pytest.raises(ValueError):
        qc.rp(0.5, "X", [0, 1])
    with pytest.raises(ValueError):
        lq.QuantumCircuit(0)
    with pytest.raises(ValueError):
        lq.KrausChannel.depolarizing(1.5)
    with pytest.raises(ValueError):
        lq.KrausChannel([np.eye(2), np.array([[0, 1], [1, 0]])])  # not complete
    pc = lq.ParametricCircuit(1)
    pc.rx(0)
    pc.rz(0)
    with pytest.raises(ValueError):
        pc.bind([0.1])


def test_fluent_chain_and_depth(lq):
    qc = lq.QuantumCircuit(2, 2)
    qc.h(0).cx(0, 1).measure_all()
    assert qc.num_qubits() == 2 and qc.depth() == 3 and "instructions=4" in repr(qc)
    pc = lq.ParametricCircuit(2)
    assert pc.ry(0) == 0 and pc.rx(1) == 1
    pc.cx(0, 1)
    assert pc.rz(0) == 2 and pc.num_parameters() == 3


def test_device_ops_fail_loudly_without_gpu(lq):
    try:
        lq.device_info()
    except RuntimeError:
        with pytest.raises(RuntimeError):
            lq.Statevector(2)
        return
    pytest.skip("a GPU is present")


def test_host_gate_matrices_equal_oracle(lq):
    """Gate.matrix() of the host layer (C++20) vs the oracle's base_matrix,
    bit for bit, for every parametrised and fixed gate kind."""
    c = cases.every_gate_circuit(6)
    G = lq.GateType
    kinds = {O.I: G.I, O.X: G.X, O.Y: G.Y, O.Z: G.Z, O.H: G.H, O.S: G.S, O.SDG: G.Sdg, O.T: G.T, O.TDG: G.Tdg,
             O.SX: G.SX, O.RX: G.RX, O.RY: G.RY, O.RZ: G.RZ, O.P: G.P, O.U: G.U, O.SWAP: G.SWAP,
             O.ISWAP: G.iSWAP, O.RXX: G.RXX, O.RYY: G.RYY, O.RZZ: G.RZZ}
    seen = 0
    for k, inst in enumerate(c.insts):
        if inst["type"] not in kinds or inst["nc"]:
            continue
        want = O.base_matrix(c, k)
        got = np.array(lq.Gate(kinds[inst["type"]], inst["params"], 0).matrix())
        assert np.array_equal(got, want), inst
        seen += 1
    assert seen >= 15


class _Program(ctypes.Structure):
    """lqcs_program (include/lqcs_device.h)."""
    _fields_ = [("ops", ctypes.c_void_p), ("n_ops", ctypes.c_uint64), ("qubits", ctypes.c_void_p),
                ("n_qubits", ctypes.c_uint64), ("cplx", ctypes.c_void_p), ("n_cplx", ctypes.c_uint64),
                ("perm", ctypes.c_void_p), ("n_perm", ctypes.c_uint64)]


def test_batched_adjoint_rejects_null_arrays_before_any_device_call():
    """lqcs_adjoint_gradient_batched checks its arrays (as the single-circuit
    entry does) and returns INVALID_ARGUMENT instead of dereferencing a null
    pointer; these checks come before the device is touched."""
    lib = ctypes.CDLL(os.path.join(LIB_DIR, "liblqcs_device.so"))
    lib.lqcs_dev_last_error.restype = ctypes.c_char_p
    u64 = ctypes.c_uint64
    nq = (ctypes.c_uint32 * 1)(4)
    progs = (_Program * 1)()
    pop = (u64 * 2)(0, 0)
    xs, zs = (u64 * 1)(1), (u64 * 1)(0)
    ys = (ctypes.c_uint32 * 1)(0)
    cs = (ctypes.c_double * 1)(1.0)
    grads, energies = (ctypes.c_double * 2)(), (ctypes.c_double * 1)()

    def call(param_op, param_off, x, z, y, c, term_off, g):
        return lib.lqcs_adjoint_gradient_batched(ctypes.c_uint32(1), nq, progs, param_op, param_off, x, z, y, c,
                                                 term_off, energies, g)

    p1, t1, t0 = (u64 * 2)(0, 2), (u64 * 2)(0, 1), (u64 * 2)(0, 0)
    assert call(None, p1, xs, zs, ys, cs, t1, grads) == 1  # parameters without param_op
    assert call(pop, p1, xs, zs, ys, cs, t1, None) == 1    # parameters without grads
    assert call(pop, p1, None, zs, ys, cs, t1, grads) == 1  # terms without x masks
    assert call(pop, p1, xs, zs, None, cs, t1, grads) == 1  # terms without n_y
    assert call(pop, p1, xs, zs, ys, None, t1, grads) == 1  # terms without coefficients
    assert b"null" in lib.lqcs_dev_last_error()
    assert call(pop, (u64 * 2)(2, 1), xs, zs, ys, cs, t0, grads) == 1  # decreasing offsets
    assert b"non-decreasing" in lib.lqcs_dev_last_error()


def test_adjoint_plan_checks_arguments_before_any_device_call():
    """lqcs_adjoint_plan_create refuses an empty batch, a null out pointer and
    the batched entry's bad arrays with INVALID_ARGUMENT before the device
    is touched; run refuses a null plan; destroy accepts one."""
    lib = ctypes.CDLL(os.path.join(LIB_DIR, "liblqcs_device.so"))
    lib.lqcs_dev_last_error.restype = ctypes.c_char_p
    u64 = ctypes.c_uint64
    nq = (ctypes.c_uint32 * 1)(4)
    progs = (_Program * 1)()
    pop = (u64 * 2)(0, 0)
    xs, zs = (u64 * 1)(1), (u64 * 1)(0)
    ys = (ctypes.c_uint32 * 1)(0)
    cs = (ctypes.c_double * 1)(1.0)
    p1, t1 = (u64 * 2)(0, 2), (u64 * 2)(0, 1)
    plan = ctypes.c_void_p()

    def create(n, param_op, param_off, term_off, out):
        return lib.lqcs_adjoint_plan_create(ctypes.c_uint32(n), nq, progs, param_op, param_off, xs, zs, ys, cs,
                                            term_off, out)

    assert create(1, pop, p1, t1, None) == 1
    assert create(0, pop, p1, t1, ctypes.byref(plan)) == 1
    assert b"at least one circuit" in lib.lqcs_dev_last_error()
    assert create(1, None, p1, t1, ctypes.byref(plan)) == 1  # parameters without param_op
    assert create(1, pop, (u64 * 2)(2, 1), t1, ctypes.byref(plan)) == 1
    assert b"non-decreasing" in lib.lqcs_dev_last_error()
    assert plan.value is None
    th = (ctypes.c_double * 2)()
    assert lib.lqcs_adjoint_plan_run(None, th, u64(2), None, th) == 1
    assert lib.lqcs_adjoint_plan_destroy(None) == 0


class _Op(ctypes.Structure):
    """lqcs_op (include/lqcs_device.h)."""
    _fields_ = [("kind", ctypes.c_uint32), ("n_controls", ctypes.c_uint32), ("n_targets", ctypes.c_uint32),
                ("qubit_offset", ctypes.c_uint32), ("data_offset", ctypes.c_uint64), ("x_mask", ctypes.c_uint64),
                ("zy_mask", ctypes.c_uint64), ("n_y", ctypes.c_uint32), ("reserved", ctypes.c_uint32),
                ("theta", ctypes.c_double)]


@pytest.mark.gpu
def test_graph_launch_refuses_a_register_it_was_not_captured_on():
    """A captured graph addresses the buffer of the register it was recorded
    on: launching it on another register (same or other width) returns LOGIC
    and changes nothing; on its own register it applies H again."""
    lib = ctypes.CDLL(os.path.join(LIB_DIR, "liblqcs_device.so"))
    lib.lqcs_dev_last_error.restype = ctypes.c_char_p
    assert lib.lqcs_dev_init(0) == 0
    a, b, c = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
    for st, n in ((a, 10), (b, 10), (c, 11)):
        assert lib.lqcs_state_create(ctypes.c_uint32(n), ctypes.byref(st)) == 0
    h = 1.0 / np.sqrt(2.0)
    ops = (_Op * 1)(_Op(0, 0, 1, 0, 0, 0, 0, 0, 0, 0.0))  # LQCS_OP_SINGLE on qubit 3
    qubits = (ctypes.c_uint32 * 1)(3)
    cplx = (ctypes.c_double * 8)(h, 0, h, 0, h, 0, -h, 0)
    prog = _Program(ctypes.cast(ops, ctypes.c_void_p), 1, ctypes.cast(qubits, ctypes.c_void_p), 1,
                    ctypes.cast(cplx, ctypes.c_void_p), 4, None, 0)
    g = ctypes.c_void_p()
    try:
        assert lib.lqcs_graph_create(a, ctypes.byref(prog), None, ctypes.byref(g)) == 0
        assert lib.lqcs_graph_launch(g, b) == 3
        assert b"captured" in lib.lqcs_dev_last_error()
        assert lib.lqcs_graph_launch(g, c) == 3
        assert lib.lqcs_graph_launch(g, a) == 0
        assert lib.lqcs_dev_synchronize() == 0
        out = np.zeros(2 << 10)
        assert lib.lqcs_state_download(a, out.ctypes.data_as(ctypes.c_void_p)) == 0
        amps = out[0::2] + 1j * out[1::2]
        want = np.zeros(1 << 10, complex)
        want[0] = want[8] = h
        assert np.max(np.abs(amps - want)) <= 1e-15
        assert lib.lqcs_state_download(b, out.ctypes.data_as(ctypes.c_void_p)) == 0
        assert out[0] == 1.0 and np.count_nonzero(out) == 1  # untouched |0>
    finally:
        if g:
            lib.lqcs_graph_destroy(g)
        for st in (a, b, c):
            lib.lqcs_state_destroy(st)
