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
