"""OpenQASM 2.0 import/export (reference src/io/qasm.cpp; its tests
tests/test_qasm.cpp and tests/python/test_bindings.py:161-170 restated).

The text-level checks run on the CPU (circuits are host objects); the ones
that compare simulated states or counts are marked gpu."""
import math

import numpy as np
import pytest


@pytest.fixture(scope="module")
def lq():
    from litequantumcomputingsimulator_amd import pylqcs

    return pylqcs


def g17(v):
    return "%.17g" % v


def sample_circuit(lq):
    """test_qasm.cpp:26-31."""
    qc = lq.QuantumCircuit(3, 3)
    qc.h(0).x(1).sdg(2).rx(0.5, 0).u(0.3, 1.2, -0.7, 1).p(0.9, 2)
    qc.cx(0, 1).cz(1, 2).cp(1.1, 0, 2).crz(-0.4, 2, 0).swap(0, 2)
    qc.ccx(0, 1, 2).cswap(2, 0, 1).barrier().measure_all()
    return qc


QISKIT_SRC = """
// generated by qiskit
OPENQASM 2.0;
include "qelib1.inc";
qreg q[2];
creg c[2];
u2(0,pi) q[0];        // = H
cx q[0],q[1];
u1(pi/4) q[1];        // = T
rz(3*pi/2) q[0];
barrier q[0],q[1];
measure q -> c;
"""


# ---- CPU: text ---------------------------------------------------------------------------
def test_export_contains_expected_header(lq):
    """test_qasm.cpp:40-50."""
    qc = lq.QuantumCircuit(1, 1)
    qc.h(0).measure(0, 0)
    text = lq.io.to_qasm(qc)
    for line in ('OPENQASM 2.0;', 'include "qelib1.inc";', "qreg q[1];", "creg c[1];", "h q[0];",
                 "measure q[0] -> c[0];"):
        assert line in text


def test_export_format_line_by_line(lq):
    qc = lq.QuantumCircuit(3)
    qc.u(0.1, -2.5, 1e-300, 2).cp(math.pi, 2, 0).reset(1).barrier()
    assert lq.io.to_qasm(qc) == (
        'OPENQASM 2.0;\ninclude "qelib1.inc";\nqreg q[3];\n'
        f"u3({g17(0.1)},{g17(-2.5)},{g17(1e-300)}) q[2];\n"
        f"cu1({g17(math.pi)}) q[2],q[0];\n"
        "reset q[1];\nbarrier q;\n")


def test_round_trip_is_exact(lq):
    """Parameters print as %.17g, so export(import(export(c))) == export(c)
    and the instruction counts agree (test_qasm.cpp:26-38)."""
    qc = sample_circuit(lq)
    rng = np.random.default_rng(0)
    for q in range(3):
        qc.u(*(float(v) for v in rng.standard_normal(3) * 10), q)
    text = lq.io.to_qasm(qc)
    parsed = lq.io.from_qasm(text)
    assert parsed.num_qubits() == qc.num_qubits()
    assert parsed.num_clbits() == qc.num_clbits()
    assert parsed.num_instructions() == qc.num_instructions()
    assert lq.io.to_qasm(parsed) == text


def test_parses_qiskit_style_source(lq):
    """test_qasm.cpp:52-74 at the text level: comments, pi expressions, u2,
    u1, a barrier with operands, whole-register measure."""
    parsed = lq.io.from_qasm(QISKIT_SRC)
    assert parsed.num_qubits() == 2 and parsed.num_clbits() == 2
    assert lq.io.to_qasm(parsed) == (
        'OPENQASM 2.0;\ninclude "qelib1.inc";\nqreg q[2];\ncreg c[2];\n'
        f"u3({g17(math.pi / 2)},0,{g17(math.pi)}) q[0];\n"
        "cx q[0],q[1];\n"
        f"u1({g17(math.pi / 4)}) q[1];\n"
        f"rz({g17(3 * math.pi / 2)}) q[0];\n"
        "barrier q;\nmeasure q[0] -> c[0];\nmeasure q[1] -> c[1];\n")


def test_expressions(lq):
    src = ('OPENQASM 2.0; qreg r[2]; rx(-(pi-1)*2/4) r[1]; ry(1.5e-3) r[0]; rz(-pi) r[1+0];'
           ' cu1(pi/2/2) r[0],r[1];')
    text = lq.io.to_qasm(lq.io.from_qasm(src))
    assert f"rx({g17(-(math.pi - 1) * 2 / 4)}) q[1];" in text
    assert f"ry({g17(1.5e-3)}) q[0];" in text
    assert f"rz({g17(-math.pi)}) q[1];" in text
    assert f"cu1({g17(math.pi / 2 / 2)}) q[0],q[1];" in text


def test_rejects_non_representable_gates(lq):
    """test_qasm.cpp:86-103."""
    qc = lq.QuantumCircuit(2)
    qc.iswap(0, 1)
    with pytest.raises(ValueError):
        lq.io.to_qasm(qc)
    qc = lq.QuantumCircuit(4)
    qc.mcx([0, 1, 2], 3)
    with pytest.raises(ValueError):
        lq.io.to_qasm(qc)
    qc = lq.QuantumCircuit(2)
    qc.permutation([1, 2, 3, 0], [0, 1])
    with pytest.raises(ValueError):
        lq.io.to_qasm(qc)


@pytest.mark.parametrize("src", [
    "OPENQASM 3.0;",                                          # test_qasm.cpp:106-109
    "OPENQASM 2.0; h q[0];",
    "OPENQASM 2.0; qreg q[1]; frobnicate q[0];",
    "OPENQASM 2.0;",                                          # no qreg
    "OPENQASM 2.0; qreg q[1]; qreg r[1];",                    # two qregs
    "OPENQASM 2.0; qreg q[1]; h q[0]; creg c[1];",            # creg after a gate
    "OPENQASM 2.0; qreg q[2]; creg c[1]; measure q -> c;",    # whole-register sizes differ
    "OPENQASM 2.0; qreg q[2]; creg c[2]; measure q[0] -> c;",  # mixed measure
    "OPENQASM 2.0; qreg q[1]; h r[0];",                       # unknown register
    "OPENQASM 2.0; qreg q[1]; rx(phi) q[0];",                 # unknown symbol
    "OPENQASM 2.0; qreg q[1]; reset q;",                      # whole-register reset
    "OPENQASM 2.0; qreg q[2]; cx q, q[1];",                   # unindexed operand
    'OPENQASM 2.0; include "qelib1.inc;',                     # unterminated string
    "OPENQASM 2.0; qreg q[1]; h q[0]",                        # missing ';'
])
def test_rejects_malformed_source(lq, src):
    with pytest.raises(ValueError):
        lq.io.from_qasm(src)


# ---- GPU: states and counts ----------------------------------------------------------------
@pytest.mark.gpu
def test_round_trip_preserves_state(lq):
    """test_qasm.cpp:26-38."""
    qc = sample_circuit(lq)
    parsed = lq.io.from_qasm(lq.io.to_qasm(qc))
    sim = lq.StatevectorSimulator()
    a = np.array(sim.run_statevector(qc).to_numpy())
    b = np.array(sim.run_statevector(parsed).to_numpy())
    assert np.max(np.abs(a - b)) <= 1e-12


@pytest.mark.gpu
def test_qiskit_style_source_state(lq):
    """test_qasm.cpp:52-74."""
    parsed = lq.io.from_qasm(QISKIT_SRC)
    expected = lq.QuantumCircuit(2, 2)
    expected.h(0).cx(0, 1).t(1).rz(3 * math.pi / 2, 0).barrier().measure_all()
    sim = lq.StatevectorSimulator()
    a = np.array(sim.run_statevector(parsed).to_numpy())
    b = np.array(sim.run_statevector(expected).to_numpy())
    assert np.max(np.abs(a - b)) <= 1e-12


@pytest.mark.gpu
def test_measure_mapping_survives_round_trip(lq):
    """test_qasm.cpp:76-84."""
    qc = lq.QuantumCircuit(2, 2)
    qc.x(0).measure(0, 1).measure(1, 0)
    parsed = lq.io.from_qasm(lq.io.to_qasm(qc))
    c1 = dict(lq.StatevectorSimulator(seed=3).run(qc, 50).counts())
    c2 = dict(lq.StatevectorSimulator(seed=3).run(parsed, 50).counts())
    assert c1 == c2 == {"10": 50}


@pytest.mark.gpu
def test_python_round_trip_counts(lq):
    """tests/python/test_bindings.py:161-170."""
    qc = lq.QuantumCircuit(2, 2)
    qc.h(0).cx(0, 1).rz(0.5, 1).measure_all()
    text = lq.io.to_qasm(qc)
    assert "OPENQASM 2.0;" in text
    parsed = lq.io.from_qasm(text)
    assert parsed.num_qubits() == 2
    a = dict(lq.StatevectorSimulator(seed=3).run(qc, shots=100).counts())
    b = dict(lq.StatevectorSimulator(seed=3).run(parsed, shots=100).counts())
    assert a == b
